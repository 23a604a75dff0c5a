// driver.cpp -- the fg2 command-line driver over the C ABI (fg2run).
//
//   fg2run MODULE [--help] [-b p1 [p2]] [-c C] [-d dev] [-l T [L1 [L2]]]
//          [-n frames [N1 [N2]]] [name=val ...] [seed=S] [INPUT]
//
// Behaviour of the reference's main() (src/main.cpp:29-120), solve()
// (src/solve.cpp:25-64), .raw snapshots (src/io.cpp:24-76) and initial
// conditions (src/init.cpp:26-42, src/*/init.h), with the module chosen at
// run time instead of at build time (Makefile:19-22):
//   - INPUT that names an existing file is loaded as a snapshot and the run
//     continues from its frame number (the first digit run of the name);
//     otherwise INPUT names an initial condition and frame 0000.raw is
//     written first;
//   - each frame advances the time by T / frames with
//     dt = min(target - t, getdt()) and writes NNNN.raw;
//   - a step that produces NaN writes ****.raw and exits with status 255.
// Initial conditions, snapshots and steps go through the C ABI (fg2_init,
// fg2_load, fg2_dump, fg2_step); the reference's libc rand() in KH and Kf is
// replaced by fg2_init's seeded stream (seed=S, default 0).  The start time
// read from a snapshot stays in FP64 (the reference narrows it to R,
// src/main.cpp:34,106), so a resumed run hits the same frame times as one
// that never stopped.
#include <hip/hip_runtime_api.h>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fg2_amd.h"

namespace {

[[noreturn]] void die(const char *fmt, ...)
{
  fflush(stdout);
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  exit(-1);
}

void usage(const char *bad)
{
  if(bad) {
    fprintf(stderr, "fg2run: illegal option \"%s\"\nTry `fg2run --help` for more information.\n", bad);
    exit(-1);
  }
  printf("Usage: fg2run MODULE [OPTION...] [INPUT_FILE]\n"
         "MI355X time-step path of the fg2 2D finite grid code\n\n"
         "  MODULE            hydro, hydro-isot or disk\n"
         "      --help        display this help and exit\n"
         "  -b                periodic boundary conditions (p/P = periodic)\n"
         "  -c                Courant number\n"
         "  -d                specify device id\n"
         "  -l                total time, box size\n"
         "  -n                number of frames and grids\n"
         "  parameter=val     set parameters in a particular scheme\n"
         "  seed=S            seed of the random initial conditions\n");
  exit(0);
}

void hip_check(hipError_t e, const char *what)
{
  if(e != hipSuccess) die("%s: %s\n", what, hipGetErrorString(e));
}

bool exists(const char *name)
{
  FILE *f = fopen(name, "r");
  if(!f) return false;
  fclose(f);
  return true;
}

int frame_of(const char *name)                             // src/misc.cpp:48-55
{
  for(const char *c = name; *c; ++c)
    if('0' <= *c && *c <= '9') return atoi(c);
  return 0;
}

std::string frame_name(int i)
{
  char b[64];
  snprintf(b, sizeof b, "%04d.raw", i);
  return b;
}

bool is_option(const char *a) { return a[0] == '-' && ((a[1] >= 'A' && a[1] <= 'Z') || (a[1] >= 'a' && a[1] <= 'z')); }

} // namespace

int main(int argc, char **argv)
{
  for(int i = 1; i < argc; ++i)
    if(!strcmp(argv[i], "--help")) usage(nullptr);
  if(argc < 2) usage(nullptr);
  const std::string module = argv[1];
  if(module != "hydro" && module != "hydro-isot" && module != "disk")
    die("fg2run: unknown module \"%s\" (hydro, hydro-isot, disk)\n", module.c_str());

  const char *input = "default";
  int n0 = 10, n1 = 1024, n2 = 1024, dev = 0, p1 = 1, p2 = 1;
  double l0 = 1.0, L1 = 1.0, L2 = 1.0, courant = 0.5;
  uint64_t seed = 0;
  std::vector<const char *> paras;
  auto more = [&](int i) { return i + 1 < argc && !is_option(argv[i + 1]); };
  for(int i = 2; i < argc; ++i) {                           // src/main.cpp:41-60
    const char *a = argv[i];
    if(strchr(a, '=')) {
      if(!strncmp(a, "seed=", 5)) seed = strtoull(a + 5, nullptr, 10);
      else paras.push_back(a);
    } else if(a[0] != '-') {
      input = a;
    } else {
      switch(a[1]) {
      case 'b':
        if(!more(i)) usage(a);
        p1 = (argv[i + 1][0] == 'p' || argv[i + 1][0] == 'P'); ++i;
        if(more(i)) { p2 = (argv[i + 1][0] == 'p' || argv[i + 1][0] == 'P'); ++i; }
        break;
      case 'c': if(!more(i)) usage(a); courant = atof(argv[++i]); break;
      case 'd': if(!more(i)) usage(a); dev = atoi(argv[++i]); break;
      case 'l':
        if(!more(i)) usage(a);
        l0 = atof(argv[++i]);
        if(more(i)) { L1 = L2 = atof(argv[++i]); if(more(i)) L2 = atof(argv[++i]); }
        break;
      case 'n':
        if(!more(i)) usage(a);
        n0 = atoi(argv[++i]);
        if(more(i)) { n1 = n2 = atoi(argv[++i]); if(more(i)) n2 = atoi(argv[++i]); }
        break;
      default: usage(a);
      }
    }
  }

  printf("2D finite grid code, MI355X time-step path (module %s)\n\n", module.c_str());
  int count = 0;
  if(hipGetDeviceCount(&count) != hipSuccess || dev >= count) die("  Device %d/%d   : does not exist, QUIT\n", dev, count);
  if(hipSetDevice(dev) != hipSuccess) die("fail to pick device, QUIT\n");
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  printf("  Device %d/%d   : \"%s\" with %gMiB global and %gKiB shared memory\n", dev, count,
         prop.name, prop.totalGlobalMem / 1048576.0, prop.sharedMemPerBlock / 1024.0);

  fg2_run *run = nullptr;
  size_t bytes = 0;
  printf("  Resolution   : %d x %d", n1, n2);
  if(fg2_setup(&run, module.c_str(), n1, n2, &bytes)) die(", fail to allocate memory, QUIT (%s)\n", fg2_last_error());
  printf(" using %.3gMiB (%.3g%%) of global memory\n", bytes / 1048576.0, 100.0 * bytes / prop.totalGlobalMem);
  if(fg2_set_geometry(run, L1, L2, p1, p2, courant)) die("%s, QUIT\n", fg2_last_error());
  for(const char *p : paras) {
    if(const char *r = fg2_para(run, p)) printf("  Parameter    : %s\n", r);
    else printf("  Fail to set  : \"%s\"\n", p);
  }

  int i;
  double t = 0.0;
  if(exists(input)) {                                        // src/main.cpp:104-115
    printf("  Input file   : \"%s\"...", input);
    fflush(stdout);
    i = frame_of(input);
    if(fg2_load(run, input, &t)) die("\n%s\n", fg2_last_error());
    fg2_config(run);
    printf("\b\b\b, start at frame %d, time %g\n", i, t);
  } else {
    printf("  Initialize   : \"%s\"\n", input);
    fg2_config(run);
    if(fg2_init(run, input, seed)) die("init: %s\n", fg2_last_error());
    i = 0;
    if(fg2_dump(run, frame_name(i).c_str(), t)) die("dump: %s\n", fg2_last_error());
  }

  printf("  Time         : %g with %d frame%s\n", l0, n0, n0 > 1 ? "s" : "");
  double flops, bits, alg;
  fg2_perf_model(run, &flops, &bits, &alg);
  const double t_final = t + l0;
  const int n = i + n0;
  printf("================================ Start Simulation ==============================\n");
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  for(const double dt_dump = (t_final - t) / (n - i); i++ < n;) {   // src/solve.cpp:32-58
    const double target = t_final - (n - i) * dt_dump;
    int m = 0;
    printf("%4d:%7.2f ->%7.2f", i, t, target);
    fflush(stdout);
    hip_check(hipEventRecord(e0, (hipStream_t)fg2_stream(run)), "hipEventRecord");
    while(t < target) {
      const double dt = std::fmin(target - t, fg2_getdt(run));
      ++m;
      if(int err = fg2_step(run, t, dt)) {
        printf(" crashed in %d step%s\n", m, m > 1 ? "s" : "");
        if(fg2_dump(run, "****.raw", t)) fprintf(stderr, "dump: %s\n", fg2_last_error());
        printf("================================ Abort Simulation ==============================\n");
        die("ERROR: %s\n", err == FG2_ERR_NAN ? "the state contains NaN" : fg2_last_error());
      }
      t = std::fmin(t + dt, target);
    }
    hip_check(hipEventRecord(e1, (hipStream_t)fg2_stream(run)), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    const double ns = ms * 1.0e6 / (m > 0 ? m : 1);
    printf(", %d step%s (%.3fms/step, %.3fGflops, %.3fGbps)\n", m, m > 1 ? "s" : "",
           1.0e-6 * ns, flops / ns, bits / ns);
    if(fg2_dump(run, frame_name(i).c_str(), t)) die("dump: %s\n", fg2_last_error());
  }
  printf("================================ Done  Simulation ==============================\n");
  fg2_free(run);
  hip_check(hipEventDestroy(e1), "hipEventDestroy");
  hip_check(hipEventDestroy(e0), "hipEventDestroy");
  return 0;
}
