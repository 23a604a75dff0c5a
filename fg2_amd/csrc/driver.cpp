// driver.cpp -- command-line front end over the C ABI (fg2run).
//
//   fg2run MODULE [--help] [-b p1 [p2]] [-c C] [-d dev] [-l T [L1 [L2]]]
//          [-n frames [N1 [N2]]] [name=val ...] [seed=S] [INPUT]
//
// Option meanings and the run sequence follow the reference's main()
// (src/main.cpp:29-120) and solve() (src/solve.cpp:25-64); snapshots are the
// reference's .raw files (src/io.cpp:24-76).  The module is an argument
// here, where the reference builds one binary per module (Makefile:19-22).
//   - An INPUT naming an existing file is read as a snapshot; the run goes on
//     from the frame number of its first digit run (src/misc.cpp:48-55).
//     Any other INPUT names an initial condition, and 0000.raw is written.
//   - Frame k ends at t_final - (frames - k) T / frames; within a frame every
//     step takes dt = min(frame end - t, getdt()), and each frame ends with
//     a NNNN.raw snapshot.
//   - A step that leaves NaN in the state writes ****.raw at that step's
//     start time and exits with status 255.
// The start time of a resumed run stays FP64 (the reference narrows it to
// R, src/main.cpp:34,106), so a resumed run reaches exactly the frame times
// of a run that never stopped.  The libc rand() of the KH and Kf initial
// conditions is fg2_init's seeded stream (seed=S, default 0).
#include <hip/hip_runtime_api.h>
#include <cctype>
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

[[noreturn]] void quit(const char *fmt, ...)
{
  fflush(stdout);
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  exit(-1);
}

[[noreturn]] void help(const char *bad)
{
  if(bad)
    quit("fg2run: illegal option \"%s\" (fg2run --help lists the options)\n", bad);
  printf("Usage: fg2run MODULE [OPTION...] [INPUT_FILE]\n"
         "Advance a 2D grid of the chosen physics module on an MI355X and write\n"
         "a .raw snapshot per frame.\n\n"
         "  MODULE             hydro | hydro-isot | disk\n"
         "  --help             print this text\n"
         "  -b P1 [P2]         boundary of dim 1 [and dim 2]: p or P periodic, else closed\n"
         "  -c C               Courant factor of the time step\n"
         "  -d DEV             HIP device index\n"
         "  -l T [L1 [L2]]     simulated time span [and box lengths]\n"
         "  -n F [N1 [N2]]     frames to write [and grid cells per side]\n"
         "  NAME=VALUE         module parameter (gamma=1.4, nus=1e-4, ...)\n"
         "  seed=S             seed of the random part of an initial condition\n"
         "  INPUT_FILE         snapshot to resume from, or the initial condition's name\n");
  exit(0);
}

void need(hipError_t e, const char *what)
{
  if(e != hipSuccess) quit("%s failed: %s\n", what, hipGetErrorString(e));
}

bool file_exists(const char *name)
{
  if(FILE *f = fopen(name, "r")) { fclose(f); return true; }
  return false;
}

// frame number of a snapshot name: its first run of digits (src/misc.cpp:48-55)
int frame_number(const char *name)
{
  const char *d = name;
  while(*d && (*d < '0' || *d > '9')) ++d;
  return *d ? atoi(d) : 0;
}

std::string snapshot_name(int frame)
{
  char b[32];
  snprintf(b, sizeof b, "%04d.raw", frame);
  return b;
}

// an argument that begins a new option: '-' followed by a letter
bool starts_option(const char *a) { return a[0] == '-' && isalpha((unsigned char)a[1]); }

struct Options {
  std::string module;
  const char *input = "default";
  int frames = 10, n1 = 1024, n2 = 1024, device = 0, p1 = 1, p2 = 1;   // src/main.cpp:33-34
  double span = 1.0, L1 = 1.0, L2 = 1.0, courant = 0.5;               // src/global.cpp:28-29
  uint64_t seed = 0;
  std::vector<const char *> assignments;
};

Options parse(int argc, char **argv)                        // src/main.cpp:41-60
{
  Options o;
  o.module = argv[1];
  if(o.module != "hydro" && o.module != "hydro-isot" && o.module != "disk")
    quit("fg2run: unknown module \"%s\"; the modules are hydro, hydro-isot and disk\n",
         o.module.c_str());
  // values that follow option argv[i]: up to `most`, stopping at the next option
  auto values = [&](int &i, int most) {
    std::vector<const char *> v;
    while((int)v.size() < most && i + 1 < argc && !starts_option(argv[i + 1])) v.push_back(argv[++i]);
    return v;
  };
  auto periodic = [](const char *s) { return s[0] == 'p' || s[0] == 'P'; };
  for(int i = 2; i < argc; ++i) {
    const char *a = argv[i];
    if(strchr(a, '=')) {
      if(!strncmp(a, "seed=", 5)) o.seed = strtoull(a + 5, nullptr, 10);
      else o.assignments.push_back(a);
      continue;
    }
    if(a[0] != '-') { o.input = a; continue; }
    std::vector<const char *> v;
    switch(a[1]) {
    case 'b':
      v = values(i, 2);
      if(v.empty()) help(a);
      o.p1 = periodic(v[0]);
      if(v.size() > 1) o.p2 = periodic(v[1]);
      break;
    case 'c':
      v = values(i, 1);
      if(v.empty()) help(a);
      o.courant = atof(v[0]);
      break;
    case 'd':
      v = values(i, 1);
      if(v.empty()) help(a);
      o.device = atoi(v[0]);
      break;
    case 'l':
      v = values(i, 3);
      if(v.empty()) help(a);
      o.span = atof(v[0]);
      if(v.size() > 1) o.L1 = o.L2 = atof(v[1]);
      if(v.size() > 2) o.L2 = atof(v[2]);
      break;
    case 'n':
      v = values(i, 3);
      if(v.empty()) help(a);
      o.frames = atoi(v[0]);
      if(v.size() > 1) o.n1 = o.n2 = atoi(v[1]);
      if(v.size() > 2) o.n2 = atoi(v[2]);
      break;
    default:
      help(a);
    }
  }
  return o;
}

const char *plural(int k) { return k == 1 ? "" : "s"; }

} // namespace

int main(int argc, char **argv)
{
  for(int i = 1; i < argc; ++i)
    if(!strcmp(argv[i], "--help")) help(nullptr);
  if(argc < 2) help(nullptr);
  const Options o = parse(argc, argv);

  printf("fg2run: module %s on an MI355X\n", o.module.c_str());
  int ndev = 0;
  if(hipGetDeviceCount(&ndev) != hipSuccess || o.device < 0 || o.device >= ndev)
    quit("fg2run: no HIP device %d (%d visible)\n", o.device, ndev);
  need(hipSetDevice(o.device), "hipSetDevice");
  hipDeviceProp_t prop;
  need(hipGetDeviceProperties(&prop, o.device), "hipGetDeviceProperties");
  printf("  device       %d of %d, %s, %.0f MiB\n", o.device, ndev, prop.name,
         prop.totalGlobalMem / 1048576.0);

  fg2_run *run = nullptr;
  size_t bytes = 0;
  if(fg2_setup(&run, o.module.c_str(), o.n1, o.n2, &bytes))
    quit("fg2run: cannot set up a %d x %d grid: %s\n", o.n1, o.n2, fg2_last_error());
  printf("  grid         %d x %d cells, %.1f MiB of device memory\n", o.n1, o.n2,
         bytes / 1048576.0);
  if(fg2_set_geometry(run, o.L1, o.L2, o.p1, o.p2, o.courant))
    quit("fg2run: %s\n", fg2_last_error());
  for(const char *a : o.assignments) {
    if(const char *r = fg2_para(run, a)) printf("  parameter    %s\n", r);
    else printf("  ignored      \"%s\" (%s)\n", a, fg2_last_error());
  }

  int frame;
  double t = 0.0;
  if(file_exists(o.input)) {                                 // src/main.cpp:104-115
    frame = frame_number(o.input);
    if(fg2_load(run, o.input, &t)) quit("fg2run: %s\n", fg2_last_error());
    fg2_config(run);
    printf("  resume       \"%s\": frame %d, t = %g\n", o.input, frame, t);
  } else {
    fg2_config(run);
    if(fg2_init(run, o.input, o.seed)) quit("fg2run: initial condition: %s\n", fg2_last_error());
    frame = 0;
    printf("  initial      \"%s\"\n", o.input);
    if(fg2_dump(run, snapshot_name(frame).c_str(), t)) quit("fg2run: %s\n", fg2_last_error());
  }
  printf("  span         %g in %d frame%s\n", o.span, o.frames, plural(o.frames));

  double flops = 0, bits = 0, alg = 0;
  fg2_perf_model(run, &flops, &bits, &alg);
  hipStream_t st = (hipStream_t)fg2_stream(run);
  hipEvent_t e0, e1;
  need(hipEventCreate(&e0), "hipEventCreate");
  need(hipEventCreate(&e1), "hipEventCreate");
  const double t_final = t + o.span;
  const int last = frame + o.frames;
  const double frame_span = (t_final - t) / o.frames;
  printf("-- run ------------------------------------------------------------------\n");
  while(frame < last) {                                      // src/solve.cpp:32-58
    ++frame;
    const double end = t_final - (last - frame) * frame_span;
    printf("  frame %4d  t %9.4f -> %9.4f", frame, t, end);
    fflush(stdout);
    int steps = 0;
    need(hipEventRecord(e0, st), "hipEventRecord");
    for(; t < end; ++steps) {
      const double dt = std::fmin(end - t, fg2_getdt(run));
      if(int err = fg2_step(run, t, dt)) {
        printf("  FAILED at step %d of the frame\n", steps + 1);
        if(fg2_dump(run, "****.raw", t)) fprintf(stderr, "fg2run: %s\n", fg2_last_error());
        printf("-- stopped: state written to ****.raw ------------------------------------\n");
        quit("fg2run: %s\n", err == FG2_ERR_NAN ? "the state contains NaN" : fg2_last_error());
      }
      t = std::fmin(t + dt, end);
    }
    need(hipEventRecord(e1, st), "hipEventRecord");
    need(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    need(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    const double s_per_step = 1e-3 * ms / (steps > 0 ? steps : 1);
    // the reference's model rates (src/config.cpp:30-36) beside the step time
    printf("  %d step%s, %.3f ms/step, model %.1f GFLOP/s %.1f GB/s\n", steps, plural(steps),
           1e3 * s_per_step, 1e-9 * flops / s_per_step, 1e-9 * bits / 8 / s_per_step);
    if(fg2_dump(run, snapshot_name(frame).c_str(), t)) quit("fg2run: %s\n", fg2_last_error());
  }
  printf("-- done -----------------------------------------------------------------\n");
  fg2_free(run);
  need(hipEventDestroy(e1), "hipEventDestroy");
  need(hipEventDestroy(e0), "hipEventDestroy");
  return 0;
}
