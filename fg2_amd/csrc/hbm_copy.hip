// hbm_copy.hip -- streams of known size over HBM: the calibration for HBM
// byte counters and the practical bandwidth ceilings.
//
//   hbm_copy [MiB] [reps]
//
// Streams MiB mebibytes (default 1024) `reps` times (default 20) per
// configuration and prints one JSON line:
//   copy_kernel   a float4 grid-stride copy (reads MiB, writes MiB).  Under
//                 rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE) its counter per
//                 dispatch divided by the bytes per launch gives the scale
//                 factor from counter units to HBM bytes; the first
//                 configuration (8 workgroups per CU) is that calibration;
//   read_kernel   a float4 grid-stride read (a sum that is never stored);
//   write_kernel  a float4 grid-stride fill.
// Each kernel is timed with 2, 3, 4 and 5 workgroups of 256 threads per CU
// (and the copy also with 8), and the fastest is the ceiling.  Every
// configuration is timed after 100 untimed launches: the GPU's clocks drop
// after a few ms without work and take some 25 ms of sustained load to
// come back.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_kernel(const float4 *__restrict__ a,
                                                   float4 *__restrict__ b, long n)
{
  for(long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// The sum is stored only if it equals `never` (the buffer holds zeros), so
// the loads are not removed and nothing is written.
__global__ __launch_bounds__(256) void read_kernel(const float4 *__restrict__ a, float *out,
                                                   long n, float never)
{
  float s = 0.0f;
  for(long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float4 v = a[i];
    s += (v.x + v.y) + (v.z + v.w);
  }
  if(s == never) out[0] = s;
}

__global__ __launch_bounds__(256) void write_kernel(float4 *__restrict__ b, long n)
{
  for(long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
}

#define CHECK(x) do { hipError_t e = (x); if(e != hipSuccess) { \
  fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e)); return 1; } } while(0)

int main(int argc, char **argv)
{
  const long mib = argc > 1 ? atol(argv[1]) : 1024;
  const int reps = argc > 2 ? atoi(argv[2]) : 20;
  if(mib < 1 || mib > 65536 || reps < 1) { fprintf(stderr, "bad arguments\n"); return 2; }
  const long bytes = mib << 20;
  const long n = bytes / 16;
  int dev, cus;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  float4 *a, *b;
  float *out;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMalloc(&out, sizeof(float)));
  CHECK(hipMemset(a, 0, bytes));
  CHECK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const char *names[3] = {"copy", "read", "write"};
  const int per_cu[5] = {8, 2, 3, 4, 5};
  double best[3] = {0, 0, 0}, calib_us = 0;
  int best_wg[3] = {0, 0, 0};
  for(int kind = 0; kind < 3; ++kind)
    for(int c = (kind == 0 ? 0 : 1); c < 5; ++c) {
      const int grid = per_cu[c] * cus;
      auto launch = [&]() {
        if(kind == 0) copy_kernel<<<grid, 256>>>(a, b, n);
        else if(kind == 1) read_kernel<<<grid, 256>>>(a, out, n, 12345.0f);
        else write_kernel<<<grid, 256>>>(b, n);
      };
      for(int k = 0; k < 100; ++k) launch();               // warm-up: the clocks ramp
      CHECK(hipGetLastError());
      CHECK(hipDeviceSynchronize());
      CHECK(hipEventRecord(e0, 0));
      for(int k = 0; k < reps; ++k) launch();
      CHECK(hipEventRecord(e1, 0));
      CHECK(hipEventSynchronize(e1));
      float ms;
      CHECK(hipEventElapsedTime(&ms, e0, e1));
      const double s = ms * 1e-3 / reps;
      const double gbps = (kind == 0 ? 2.0 : 1.0) * bytes / s / 1e9;
      if(kind == 0 && c == 0) calib_us = s * 1e6;
      if(gbps > best[kind]) { best[kind] = gbps; best_wg[kind] = per_cu[c]; }
    }
  printf("{\"kernel\": \"copy_kernel\", \"bytes_read_per_launch\": %ld, \"bytes_written_per_launch\": %ld, "
         "\"calibration_launch_us\": %.3f, \"calibration_workgroups_per_cu\": 8",
         bytes, bytes, calib_us);
  for(int kind = 0; kind < 3; ++kind)
    printf(", \"%s\": {\"GBps\": %.1f, \"workgroups_per_cu\": %d}", names[kind], best[kind],
           best_wg[kind]);
  printf(", \"note\": \"GBps: bytes read plus bytes written per second, the fastest of the "
         "workgroup counts tried\"}\n");
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  CHECK(hipFree(out));
  return 0;
}
