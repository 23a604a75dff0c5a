// hbm_copy.hip -- a plain device-to-device copy of known size, the
// calibration for HBM byte counters and the practical bandwidth ceiling.
//
//   hbm_copy [MiB] [reps]
//
// Copies MiB mebibytes (default 1024) from one buffer to another `reps`
// times (default 20) per kernel and prints one JSON line: bytes read and
// written per launch and the achieved GB/s (read + write) from hipEvent
// timing, for two kernels:
//   copy_kernel           float4 grid-stride loop, one load in flight per
//                         thread (the counter calibration: under rocprofv3
//                         --pmc FETCH_SIZE or WRITE_SIZE its counter divided
//                         by the bytes per launch gives the scale factor from
//                         counter units to HBM bytes);
//   copy_unrolled_kernel  each thread moves 4 float4 per trip with
//                         non-temporal loads and stores, so 4 loads are in
//                         flight (the practical ceiling).
// Both are timed after 200 untimed launches: the GPU's clocks drop after a
// few ms without work and take some 25 ms of sustained load to come back.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_kernel(const float4 *__restrict__ a,
                                                   float4 *__restrict__ b, long n)
{
  for(long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

typedef float v4f __attribute__((ext_vector_type(4)));

// n4: 16-byte elements, a multiple of 4 * 256 * gridDim.x (the caller sizes
// the buffer so), so every trip of every thread is in bounds.
__global__ __launch_bounds__(256) void copy_unrolled_kernel(const v4f *__restrict__ a,
                                                            v4f *__restrict__ b, long n4)
{
  const long stride = (long)gridDim.x * 256;
  for(long base = blockIdx.x * 256L + threadIdx.x; base < n4; base += 4 * stride) {
    v4f x[4];
#pragma unroll
    for(int k = 0; k < 4; ++k) x[k] = __builtin_nontemporal_load(a + base + k * stride);
#pragma unroll
    for(int k = 0; k < 4; ++k) __builtin_nontemporal_store(x[k], b + base + k * stride);
  }
}

#define CHECK(x) do { hipError_t e = (x); if(e != hipSuccess) { \
  fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e)); return 1; } } while(0)

int main(int argc, char **argv)
{
  const long mib = argc > 1 ? atol(argv[1]) : 1024;
  const int reps = argc > 2 ? atoi(argv[2]) : 20;
  if(mib < 1 || mib > 65536 || reps < 1) { fprintf(stderr, "bad arguments\n"); return 2; }
  int dev, cus;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int grid = cus * 8;
  // round the size to whole trips of the unrolled kernel (4 float4 per
  // thread of the grid), which the plain kernel copies as well
  const long trip = 4L * 256 * grid * 16;                  // bytes
  const long bytes = ((mib << 20) + trip - 1) / trip * trip;
  const long n = bytes / 16;
  float4 *a, *b;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMemset(a, 0, bytes));
  CHECK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  float ms[2];
  for(int kind = 0; kind < 2; ++kind) {
    auto launch = [&]() {
      if(kind == 0) copy_kernel<<<grid, 256>>>(a, b, n);
      else copy_unrolled_kernel<<<grid, 256>>>((const v4f *)a, (v4f *)b, n);
    };
    for(int k = 0; k < 200; ++k) launch();                 // warm-up: the clocks ramp
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0, 0));
    for(int k = 0; k < reps; ++k) launch();
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms[kind], e0, e1));
  }
  const double s0 = ms[0] * 1e-3 / reps, s1 = ms[1] * 1e-3 / reps;
  printf("{\"kernel\": \"copy_kernel\", \"bytes_read_per_launch\": %ld, \"bytes_written_per_launch\": %ld, "
         "\"avg_launch_us\": %.3f, \"GBps_read_plus_write\": %.1f, "
         "\"unrolled\": {\"kernel\": \"copy_unrolled_kernel\", \"avg_launch_us\": %.3f, "
         "\"GBps_read_plus_write\": %.1f}}\n",
         bytes, bytes, s0 * 1e6, 2.0 * bytes / s0 / 1e9, s1 * 1e6, 2.0 * bytes / s1 / 1e9);
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  return 0;
}
