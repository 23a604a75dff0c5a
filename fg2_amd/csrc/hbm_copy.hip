// hbm_copy.hip -- a plain device-to-device copy of known size, the
// calibration for HBM byte counters and the practical bandwidth ceiling.
//
//   hbm_copy [MiB] [reps]
//
// Copies MiB mebibytes (default 1024) from one buffer to another `reps`
// times (default 20) with a float4 grid-stride kernel and prints one JSON
// line: bytes read and written per launch and the achieved GB/s (read +
// write) from hipEvent timing.  Under `rocprofv3 --pmc FETCH_SIZE` (or
// WRITE_SIZE) the counter of `copy_kernel` divided by the bytes per launch
// gives the scale factor from counter units to HBM bytes.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_kernel(const float4 *__restrict__ a,
                                                   float4 *__restrict__ b, long n)
{
  for(long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = a[i];
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
  float4 *a, *b;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMemset(a, 0, bytes));
  CHECK(hipMemset(b, 0, bytes));
  int dev, cus;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int grid = cus * 8;
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  copy_kernel<<<grid, 256>>>(a, b, n);                     // warm-up
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(e0, 0));
  for(int k = 0; k < reps; ++k) copy_kernel<<<grid, 256>>>(a, b, n);
  CHECK(hipEventRecord(e1, 0));
  CHECK(hipEventSynchronize(e1));
  float ms;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  const double s = ms * 1e-3 / reps;
  printf("{\"kernel\": \"copy_kernel\", \"bytes_read_per_launch\": %ld, \"bytes_written_per_launch\": %ld, "
         "\"avg_launch_us\": %.3f, \"GBps_read_plus_write\": %.1f}\n",
         bytes, bytes, s * 1e6, 2.0 * bytes / s / 1e9);
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  return 0;
}
