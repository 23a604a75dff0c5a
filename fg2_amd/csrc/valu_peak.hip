// valu_peak.hip -- the practical FP32 VALU issue ceiling of the GPU.
//
//   valu_peak [iters]
//
// Every lane runs 8 independent chains of FP32 fused multiply-adds (packed
// FP32 is off in this build, so each is one v_fmac_f32 wave-instruction) for
// `iters` x 8 steps, with 4, 8 and 16 workgroups of 256 threads per CU, and
// prints one JSON line: FMA wave-instructions per second of each
// configuration against 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles = 614.4 G/s
// (a wave64 FP32 instruction holds a 16-lane SIMD for 4 cycles).  Each
// configuration is timed after 50 untimed launches (the clocks ramp under
// sustained load).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void fma_kernel(float *out, int iters, float a, float b)
{
  float x[8];
#pragma unroll
  for(int c = 0; c < 8; ++c) x[c] = threadIdx.x * 1e-3f + c;
  for(int k = 0; k < iters; ++k) {
#pragma unroll
    for(int u = 0; u < 8; ++u)
#pragma unroll
      for(int c = 0; c < 8; ++c) x[c] = __builtin_fmaf(x[c], a, b);
  }
  float s = 0.0f;
#pragma unroll
  for(int c = 0; c < 8; ++c) s += x[c];
  if(s == 1234.5f) out[0] = s;              // keeps the chains; never true here
}

#define CHECK(x) do { hipError_t e = (x); if(e != hipSuccess) { \
  fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e)); return 1; } } while(0)

int main(int argc, char **argv)
{
  const int iters = argc > 1 ? atoi(argv[1]) : 4096;
  if(iters < 1 || iters > (1 << 20)) { fprintf(stderr, "bad arguments\n"); return 2; }
  int dev, cus;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  float *out;
  CHECK(hipMalloc(&out, sizeof(float)));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const double peak = 256.0 * 4 * 2.4e9 / 4;
  printf("{\"kernel\": \"fma_kernel\", \"peak_G_per_s\": %.1f", peak / 1e9);
  for(int per_cu : {4, 8, 16}) {
    const int grid = cus * per_cu, reps = 20;
    for(int k = 0; k < 50; ++k) fma_kernel<<<grid, 256>>>(out, iters, 0.999f, 0.001f);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0, 0));
    for(int k = 0; k < reps; ++k) fma_kernel<<<grid, 256>>>(out, iters, 0.999f, 0.001f);
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double s = ms * 1e-3 / reps;
    const double insts = (double)grid * 4 * iters * 64;      // waves x FMAs per lane
    printf(", \"workgroups_per_cu_%d\": {\"G_per_s\": %.1f, \"frac\": %.3f}", per_cu,
           insts / s / 1e9, insts / s / peak);
  }
  printf("}\n");
  CHECK(hipFree(out));
  return 0;
}
