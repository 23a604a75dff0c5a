// ics.hip -- the reference's initial conditions, evaluated on the device.
//
// init() (src/init.cpp:26-42) evaluates a module's IC function at the cell
// centres x = l1 (i + 1/2) / n1, y = l2 (j + 1/2) / n2 (R-typed, so rounded
// to FP32) on the host and copies the result up.  Here one thread per cell
// evaluates the same function in FP64 from the FP32 coordinates and FP32
// parameters (the reference reads them back from its __constant__ R
// parameters) and stores FP32 into the run's layout; no host loop and no
// host-to-device copy of a 12.8 M-cell state.
//
// The libc rand() calls of KH and Kf are replaced by a counter-based stream
// (ic_draw, stencil.h): draw number q of a run seeded with S is the (q+1)-th
// output of SplitMix64 started at state S, so a sequential host loop that
// draws in the reference's order (one draw per cell for KH, two for Kf,
// row-major over i then j) reproduces it exactly.  The 53-bit draw is
// centred, so it lies in (0, 1) and log(r1) in Kf is finite (the
// reference's rand() / RAND_MAX can be 0).  Each module's ic() is its
// init.h, evaluated with every FP64 operation rounded on its own.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>

#include "launch.h"
#include "physics.h"

namespace fg2 {

namespace {

template<class M>
__global__ __launch_bounds__(256) void ic_kernel(IcArgs A)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if(j >= A.n2) return;
  // cell centres in R (src/init.cpp:31-34)
  const float x = (float)((double)A.l1 * (i + 0.5) / A.n1);
  const float y = (float)((double)A.l2 * (j + 0.5) / A.n2);
  float c[M::NV];
  M::ic(A, i, j, x, y, c);
  float *p = A.u + (long)i * A.s + (long)j * M::NV;
  for(int k = 0; k < M::NV; ++k) p[k] = c[k];
}

} // namespace

hipError_t launch_ic(int module, int ic, float *u, long s, int n1, int n2, float l1, float l2,
                     const float *par, unsigned long long seed, hipStream_t st)
{
  IcArgs A{};
  A.u = u; A.s = s; A.n1 = n1; A.n2 = n2; A.l1 = l1; A.l2 = l2;
  memcpy(A.par, par, sizeof A.par);
  A.seed = seed;
  A.ic = ic;
  const dim3 grid((n2 + 255) / 256, n1);
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    ic_kernel<M><<<grid, 256, 0, st>>>(A);
    return hipGetLastError();
  });
}

} // namespace fg2
