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
// The libc rand() calls of KH and Kf are replaced by a counter-based stream:
// draw number q (0, 1, 2, ...) of a run seeded with S is the (q+1)-th output
// of SplitMix64 started at state S, so a sequential host loop that draws in
// the reference's order (one draw per cell for KH, two for Kf, row-major
// over i then j) reproduces it exactly.  The 53-bit draw is centred,
// ((z >> 11) + 1/2) 2^-53, so it lies in (0, 1) and log(r1) in Kf is finite
// (the reference's rand() / RAND_MAX can be 0).
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>

#include "fg2_layout.h"
#include "launch.h"

namespace fg2 {

namespace {

// Round every FP64 operation on its own (no fused multiply-adds), so a
// host evaluation with the same formulas lands on the same doubles up to
// the math library's last-place differences.
#pragma clang fp contract(off)

__device__ __forceinline__ double draw(uint64_t seed, uint64_t q)
{
  uint64_t z = seed + (q + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((double)(z >> 11) + 0.5) * 0x1.0p-53;
}

__device__ __forceinline__ double fermi_dirac(double a, double d) { return 1.0 / (exp(a / d) + 1.0); }

struct IcArgs {
  float *u;
  long s;
  int n1, n2;
  float l1, l2;
  float par[16];
  uint64_t seed;
  int ic;
};

// src/hydro/init.h:26-112: (lnd, u1, u2, lne)
__device__ void hydro_ic(const IcArgs &A, int i, int j, float x, float y, float *c)
{
  const double g = A.par[0];
  // the centre shift is stored back into the R-typed argument (init.h:30-31)
  const double xx = (float)((double)x - 0.5 * A.l1), yy = (float)((double)y - 0.5 * A.l2);
  double d = 1.0, u = 0.0, e = 1.0;
  switch(A.ic) {
  case IC_BLAST:
    e = ((xx * xx + yy * yy < 0.01) ? 1.0 : 0.01) / (g - 1.0);
    break;
  case IC_BOW: {
    const double r = sqrt(xx * xx + yy * yy);
    d = 9.9 * fermi_dirac(r - 0.01, 0.001) + 0.1;
    u = (10.0 / d - 1.0) / 99.0;
    e = 0.01 / d / (g - 1.0);
    break;
  }
  case IC_IMPLOS: {
    const double f = fermi_dirac(0.15 - fabs(xx) - fabs(yy), 0.001);
    d = 0.125 + 0.875 * f;
    e = (0.14 + 0.86 * f) / d / (g - 1.0);
    break;
  }
  case IC_KH: {
    const double r = draw(A.seed, (uint64_t)i * A.n2 + j);
    if(fabs(yy) < 0.25) { d = 2.0; u = 0.5 + 0.01 * r - 0.005; }
    else                { d = 1.0; u = -0.5 + 0.01 * r - 0.005; }
    e = 2.5 / d / (g - 1.0);
    break;
  }
  case IC_SOD: {
    double P;
    if(fmod(xx - 2.0 * yy + 1.5, 1.0) < 0.5) { d = 1.0; P = 1.0; }
    else                                     { d = 0.125; P = 0.1; }
    e = P / d / (g - 1.0);
    break;
  }
  default:                                     // zeros
    c[0] = c[1] = c[2] = c[3] = 0.0f;
    return;
  }
  c[0] = (float)log(d); c[1] = (float)u; c[2] = 0.0f; c[3] = (float)log(e);
}

// src/hydro-isot/init.h:21-49: (lnd, u1, u2)
__device__ void isot_ic(const IcArgs &A, int i, int j, float x, float *c)
{
  if(A.ic != IC_KF) { c[0] = c[1] = c[2] = 0.0f; return; }
  const double nus = A.par[1], fi = A.par[4], ki = A.par[5];
  const uint64_t q = 2 * ((uint64_t)i * A.n2 + j);
  const double r1 = draw(A.seed, q), r2 = draw(A.seed, q + 1);
  const double rad = sqrt(-2.0 * log(r1));
  const double q1 = rad * sin(2.0 * M_PI * r2);
  const double q2 = rad * cos(2.0 * M_PI * r2);
  const double a = fi / (nus * ki * ki);
  c[0] = 0.0f;
  c[1] = (float)(1.0e-9 * q1);
  c[2] = (float)(1.0e-9 * q2 + a * cos(ki * (double)x));
}

// src/disk/init.h:25-68: hydrostatic torus of Hawley (2000), (lnd, ur, uz, Omg, lne, lner)
__device__ void disk_ic(const IcArgs &A, float lnr, float theta, float *c)
{
  const double M = A.par[0], rin = A.par[2], G = A.par[3], ar = A.par[8];
  const double r = rin * exp((double)lnr), sn = sin((double)theta), cyl = r * sn;
  const double q0 = 2.0, r0 = 16.0, d0 = 1.0, e0 = 0.01, bg = 0.0001;
  const double g1 = G - 1.0, q1 = 2.0 * q0 - 2.0;
  const double lK = sqrt(M * pow(r0, q1 - 1.0));
  const double K = (g1 * e0) / pow(d0, g1);
  double den = G * e0 - M / r0 + lK * lK / (pow(r0, q1) * q1) + M / r - lK * lK / (pow(cyl, q1) * q1);
  den = den > 0.0 ? pow(den * g1 / (G * K), 1.0 / g1) : -pow(-den * g1 / (G * K), 1.0 / g1);
  double omg = lK * pow(cyl, -q0);
  if(den < bg) { den = bg; omg *= sn * sn; }
  const double eng = K * pow(den, g1) / g1;
  const double tbg = K * pow(bg, g1);
  const double rad = ar * tbg * tbg * tbg * tbg / bg;
  c[0] = (float)log(den); c[1] = 0.0f; c[2] = 0.0f; c[3] = (float)omg;
  c[4] = (float)log(eng); c[5] = (float)log(rad);
}

__global__ __launch_bounds__(256) void ic_kernel(IcArgs A, int module)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if(j >= A.n2) return;
  // cell centres in R (src/init.cpp:31-34)
  const float x = (float)((double)A.l1 * (i + 0.5) / A.n1);
  const float y = (float)((double)A.l2 * (j + 0.5) / A.n2);
  float c[6];
  int nv;
  if(module == HYDRO)           { hydro_ic(A, i, j, x, y, c); nv = 4; }
  else if(module == HYDRO_ISOT) { isot_ic(A, i, j, x, c);     nv = 3; }
  else                          { disk_ic(A, x, y, c);         nv = 6; }
  float *p = A.u + (long)i * A.s + (long)j * nv;
  for(int k = 0; k < nv; ++k) p[k] = c[k];
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
  ic_kernel<<<grid, 256, 0, st>>>(A, module);
  return hipGetLastError();
}

} // namespace fg2
