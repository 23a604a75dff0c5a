// stencil.h -- the 6th-order operators every physics module builds its
// right-hand side from, and the types of the module interface (physics.h).
//
// Every kernel (plain kick, fused substage, one-pass step, batched)
// evaluates a module's right-hand side from the same stencil primitives,
// with explicit FMAs in the operators.  The per-cell instruction sequence
// therefore does not depend on which kernel, tile shape or batch a cell is
// computed in, and the paths agree bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "fg2_layout.h"

namespace fg2 {

// Stencil primitives of one cell.  For variable q and k = 1..3:
//   dv[k-1][q] = f(i+k, j) - f(i-k, j)      sv = f(i+k, j) + f(i-k, j)
//   dh[k-1][q] = f(i, j+k) - f(i, j-k)      sh = f(i, j+k) + f(i, j-k)
//   mx[k-1][d] = (f(i+k,j+k) - f(i-k,j+k)) - (f(i+k,j-k) - f(i-k,j-k))
// mx equals the reference's _MIX_(x,k,k) (src/deriv.h:36-37) up to the
// association of the four-term sum.
template<int NV, int ND> struct Prim {
  float c[NV];
  float dv[3][NV], sv[3][NV], dh[3][NV], sh[3][NV];
  float mx[3][ND];
};

// 6th-order operators of src/deriv.h:41-63 with 1/Delta folded into the
// coefficients (Params::a1, b1, m12 ...).
__device__ __forceinline__ float op_d(const float a[3], float d0, float d1, float d2)
{ return __builtin_fmaf(a[2], d2, __builtin_fmaf(a[1], d1, a[0] * d0)); }

__device__ __forceinline__ float op_dd(const float b[4], float c, float s0, float s1, float s2)
{ return __builtin_fmaf(b[3], s2, __builtin_fmaf(b[2], s1, __builtin_fmaf(b[1], s0, b[0] * c))); }

template<class Pr> __device__ __forceinline__ float D1 (const Params &P, const Pr &x, int q) { return op_d(P.a1, x.dv[0][q], x.dv[1][q], x.dv[2][q]); }
template<class Pr> __device__ __forceinline__ float D2 (const Params &P, const Pr &x, int q) { return op_d(P.a2, x.dh[0][q], x.dh[1][q], x.dh[2][q]); }
template<class Pr> __device__ __forceinline__ float D11(const Params &P, const Pr &x, int q) { return op_dd(P.b1, x.c[q], x.sv[0][q], x.sv[1][q], x.sv[2][q]); }
template<class Pr> __device__ __forceinline__ float D22(const Params &P, const Pr &x, int q) { return op_dd(P.b2, x.c[q], x.sh[0][q], x.sh[1][q], x.sh[2][q]); }
template<class Pr> __device__ __forceinline__ float D12(const Params &P, const Pr &x, int d) { return op_d(P.m12, x.mx[0][d], x.mx[1][d], x.mx[2][d]); }
// D11 + D22 in one chain: (b1[0] + b2[0]) f0 + sum_k b1[k] sv_k + b2[k] sh_k
template<class Pr> __device__ __forceinline__ float LAP(const Params &P, const Pr &x, int q)
{
  float a = P.blap0 * x.c[q];
#pragma unroll
  for(int k = 0; k < 3; ++k) {
    a = __builtin_fmaf(P.b1[k + 1], x.sv[k][q], a);
    a = __builtin_fmaf(P.b2[k + 1], x.sh[k][q], a);
  }
  return a;
}

constexpr float THIRD = 1.0f / 3.0f;

// ---- types of the module interface ---------------------------------------

// A module parameter and its reference default (src/<module>/eqns.cuh).
struct ParamDef { const char *name; double value; };

// The geometry a module's settings() may impose (src/<module>/config.h).
struct Geometry { double l1, l2; int p1, p2; };

// Arguments of an initial condition: the state, its grid, box and FP32
// parameters, and the seed of its random stream.
struct IcArgs {
  float *u;
  long s;
  int n1, n2;
  float l1, l2;
  float par[16];
  uint64_t seed;
  int ic;                      // the module's own number of the named condition
};

// Draw number q (0, 1, 2, ...) of an initial condition's random stream
// seeded with S: the (q+1)-th output of SplitMix64 started at state S,
// centred, ((z >> 11) + 1/2) 2^-53, so it lies in (0, 1).  A sequential
// host loop drawing in the reference's order reproduces it exactly.
__device__ __forceinline__ double ic_draw(uint64_t seed, uint64_t q)
{
  uint64_t z = seed + (q + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((double)(z >> 11) + 0.5) * 0x1.0p-53;
}

__device__ __forceinline__ double fermi_dirac(double a, double d) { return 1.0 / (exp(a / d) + 1.0); }

} // namespace fg2
