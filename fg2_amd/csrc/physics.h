// physics.h -- the three physics modules on the device (FP32).
//
// Every kernel (plain kick, fused substage, one-pass step, batched) evaluates
// the right-hand side through M::rhs() below from the same stencil
// primitives, with explicit FMAs in the operators.  The per-cell instruction
// sequence therefore does not depend on which kernel, tile shape or batch a
// cell is computed in, and the paths agree bit for bit.
//
// Module interface mirrored from the reference (src/<module>/):
//   NV        number of variables, NVAR = sizeof(S)/sizeof(R)  (src/fg2.h:47-48)
//   ODD       variable negated by the reflective dim-2 boundary (src/*/bcond.cuh)
//   ND, D12V  variables whose mixed derivative D12 the RHS uses
//   rhs()     the module's eqns()                             (src/*/eqns.cuh)
//   first_order() the point-wise pre-step                     (src/*/Euler.cuh)
//   FLOP      eqns_flop of the reference's perf model          (src/*/config.h)
#pragma once
#include <hip/hip_runtime.h>
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

// ---------------------------------------------------------------------------
// hydro: (lnd, u1, u2, lne); src/hydro/eqns.cuh:19-92.
// par: gamma nus nub kappa dd
// ---------------------------------------------------------------------------
struct Hydro {
  static constexpr int NV = 4, ODD = 2, ND = 2, FLOP = 289, ID = HYDRO;
  static constexpr int D12V[ND] = {1, 2};
  using P_t = Prim<NV, ND>;

  static __device__ __forceinline__ void rhs(const Params &P, const P_t &x, int, int, float dt[NV])
  {
    const float gamma = P.par[0], nus = P.par[1], nub = P.par[2], kappa = P.par[3], dd = P.par[4];
    float d1[NV], d2[NV];
#pragma unroll
    for(int q = 0; q < NV; ++q) { d1[q] = D1(P, x, q); d2[q] = D2(P, x, q); }
    const float u1 = x.c[1], u2 = x.c[2];
#pragma unroll
    for(int q = 0; q < NV; ++q) dt[q] = -(u1 * d1[q] + u2 * d2[q]);     // advection

    const float gamma1 = gamma - 1.0f;                                     // pressure
    const float temp   = gamma1 * __expf(x.c[3]);
    const float div_u  = d1[1] + d2[2];
    dt[0] -= div_u;
    dt[1] -= temp * (d1[0] + d1[3]);
    dt[2] -= temp * (d2[0] + d2[3]);
    dt[3] -= div_u * gamma1;

    const float s11 = d1[1] - div_u * THIRD;                              // viscous heating
    const float s12 = (d1[2] + d2[1]) * 0.5f;
    const float s22 = d2[2] - div_u * THIRD;
    const float two_nus = 2.0f * nus;
    dt[1] += two_nus * (s11 * d1[0] + s12 * d2[0]);
    dt[2] += two_nus * (s12 * d1[0] + s22 * d2[0]);
    dt[3] += (gamma1 * __builtin_amdgcn_rcpf(temp)) *
      (two_nus * (s11 * s11 + 2.0f * s12 * s12 + s22 * s22) + nub * div_u * div_u);

    const float d11_u1 = D11(P, x, 1), d11_u2 = D11(P, x, 2);             // viscosity
    const float d12_u1 = D12(P, x, 0), d12_u2 = D12(P, x, 1);
    const float d22_u1 = D22(P, x, 1), d22_u2 = D22(P, x, 2);
    const float mixed  = nus * THIRD + nub;
    dt[1] += nus * (d11_u1 + d22_u1) + mixed * (d11_u1 + d12_u2);
    dt[2] += nus * (d11_u2 + d22_u2) + mixed * (d12_u1 + d22_u2);

    const float d_lne_2 = d1[3] * d1[3] + d2[3] * d2[3];                  // diffusion
    const float ed      = gamma1 * kappa;
    // dd = 0 (the default) contributes 0 * (finite) = 0: skip the work
    if(dd != 0.0f) {
      const float d_lnd_2 = d1[0] * d1[0] + d2[0] * d2[0];
      dt[0] += dd * (LAP(P, x, 0) + d_lnd_2);
    }
    dt[3] += ed * (LAP(P, x, 3) + d_lne_2);
  }

  static __device__ __forceinline__ void first_order(const Params &, float *, float, int, int) {}
  static constexpr bool HAS_FIRST_ORDER = false;

  // Advective rate for the CFL step: (|u1| + cs) k/D1 + (|u2| + cs) k/D2,
  // cs^2 = gamma (gamma - 1) e.
  static __device__ __forceinline__ float lam_adv(const Params &P, const float *c, int, int,
                                                  float k1a, float k1b)
  {
    const float g = P.par[0];
    const float cs = sqrtf(g * (g - 1.0f) * expf(c[3]));
    return (fabsf(c[1]) + cs) * k1a + (fabsf(c[2]) + cs) * k1b;
  }
  static __device__ __forceinline__ float lam_diff(const Params &, const float *, int, int) { return 0.0f; }
  static constexpr bool CELL_DIFF = false;
};

// ---------------------------------------------------------------------------
// hydro-isot: (lnd, u1, u2); src/hydro-isot/eqns.cuh:19-87, Euler.cuh:21-44.
// par: dd nus nub cs2 f k
// ---------------------------------------------------------------------------
struct HydroIsot {
  static constexpr int NV = 3, ODD = 2, ND = 2, FLOP = 214, ID = HYDRO_ISOT;
  static constexpr int D12V[ND] = {1, 2};
  using P_t = Prim<NV, ND>;

  static __device__ __forceinline__ void rhs(const Params &P, const P_t &x, int i, int, float dt[NV])
  {
    const float dd = P.par[0], nus = P.par[1], nub = P.par[2], cs2 = P.par[3];
    const float ff = P.par[4], kk = P.par[5];
    float d1[NV], d2[NV];
#pragma unroll
    for(int q = 0; q < NV; ++q) { d1[q] = D1(P, x, q); d2[q] = D2(P, x, q); }
    const float u1 = x.c[1], u2 = x.c[2];
#pragma unroll
    for(int q = 0; q < NV; ++q) dt[q] = -(u1 * d1[q] + u2 * d2[q]);

    const float div_u = d1[1] + d2[2];
    dt[0] -= div_u;
    dt[1] -= cs2 * d1[0];
    dt[2] -= cs2 * d2[0];

    const float s11 = d1[1] - div_u * THIRD;
    const float s12 = (d1[2] + d2[1]) * 0.5f;
    const float s22 = d2[2] - div_u * THIRD;
    const float two_nus = 2.0f * nus;
    dt[1] += two_nus * (s11 * d1[0] + s12 * d2[0]);
    dt[2] += two_nus * (s12 * d1[0] + s22 * d2[0]);

    const float d11_u1 = D11(P, x, 1), d11_u2 = D11(P, x, 2);
    const float d12_u1 = D12(P, x, 0), d12_u2 = D12(P, x, 1);
    const float d22_u1 = D22(P, x, 1), d22_u2 = D22(P, x, 2);
    const float mixed  = nus * THIRD + nub;
    dt[1] += nus * (d11_u1 + d22_u1) + mixed * (d11_u1 + d12_u2);
    dt[2] += nus * (d11_u2 + d22_u2) + mixed * (d12_u1 + d22_u2);

    if(dd != 0.0f)                         // dd = 0 (the default) adds 0
      dt[0] += dd * (LAP(P, x, 0) + d1[0] * d1[0] + d2[0] * d2[0]);

    if(ff * kk < 0.0f) dt[2] += ff * cosf(kk * (float)i * P.Delta1);     // Kolmogorov forcing
  }

  // Stochastic forcing, applied when f k > 0 (src/hydro-isot/Euler.cuh:21-37).
  static constexpr bool HAS_FIRST_ORDER = true;
  static __device__ __forceinline__ void first_order(const Params &P, float *u, float dt, int i, int j)
  {
    const float ff = P.par[4], kk = P.par[5];
    if(!(ff * kk > 0.0f)) return;
    const float m1 = kk * cosf(P.rand1), m2 = kk * sinf(P.rand1);
    const int k1 = (int)(m1 + 0.5f), k2 = (int)(m2 + 0.5f);
    const float phi = 6.2831853071795865f *
      ((float)(k1 * i) * P.Delta1 + (float)(k2 * j) * P.Delta2);
    const float amp = sqrtf(dt) * (ff * cosf(phi + P.rand2)) / (kk * expf(u[0]));
    u[1] += amp * m1;
    u[2] += amp * m2;
  }

  static __device__ __forceinline__ float lam_adv(const Params &P, const float *c, int, int,
                                                  float k1a, float k1b)
  {
    const float cs = sqrtf(P.par[3]);
    return (fabsf(c[1]) + cs) * k1a + (fabsf(c[2]) + cs) * k1b;
  }
  static __device__ __forceinline__ float lam_diff(const Params &, const float *, int, int) { return 0.0f; }
  static constexpr bool CELL_DIFF = false;
};

// ---------------------------------------------------------------------------
// disk: (lnd, ur, uz, Omg, lne, lner) on the (log r, theta) plane;
// src/disk/eqns.cuh:19-247, Euler.cuh:19-69.
// par: M rS rin gamma nus nub kappa alpha ar es ff dd nu ed rd eps (eps: an
// added gravity softening length, 0 = the reference's force law)
// ---------------------------------------------------------------------------
struct Disk {
  static constexpr int NV = 6, ODD = 1, ND = 4, FLOP = 733, ID = DISK;
  static constexpr int D12V[ND] = {0, 1, 2, 5};
  using P_t = Prim<NV, ND>;

  // Row quantities of the (log r, theta) grid, from the run's table (one
  // scalar load: the row index is uniform across a wavefront).
  static __device__ __forceinline__ float4 row(const Params &P, int i)
  { return reinterpret_cast<const float4 *>(P.rowtab)[i]; }

  static __device__ __forceinline__ void rhs(const Params &P, const P_t &x, int i, int j, float dt[NV])
  {
    enum { LND, UR, UZ, OMG, LNE, LNER };
    const float M = P.par[0], g = P.par[3];
    const float nus = P.par[4], nub = P.par[5], kap = P.par[6], alpha = P.par[7];
    const float ar = P.par[8], es = P.par[9], ff = P.par[10];
    const float pdd = P.par[11], pnu = P.par[12], ped = P.par[13], prd = P.par[14];

    // sin and cos of the column and 1/sin depend on j only (hoisted out of a
    // row march); sphr, 1/sphr, sphr^2 and the gravity come from the table.
    // Reciprocals are v_rcp_f32 (1 ulp) instead of IEEE divisions.
    float sint, cost;
    sincosf(((float)j + 0.5f) * P.Delta2, &sint, &cost);
    const float4 rw = row(P, i);
    const float sphr = rw.x, isphr = rw.y, sphr_2 = rw.z, gr = rw.w;
    const float r = sphr * sint;
    const float ir = isphr * __builtin_amdgcn_rcpf(sint);

    const float den = expf(x.c[LND]), eng = expf(x.c[LNE]), rad = expf(x.c[LNER]);
    const float ieng = __builtin_amdgcn_rcpf(eng);
    const float gamma1 = g - 1.0f;
    const float temp = gamma1 * eng;
    const float nuSS = alpha * g * temp * r * __builtin_amdgcn_sqrtf(r * __builtin_amdgcn_rcpf(M));

    float dr[NV], dz[NV], lap[NV], d11[NV], d22[NV];
#pragma unroll
    for(int q = 0; q < NV; ++q) {
      const float a = D1(P, x, q), b = D2(P, x, q);
      dr[q] = (sint * a + cost * b) * isphr;
      dz[q] = (cost * a - sint * b) * isphr;
      d11[q] = D11(P, x, q);
      d22[q] = D22(P, x, q);
    }
    const float d12_lnd = D12(P, x, 0), d12_ur = D12(P, x, 1);
    const float d12_uz = D12(P, x, 2), d12_lner = D12(P, x, 3);
    const float cr = cost * isphr, sr = sint * isphr;
    const float cc = cr * cr, cs = cr * sr, ss = sr * sr;
    const float c2 = cc - ss, s2 = 2.0f * cs;
    const float isphr_2 = isphr * isphr;
#pragma unroll
    for(int q = 0; q < NV; ++q) lap[q] = (d11[q] + d22[q]) * isphr_2;

    const float drr_lnd  = ss * d11[LND]  + cc * d22[LND]  + s2 * d12_lnd  + cr * dz[LND]  - sr * dr[LND];
    const float drr_ur   = ss * d11[UR]   + cc * d22[UR]   + s2 * d12_ur   + cr * dz[UR]   - sr * dr[UR];
    const float drr_lner = ss * d11[LNER] + cc * d22[LNER] + s2 * d12_lner + cr * dz[LNER] - sr * dr[LNER];
    const float drz_lnd  = cs * (d11[LND]  - d22[LND])  + c2 * d12_lnd  - cr * dr[LND]  - sr * dz[LND];
    const float drz_ur   = cs * (d11[UR]   - d22[UR])   + c2 * d12_ur   - cr * dr[UR]   - sr * dz[UR];
    const float drz_uz   = cs * (d11[UZ]   - d22[UZ])   + c2 * d12_uz   - cr * dr[UZ]   - sr * dz[UZ];
    const float drz_lner = cs * (d11[LNER] - d22[LNER]) + c2 * d12_lner - cr * dr[LNER] - sr * dz[LNER];
    const float dzz_lnd  = cc * d11[LND]  + ss * d22[LND]  - s2 * d12_lnd  - cr * dz[LND]  + sr * dr[LND];
    const float dzz_uz   = cc * d11[UZ]   + ss * d22[UZ]   - s2 * d12_uz   - cr * dz[UZ]   + sr * dr[UZ];
    const float dzz_lner = cc * d11[LNER] + ss * d22[LNER] - s2 * d12_lner - cr * dz[LNER] + sr * dr[LNER];

    const float ur = x.c[UR], uz = x.c[UZ], Omg = x.c[OMG];
    dt[LND]  = -(ur * dr[LND]  + uz * dz[LND]);                          // advection, pseudo-force
    dt[UR]   = -(ur * dr[UR]   + uz * dz[UR] - Omg * Omg * r);
    dt[UZ]   = -(ur * dr[UZ]   + uz * dz[UZ]);
    dt[OMG]  = -(ur * dr[OMG]  + uz * dz[OMG] + 2.0f * ur * Omg * ir);
    dt[LNE]  = -(ur * dr[LNE]  + uz * dz[LNE]);
    dt[LNER] = -(ur * dr[LNER] + uz * dz[LNER]);

    {                                                                     // pressure, viscous heating
      const float ur_r = ur * ir;
      const float div_u = dr[UR] + dz[UZ] + ur_r;
      dt[LND] -= div_u;
      dt[UR]  -= temp * (dr[LND] + dr[LNE]);
      dt[UZ]  -= temp * (dz[LND] + dz[LNE]);
      dt[LNE] -= div_u * gamma1;
      const float Srr = dr[UR] - div_u * THIRD;
      const float Szz = dz[UZ] - div_u * THIRD;
      const float Spp = ur_r - div_u * THIRD;
      const float Srz = (dz[UR] + dr[UZ]) * 0.5f;
      const float Szp = r * dz[OMG] * 0.5f;
      const float Spr = r * dr[OMG] * 0.5f;
      const float two_nus = 2.0f * (nus + nuSS);
      dt[UR]  += two_nus * (Srr * dr[LND] + Srz * dz[LND]);
      dt[UZ]  += two_nus * (Srz * dr[LND] + Szz * dz[LND]);
      dt[OMG] += two_nus * (Spr * dr[LND] + Szp * dz[LND]) * ir;
      dt[LNE] += (two_nus * (Srr * Srr + Szz * Szz + Spp * Spp +
                  2.0f * (Srz * Srz + Szp * Szp + Spr * Spr)) +
                 nub * div_u * div_u) * ieng;
    }
    {                                                                     // velocity diffusion
      const float tmp1 = (nus + nuSS) + pnu * sphr_2;
      const float tmp2 = (nus + nuSS) * ir;
      dt[UR]  += tmp1 * lap[UR]  + tmp2 * (dr[UR] - ur * ir);
      dt[UZ]  += tmp1 * lap[UZ]  + tmp2 * dr[UZ];
      dt[OMG] += tmp1 * lap[OMG] + tmp2 * (dr[OMG] * 3.0f);
      const float mixed = (nus + nuSS) * THIRD + nub;
      dt[UR] += mixed * (drr_ur + drz_uz + (dr[UR] - ur * ir) * ir);
      dt[UZ] += mixed * (drz_ur + dzz_uz + dz[UR] * ir);
    }
    {                                                                     // diffusion, radiation
      const float lap_d_d = lap[LND] + dr[LND] * dr[LND] + dz[LND] * dz[LND];
      const float lap_e_e = lap[LNE] + dr[LNE] * dr[LNE] + dz[LNE] * dz[LNE];
      dt[LND] += lap_d_d * pdd * sphr_2;
      dt[LNE] += lap_e_e * (ped * sphr_2 + kap * gamma1 + g * nuSS);

      // free-free opacity and cooling vanish with ff = 0 (the default): the
      // uniform branch skips powf
      float kappa = 0.0f, cooling = 0.0f;
      if(ff != 0.0f) {
        kappa = ff * den * den * powf(temp, -3.5f);
        const float temp2 = temp * temp;
        cooling = kappa * (ar * temp2 * temp2 * __builtin_amdgcn_rcpf(den) - rad);
      }
      const float chi = es * den + kappa;
      const float ichi = __builtin_amdgcn_rcpf(chi);
      const float d_lner_2 = dr[LNER] * dr[LNER] + dz[LNER] * dz[LNER];
      const float bigR = __builtin_amdgcn_sqrtf(d_lner_2) * ichi;
      const float denom = 6.0f + (3.0f + bigR) * bigR;
      const float idenom = __builtin_amdgcn_rcpf(denom);
      const float Lambda = (2.0f + bigR) * idenom;
      dt[UR] -= Lambda * rad * (dr[LND] + dr[LNE]);
      dt[UZ] -= Lambda * rad * (dz[LND] + dz[LNE]);

      dt[LNE]  -= cooling * ieng;
      dt[LNER] += cooling * __builtin_amdgcn_rcpf(rad);

      const float d_lnd_d_lner = dr[LND] * dr[LNER] + dz[LND] * dz[LNER];
      const float lap_er_er = lap[LNER] + d_lner_2;
      const float nu_rad = Lambda * ichi;
      const float Lambdap = -bigR * (4.0f + bigR) * (idenom * idenom);
      const float dr_lnEr = dr[LND] + dr[LNER];
      const float dz_lnEr = dz[LND] + dz[LNER];
      const float drr_lnEr = drr_lnd + drr_lner;
      const float drz_lnEr = drz_lnd + drz_lner;
      const float dzz_lnEr = dzz_lnd + dzz_lner;
      const float dr_chi = chi * dr[LND] - 3.5f * kappa * dr[LNE];
      const float dz_chi = chi * dz[LND] - 3.5f * kappa * dz[LNE];
      const float ibr = __builtin_amdgcn_rcpf(1e-16f + bigR);
      const float dr_bigR = ((dr_lnEr * drr_lnEr + dz_lnEr * drz_lnEr) * ibr * ichi - bigR * dr_chi) * ichi;
      const float dz_bigR = ((dr_lnEr * drz_lnEr + dz_lnEr * dzz_lnEr) * ibr * ichi - bigR * dz_chi) * ichi;
      dt[LNER] += lap_er_er * prd * sphr_2 +
        nu_rad * (lap_d_d + 2.0f * d_lnd_d_lner + lap_er_er) +
        dr_lnEr * (Lambdap * dr_bigR - nu_rad * dr_chi) * ichi +
        dz_lnEr * (Lambdap * dz_bigR - nu_rad * dz_chi) * ichi;
    }
    dt[UR] -= sint * gr;                                                  // gravity
    dt[UZ] -= cost * gr;
  }

  // Radiation-equilibrium pre-step (src/disk/Euler.cuh:19-69).
  static constexpr bool HAS_FIRST_ORDER = true;
  static __device__ __forceinline__ void first_order(const Params &P, float *u, float dt, int, int)
  {
    const float g = P.par[3], ar = P.par[8], ff = P.par[10];
    const float den = expf(u[0]), eng = expf(u[4]), rad = expf(u[5]);
    const float gamma1 = g - 1.0f;
    const float g2 = gamma1 * gamma1;
    const float scale = powf(ar * g2 * g2 / den, THIRD);
    const float x0 = (eng + rad) * scale;
    const float a = powf(sqrtf(0.00390625f + x0 * x0 * x0 / 27.0f) + 0.0625f, THIRD);
    const float b = a - x0 / (3.0f * a);
    const float c = sqrtf(2.0f * b);
    const float eq_eng = (sqrtf(0.5f / c - 0.5f * b) - 0.5f * c) / scale;
    const float dt_eng = (eq_eng - eng) / dt;
    const float temp = gamma1 * eng;
    const float temp2 = temp * temp;
    const float kappa = ff * den * den * powf(temp, -3.5f);
    const float cooling = kappa * (ar * temp2 * temp2 / den - rad);
    if(cooling * cooling > dt_eng * dt_eng) {
      u[4] = logf(eq_eng);
      u[5] = logf(eng + rad - eq_eng);
    }
  }

  // Advective rate with the (log r, theta) metric: (|ur| + |uz| + cs)(k/D1 + k/D2)/sphr,
  // cs^2 = gamma (gamma-1) e + er/3.
  static __device__ __forceinline__ float lam_adv(const Params &P, const float *c, int i, int,
                                                  float k1a, float k1b)
  {
    const float g = P.par[3];
    const float temp = (g - 1.0f) * expf(c[4]);
    const float cs = sqrtf(g * temp + expf(c[5]) * THIRD);
    return (fabsf(c[1]) + fabsf(c[2]) + cs) * (k1a + k1b) * row(P, i).y;
  }
  // Largest local diffusivity in index-space units (coefficient / sphr^2).
  static constexpr bool CELL_DIFF = true;
  static __device__ __forceinline__ float lam_diff(const Params &P, const float *c, int i, int j)
  {
    const float M = P.par[0], g = P.par[3];
    const float4 rw = row(P, i);
    const float r = rw.x * sinf(((float)j + 0.5f) * P.Delta2);
    const float s2 = rw.z;
    const float temp = (g - 1.0f) * expf(c[4]);
    const float nuSS = P.par[7] * g * temp * r * sqrtf(r / M);
    const float nuv = P.par[4] + nuSS;
    float dmax = nuv * (4.0f / 3.0f) + P.par[5] + P.par[12] * s2;
    dmax = fmaxf(dmax, P.par[11] * s2);
    dmax = fmaxf(dmax, P.par[13] * s2 + P.par[6] * (g - 1.0f) + g * nuSS);
    dmax = fmaxf(dmax, P.par[14] * s2 + 1.0f / (3.0f * P.par[9] * expf(c[0])));
    return dmax / s2;
  }
};

} // namespace fg2
