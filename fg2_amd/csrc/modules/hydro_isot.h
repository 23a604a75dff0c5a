// modules/hydro_isot.h -- isothermal hydrodynamics, state (lnd, u1, u2):
// the reference's src/hydro-isot/ (state.h, eqns.cuh:19-87, Euler.cuh:21-44,
// bcond.cuh, config.h:21-26, init.h:21-49) as one module of this library
// (the interface: physics.h).
#pragma once
#include "../stencil.h"

namespace fg2 {

struct HydroIsot {
  static constexpr const char *NAME = "hydro-isot";
  static constexpr const char *FIELDS = "lnd,u1,u2";   // the state variables (state.h)
  static constexpr int NV = 3, ODD = 2, ND = 2, FLOP = 214;
  static constexpr int D12V[ND] = {1, 2};
  static constexpr bool ONE_PASS = true, ROW_TABLE = false, HAS_FIRST_ORDER = true,
                        RANDOM_PHASES = true, CELL_DIFF = false;
  // the right-hand side takes the stencil primitives whole (rhs)
  static constexpr bool DERIV_FIRST = false;
  // the periodic one-pass step syncs its workgroup once per stage: once per
  // row would cost it its third wavefront per SIMD and 17%
  static constexpr bool STEP_ONE_BARRIER = false;
  using P_t = Prim<NV, ND>;

  // par: dd nus nub cs2 f k (src/hydro-isot/eqns.cuh:19-25)
  static constexpr ParamDef PARAMS[] = {
    {"dd", 0.0}, {"nus", 2.0e-4}, {"nub", 0.0}, {"cs2", 1.0}, {"f", 0.0}, {"k", -8.0}};

  // config.h settings(): a 2 pi periodic box, in R
  static void settings(Geometry &g, int, int) { g.l1 = g.l2 = (float)(2.0 * M_PI); }
  // the stochastic forcing runs when f k > 0 (Euler.cuh:21-37); f k < 0 is
  // the Kolmogorov forcing inside eqns
  static __host__ __device__ bool first_order_on(const float *par) { return par[4] * par[5] > 0.0f; }
  // D_max of the CFL rule: max(4 nus / 3 + nub, dd)
  static double diffusivity(const double *p)
  {
    const double a = p[1] * 4.0 / 3.0 + p[2];
    return a > p[0] ? a : p[0];
  }

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
  static __device__ __forceinline__ void first_order(const Params &P, float *u, float dt, int i, int j)
  {
    const float ff = P.par[4], kk = P.par[5];
    if(!first_order_on(P.par)) return;
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

  static __device__ __forceinline__ void sums(const Params &, const float *c, int, int, double acc[4])
  {
    const double rho = exp((double)c[0]);
    acc[0] += rho;
    acc[1] += rho * c[1];
    acc[2] += rho * c[2];
  }

  // init.h pick(): "Kf" (1), anything else zeros
  static int ic_of(const char *n) { return (n[0] == 'K' && n[1] == 'f' && !n[2]) ? 1 : 0; }
  // src/hydro-isot/init.h:21-49: Kolmogorov flow with Box-Muller noise
  static __device__ void ic(const IcArgs &A, int i, int j, float x, float, float *c)
  {
#pragma clang fp contract(off)
    if(A.ic != 1) { c[0] = c[1] = c[2] = 0.0f; return; }
    const double nus = A.par[1], fi = A.par[4], ki = A.par[5];
    const uint64_t q = 2 * ((uint64_t)i * A.n2 + j);
    const double r1 = ic_draw(A.seed, q), r2 = ic_draw(A.seed, q + 1);
    const double rad = sqrt(-2.0 * log(r1));
    const double q1 = rad * sin(2.0 * M_PI * r2);
    const double q2 = rad * cos(2.0 * M_PI * r2);
    const double a = fi / (nus * ki * ki);
    c[0] = 0.0f;
    c[1] = (float)(1.0e-9 * q1);
    c[2] = (float)(1.0e-9 * q2 + a * cos(ki * (double)x));
  }
};

} // namespace fg2
