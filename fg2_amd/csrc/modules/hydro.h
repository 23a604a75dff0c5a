// modules/hydro.h -- ideal-gas hydrodynamics, state (lnd, u1, u2, lne):
// the reference's src/hydro/ (state.h, eqns.cuh:19-92, bcond.cuh, config.h,
// init.h:26-112) as one module of this library (the interface: physics.h).
#pragma once
#include "../stencil.h"

namespace fg2 {

struct Hydro {
  static constexpr const char *NAME = "hydro";
  static constexpr const char *FIELDS = "lnd,u1,u2,lne";   // the state variables (state.h)
  static constexpr int NV = 4, ODD = 2, ND = 2, FLOP = 289;
  static constexpr int D12V[ND] = {1, 2};
  static constexpr bool ONE_PASS = true, ROW_TABLE = false, HAS_FIRST_ORDER = false,
                        RANDOM_PHASES = false, CELL_DIFF = false;
  // the right-hand side takes the stencil primitives whole (rhs)
  static constexpr bool DERIV_FIRST = false;
  // the one-pass step syncs its workgroup once per row (at two wavefronts
  // per SIMD, as with one barrier per stage, and 8% faster)
  static constexpr bool STEP_ONE_BARRIER = true;
  using P_t = Prim<NV, ND>;

  // par: gamma nus nub kappa dd (src/hydro/eqns.cuh:19-24)
  static constexpr ParamDef PARAMS[] = {
    {"gamma", 5.0 / 3.0}, {"nus", 2.0e-4}, {"nub", 0.0}, {"kappa", 5.0e-4}, {"dd", 0.0}};

  static void settings(Geometry &, int, int) {}
  static __host__ __device__ bool first_order_on(const float *) { return false; }
  // D_max of the CFL rule: max(4 nus / 3 + nub, (gamma - 1) kappa, dd)
  static double diffusivity(const double *p)
  {
    const double a = p[1] * 4.0 / 3.0 + p[2], b = (p[0] - 1.0) * p[3];
    return a > b ? (a > p[4] ? a : p[4]) : (b > p[4] ? b : p[4]);
  }

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

  // FP64 sums of rho, rho u1, rho u2 at one cell (fg2_sums)
  static __device__ __forceinline__ void sums(const Params &, const float *c, int, int, double acc[4])
  {
    const double rho = exp((double)c[0]);
    acc[0] += rho;
    acc[1] += rho * c[1];
    acc[2] += rho * c[2];
  }

  // init.h pick(): the named conditions (0 = zeros)
  static int ic_of(const char *n)
  {
    const char *names[] = {"blast", "bow", "implos", "KH", "Sod"};
    for(int k = 0; k < 5; ++k) {
      const char *a = names[k], *b = n;
      while(*a && *a == *b) { ++a; ++b; }
      if(!*a && !*b) return k + 1;
    }
    return 0;
  }
  // src/hydro/init.h:26-112 at cell (i, j), centre (x, y), in FP64 with every
  // operation rounded on its own
  static __device__ void ic(const IcArgs &A, int i, int j, float x, float y, float *c)
  {
#pragma clang fp contract(off)
    const double g = A.par[0];
    // the centre shift is stored back into the R-typed argument (init.h:30-31)
    const double xx = (float)((double)x - 0.5 * A.l1), yy = (float)((double)y - 0.5 * A.l2);
    double d = 1.0, u = 0.0, e = 1.0;
    switch(A.ic) {
    case 1:                                      // blast
      e = ((xx * xx + yy * yy < 0.01) ? 1.0 : 0.01) / (g - 1.0);
      break;
    case 2: {                                    // bow
      const double r = sqrt(xx * xx + yy * yy);
      d = 9.9 * fermi_dirac(r - 0.01, 0.001) + 0.1;
      u = (10.0 / d - 1.0) / 99.0;
      e = 0.01 / d / (g - 1.0);
      break;
    }
    case 3: {                                    // implos
      const double f = fermi_dirac(0.15 - fabs(xx) - fabs(yy), 0.001);
      d = 0.125 + 0.875 * f;
      e = (0.14 + 0.86 * f) / d / (g - 1.0);
      break;
    }
    case 4: {                                    // KH
      const double r = ic_draw(A.seed, (uint64_t)i * A.n2 + j);
      if(fabs(yy) < 0.25) { d = 2.0; u = 0.5 + 0.01 * r - 0.005; }
      else                { d = 1.0; u = -0.5 + 0.01 * r - 0.005; }
      e = 2.5 / d / (g - 1.0);
      break;
    }
    case 5: {                                    // Sod
      double Pr;
      if(fmod(xx - 2.0 * yy + 1.5, 1.0) < 0.5) { d = 1.0; Pr = 1.0; }
      else                                     { d = 0.125; Pr = 0.1; }
      e = Pr / d / (g - 1.0);
      break;
    }
    default:                                     // zeros
      c[0] = c[1] = c[2] = c[3] = 0.0f;
      return;
    }
    c[0] = (float)log(d); c[1] = (float)u; c[2] = 0.0f; c[3] = (float)log(e);
  }
};

} // namespace fg2
