// modules/disk.h -- an accretion disk on the (log r, theta) plane, state
// (lnd, ur, uz, Omg, lne, lner) with flux-limited radiative diffusion: the
// reference's src/disk/ (state.h, eqns.cuh:19-247, Euler.cuh:19-69,
// bcond.cuh, config.h:21-34, init.h:25-68) as one module of this library
// (the interface: physics.h).
#pragma once
#include <cmath>
#include "../stencil.h"

namespace fg2 {

struct Disk {
  static constexpr const char *NAME = "disk";
  static constexpr const char *FIELDS = "lnd,ur,uz,Omg,lne,lner";   // the state variables (state.h)
  static constexpr int NV = 6, ODD = 1, ND = 4, FLOP = 733;
  static constexpr int D12V[ND] = {0, 1, 2, 5};
  // the one-pass step's register rings would hold 3 x 8 x 6 floats per lane
  // and spill, so disk runs step by substages
  static constexpr bool ONE_PASS = false, ROW_TABLE = true, HAS_FIRST_ORDER = true,
                        RANDOM_PHASES = false, CELL_DIFF = true;
  using P_t = Prim<NV, ND>;

  // par: M rS rin gamma nus nub kappa alpha ar es ff dd nu ed rd
  // (src/disk/eqns.cuh:19-36), and eps, a gravity softening length that the
  // reference does not have (0 keeps its force law)
  static constexpr ParamDef PARAMS[] = {
    {"M", 1.0}, {"rS", 2.0}, {"rin", 3.0}, {"gamma", 5.0 / 3.0}, {"nus", 0.0}, {"nub", 0.0},
    {"kappa", 0.0}, {"alpha", 0.01}, {"ar", 1.0e8}, {"es", 1.0}, {"ff", 0.0}, {"dd", 1.0e-4},
    {"nu", 2.0e-4}, {"ed", 5.0e-4}, {"rd", 5.0e-4}, {"eps", 0.0}};

  // config.h settings() (src/disk/config.h:21-34): theta spans [0, pi] in R,
  // log r takes cells of the same aspect at the equator, Neumann rows and
  // reflecting columns
  static void settings(Geometry &g, int n1, int n2)
  {
    g.l2 = (float)M_PI;
    const float hd2 = (float)(0.5 * g.l2 / n2);
    const float hd1 = (float)std::log(hd2 + std::sqrt(hd2 * (double)hd2 + 1.0));
    g.l1 = (float)(2.0 * hd1 * n1);
    g.p1 = 0; g.p2 = 0;
  }
  // the radiation-equilibrium pre-step runs when ff != 0 (Euler.cuh:19-69)
  static __host__ __device__ bool first_order_on(const float *par) { return par[10] != 0.0f; }
  // the diffusive rate is per cell (lam_diff)
  static double diffusivity(const double *) { return -1.0; }

  // Row table entry of row i (Params::rowtab): the row's spherical radius
  // (src/disk/eqns.cuh:52-54), its inverse and square, and the gravity of
  // the point mass (src/disk/eqns.cuh:238-239).  A softening length eps > 0
  // replaces M / d^2 by the Plummer force M d / (d^2 + eps^2)^(3/2), d =
  // sphr - rS.
  static __device__ __forceinline__ float4 row_entry(const Params &P, int i)
  {
    const float M = P.par[0], rS = P.par[1], rin = P.par[2], eps = P.par[15];
    const float sphr = rin * expf(((float)i + 0.5f) * P.Delta1);
    const float tmp = sphr - rS;
    float gr;
    if(eps == 0.0f) {
      gr = M / (tmp * tmp);
    } else {
      const float d2 = tmp * tmp + eps * eps;
      gr = M * tmp / (d2 * sqrtf(d2));
    }
    return make_float4(sphr, 1.0f / sphr, sphr * sphr, gr);
  }

  // Row quantities of the (log r, theta) grid, from the run's table (one
  // scalar load: the row index is uniform across a wavefront).
  static __device__ __forceinline__ float4 row(const Params &P, int i)
  {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 r = ((__attribute__((address_space(4))) const v4 *)P.rowtab)[i];
    return make_float4(r.x, r.y, r.z, r.w);
  }

  // The right-hand side in two parts, so that a kernel can take each
  // variable's derivatives as soon as its stencil is read (deriv_var) and
  // keep only those, not the 90 stencil primitives, until rhs_d runs: D1
  // and D2 of every variable, D11 and D22 of all but Omg and lne, whose
  // viscous terms need only their Laplacian, and D12 of the D12V variables.
  static constexpr bool DERIV_FIRST = true;
  static constexpr bool STEP_ONE_BARRIER = false;   // no one-pass step (ONE_PASS)
  enum { LND, UR, UZ, OMG, LNE, LNER };
  struct Deriv { float a[NV], b[NV], d11[NV], d22[NV], lap[NV], d12[ND]; };
  static __device__ __forceinline__ void deriv_var(const Params &P, const P_t &x, int q, Deriv &D)
  {
    D.a[q] = D1(P, x, q);
    D.b[q] = D2(P, x, q);
    if(q == OMG || q == LNE) {
      D.lap[q] = LAP(P, x, q);
    } else {
      D.d11[q] = D11(P, x, q);
      D.d22[q] = D22(P, x, q);
    }
  }
  static __device__ __forceinline__ void deriv_mixed(const Params &P, const P_t &x, Deriv &D)
  {
#pragma unroll
    for(int d = 0; d < ND; ++d) D.d12[d] = D12(P, x, d);
  }

  static __device__ __forceinline__ void rhs(const Params &P, const P_t &x, int i, int j, float dt[NV])
  {
    Deriv D;
#pragma unroll
    for(int q = 0; q < NV; ++q) deriv_var(P, x, q, D);
    deriv_mixed(P, x, D);
    rhs_d(P, x.c, D, i, j, dt);
  }

  // The right-hand side from the centre values xc and the derivatives D.
  static __device__ __forceinline__ void rhs_d(const Params &P, const float (&xc)[NV], const Deriv &D,
                                               int i, int j, float dt[NV])
  {
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

    const float den = expf(xc[LND]), eng = expf(xc[LNE]), rad = expf(xc[LNER]);
    const float ieng = __builtin_amdgcn_rcpf(eng);
    const float gamma1 = g - 1.0f;
    const float temp = gamma1 * eng;
    const float nuSS = alpha * g * temp * r * __builtin_amdgcn_sqrtf(r * __builtin_amdgcn_rcpf(M));

    // d/dr = (sin d/dlnr + cos d/dtheta) / sphr, d/dz = (cos d/dlnr - sin
    // d/dtheta) / sphr, with 1/sphr folded into the weights
    const float cr = cost * isphr, sr = sint * isphr;
    float dr[NV], dz[NV], lap[NV], d11[NV], d22[NV];
#pragma unroll
    for(int q = 0; q < NV; ++q) {
      const float a = D.a[q], b = D.b[q];
      dr[q] = sr * a + cr * b;
      dz[q] = cr * a - sr * b;
      if(q != OMG && q != LNE) {          // Omg and lne need only their Laplacian
        d11[q] = D.d11[q];
        d22[q] = D.d22[q];
      }
    }
    const float d12_lnd = D.d12[0], d12_ur = D.d12[1];
    const float d12_uz = D.d12[2], d12_lner = D.d12[3];
    const float cc = cr * cr, cs = cr * sr, ss = sr * sr;
    const float c2 = cc - ss, s2 = 2.0f * cs;
    const float isphr_2 = isphr * isphr;
#pragma unroll
    for(int q = 0; q < NV; ++q)
      lap[q] = (q == OMG || q == LNE ? D.lap[q] : d11[q] + d22[q]) * isphr_2;

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

    const float ur = xc[UR], uz = xc[UZ], Omg = xc[OMG];
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
  // FP64 sums of rho, rho ur, rho uz and rho Omg R^2 at one cell, with the
  // volume element of the (log r, theta) grid, r^3 sin(theta) dlnr dtheta, in
  // index space (fg2_sums)
  static __device__ __forceinline__ void sums(const Params &P, const float *c, int i, int j,
                                              double acc[4])
  {
    const double rho = exp((double)c[0]);
    const double sphr = (double)P.par[2] * exp((i + 0.5) * (double)P.Delta1);
    const double th = (j + 0.5) * (double)P.Delta2;
    const double R = sphr * sin(th);
    const double dV = sphr * sphr * sphr * sin(th);
    acc[0] += rho * dV;
    acc[1] += rho * c[1] * dV;
    acc[2] += rho * c[2] * dV;
    acc[3] += rho * c[3] * R * R * dV;
  }

  // init.h pick(): the torus, whatever the name
  static int ic_of(const char *) { return 1; }
  // src/disk/init.h:25-68: hydrostatic torus of Hawley (2000) at log r = x,
  // theta = y
  static __device__ void ic(const IcArgs &A, int, int, float lnr, float theta, float *c)
  {
#pragma clang fp contract(off)
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
};

} // namespace fg2
