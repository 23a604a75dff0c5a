// oracle.cpp -- host restatement of fg2's time-step path.
//
// TEST INFRASTRUCTURE ONLY.  This file is what GPU results are compared with: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// library built from it.  The product (fg2_amd/) never links or calls it.
//
// It restates, in plain C++ templated on the storage type R (double = the
// FP64 oracle, float = the FP32 CPU baseline), the reference's algorithm:
//   - ghosted AoS storage, HALF = 3 ghost cells per side   (src/setup.cpp:60-78)
//   - boundary fill, dim 1 first then dim 2               (src/bcond.cu:22-86)
//   - 6th-order operators D1 D2 D11 D22 D12               (src/deriv.h:41-63)
//   - module right-hand sides                             (src/hydro/eqns.cuh:26-92,
//                                                          src/hydro-isot/eqns.cuh:27-87,
//                                                          src/disk/eqns.cuh:38-247)
//   - kick: v <- F(u) + beta v                            (src/kick.cu:41-57,98)
//   - drift: u <- u + dt v, NaN flag                      (src/drift.cu:23-32)
//   - point-wise first_order pre-step                     (src/kick.cu:116-138,
//                                                          src/*/Euler.cuh)
//   - low-storage RK3 step                                (src/step.cpp:22-43)
//   - closed-form time step                               (src/getdt.cpp:22-29)
//   - Delta = l/n, module settings()                      (src/config.cpp:24-41,
//                                                          src/*/config.h)
//   - initial conditions at the cell centres              (src/init.cpp:26-42,
//                                                          src/*/init.h)
// and a CFL time step with advective and diffusive limits, which the
// reference only has as a TODO (src/getdt.cpp:26-28).
//
// Pinning status of each function is recorded in the oracle's Python header
// (oracle/oracle.py) and in the tests that check it.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <limits>

namespace {

constexpr int HALF = 3;
enum Module { HYDRO = 0, HYDRO_ISOT = 1, DISK = 2 };

int nvar_of(int m) { return m == HYDRO ? 4 : m == HYDRO_ISOT ? 3 : 6; }

// ---------------------------------------------------------------------------
// Parameters (the module __constant__ para_* symbols and their defaults)
// src/hydro/eqns.cuh:19-24, src/hydro-isot/eqns.cuh:19-25, src/disk/eqns.cuh:19-36
// ---------------------------------------------------------------------------
struct Param { const char *name; double value; };

std::vector<Param> default_params(int m)
{
  if(m == HYDRO)
    return {{"gamma", 5.0 / 3.0}, {"nus", 2.0e-4}, {"nub", 0.0},
            {"kappa", 5.0e-4}, {"dd", 0.0}};
  if(m == HYDRO_ISOT)
    return {{"dd", 0.0}, {"nus", 2.0e-4}, {"nub", 0.0}, {"cs2", 1.0},
            {"f", 0.0}, {"k", -8.0}};
  return {{"M", 1.0}, {"rS", 2.0}, {"rin", 3.0}, {"gamma", 5.0 / 3.0},
          {"nus", 0.0}, {"nub", 0.0}, {"kappa", 0.0}, {"alpha", 0.01},
          {"ar", 1.0e8}, {"es", 1.0}, {"ff", 0.0}, {"dd", 1.0e-4},
          {"nu", 2.0e-4}, {"ed", 5.0e-4}, {"rd", 5.0e-4},
          {"eps", 0.0}};   // added softening length; 0 = the reference's gravity
}

// SplitMix64 (Steele, Lea, Flood 2014): the seeded stand-in for the libc
// rand() calls of the hydro-isot RANDOMIZE macro (src/hydro-isot/Euler.cuh:39-44).
uint64_t splitmix64(uint64_t &state)
{
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double unit_double(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

template<class R> struct Sim {
  int module, nv, n1, n2;
  long w;                 // ghosted row length in cells, n2 + 2 HALF
  std::vector<R> u, v;    // (n1 + 2 HALF) x (n2 + 2 HALF) x nv
  std::vector<Param> par;
  double l1 = 1.0, l2 = 1.0, c = 0.5;  // src/global.cpp:28-29
  int p1 = 1, p2 = 1;
  R Delta1 = 0, Delta2 = 0;
  int nan_flag = 0;
  uint64_t rng = 0;
  double rand1 = 0, rand2 = 0;

  Sim(int m, int a, int b) : module(m), nv(nvar_of(m)), n1(a), n2(b),
    w(b + 2 * HALF),
    u((size_t)(a + 2 * HALF) * (b + 2 * HALF) * nvar_of(m),
      -std::numeric_limits<R>::max()),                 // src/setup.cpp:80-83
    v(u.size(), -std::numeric_limits<R>::max()),
    par(default_params(m)) {}

  size_t idx(long i, long j) const
  { return ((size_t)(i + HALF) * w + (j + HALF)) * nv; }
  R *cell(std::vector<R> &a, long i, long j) { return &a[idx(i, j)]; }

  double get(const char *name) const
  {
    for(auto &p : par) if(!strcmp(p.name, name)) return p.value;
    return std::nan("");
  }
  // The reference keeps parameters in R-typed constants, so a parameter is
  // rounded to R before use (src/para.cpp:33-35).  pr[] holds them in the
  // order of default_params(), refreshed by refresh() before each sweep.
  R pr[16] = {};
  void refresh() { for(size_t k = 0; k < par.size() && k < 16; ++k) pr[k] = (R)par[k].value; }
  R P(const char *name) const { return (R)get(name); }
};

// ---------------------------------------------------------------------------
// Stencil access and the 6th-order operators, src/deriv.h:33-63.
// f(a, b) is the value of variable q at cell (i + a, j + b).
// ---------------------------------------------------------------------------
template<class R> struct Stencil {
  const R *c;  long s1, s2;  R D1, D2;
  R f(int q, int a, int b) const { return c[a * s1 + b * s2 + q]; }
  R dif(int q, int a, int b) const { return f(q, a, b) - f(q, -a, -b); }
  R sum(int q, int a, int b) const { return f(q, a, b) + f(q, -a, -b); }
  R mix(int q, int a, int b) const
  { return f(q, a, b) - f(q, -a, b) + f(q, -a, -b) - f(q, a, -b); }

  R d1(int q) const
  { return ((R)(45./60.) * dif(q,1,0) + (R)(-9./60.) * dif(q,2,0)
          + (R)(1./60.) * dif(q,3,0)) / D1; }
  R d2(int q) const
  { return ((R)(45./60.) * dif(q,0,1) + (R)(-9./60.) * dif(q,0,2)
          + (R)(1./60.) * dif(q,0,3)) / D2; }
  R d11(int q) const
  { return ((R)(-490./180.) * f(q,0,0) + (R)(270./180.) * sum(q,1,0)
          + (R)(-27./180.) * sum(q,2,0) + (R)(2./180.) * sum(q,3,0)) / (D1 * D1); }
  R d22(int q) const
  { return ((R)(-490./180.) * f(q,0,0) + (R)(270./180.) * sum(q,0,1)
          + (R)(-27./180.) * sum(q,0,2) + (R)(2./180.) * sum(q,0,3)) / (D2 * D2); }
  R d12(int q) const
  { return ((R)(270./720.) * mix(q,1,1) + (R)(-27./720.) * mix(q,2,2)
          + (R)(2./720.) * mix(q,3,3)) / (D1 * D2); }
};

// ---------------------------------------------------------------------------
// Right-hand sides
// ---------------------------------------------------------------------------

// hydro: variables (lnd, u1, u2, lne); src/hydro/eqns.cuh:26-92
template<class R> void rhs_hydro(const Sim<R> &S, const Stencil<R> &st, R *out)
{
  enum { LND, U1, U2, LNE };
  // gamma nus nub kappa dd
  const R g = S.pr[0], nus = S.pr[1], nub = S.pr[2], kap = S.pr[3], dd = S.pr[4];
  R d1[4], d2[4], dt[4] = {0, 0, 0, 0};
  for(int q = 0; q < 4; ++q) { d1[q] = st.d1(q); d2[q] = st.d2(q); }
  const R u1 = st.f(U1,0,0), u2 = st.f(U2,0,0);

  for(int q = 0; q < 4; ++q) dt[q] -= u1 * d1[q] + u2 * d2[q];   // :37-46

  const R gamma1 = g - (R)1;                                       // :49-56
  const R temp   = gamma1 * std::exp(st.f(LNE,0,0));
  const R div_u  = d1[U1] + d2[U2];
  dt[LND] -= div_u;
  dt[U1]  -= temp * (d1[LND] + d1[LNE]);
  dt[U2]  -= temp * (d2[LND] + d2[LNE]);
  dt[LNE] -= div_u * gamma1;

  const R s11 = d1[U1] - div_u / (R)3;                             // :61-70
  const R s12 = (d1[U2] + d2[U1]) / (R)2;
  const R s22 = d2[U2] - div_u / (R)3;
  const R two_nus = (R)2 * nus;
  dt[U1]  += two_nus * (s11 * d1[LND] + s12 * d2[LND]);
  dt[U2]  += two_nus * (s12 * d1[LND] + s22 * d2[LND]);
  dt[LNE] += (gamma1 / temp) *
    (two_nus * (s11 * s11 + (R)2 * s12 * s12 + s22 * s22) + nub * div_u * div_u);

  const R d11_u1 = st.d11(U1), d11_u2 = st.d11(U2);                // :73-80
  const R d12_u1 = st.d12(U1), d12_u2 = st.d12(U2);
  const R d22_u1 = st.d22(U1), d22_u2 = st.d22(U2);
  const R mixed  = nus / (R)3 + nub;
  dt[U1] += nus * (d11_u1 + d22_u1) + mixed * (d11_u1 + d12_u2);
  dt[U2] += nus * (d11_u2 + d22_u2) + mixed * (d12_u1 + d22_u2);

  const R d_lnd_2 = d1[LND] * d1[LND] + d2[LND] * d2[LND];          // :83-89
  const R d_lne_2 = d1[LNE] * d1[LNE] + d2[LNE] * d2[LNE];
  const R ed      = (g - (R)1) * kap;
  dt[LND] += dd * (st.d11(LND) + st.d22(LND) + d_lnd_2);
  dt[LNE] += ed * (st.d11(LNE) + st.d22(LNE) + d_lne_2);

  for(int q = 0; q < 4; ++q) out[q] = dt[q];
}

// hydro-isot: variables (lnd, u1, u2); src/hydro-isot/eqns.cuh:27-87
template<class R> void rhs_isot(const Sim<R> &S, const Stencil<R> &st,
                                long i, R *out)
{
  enum { LND, U1, U2 };
  // dd nus nub cs2 f k
  const R dd = S.pr[0], nus = S.pr[1], nub = S.pr[2];
  const R cs2 = S.pr[3], ff = S.pr[4], kk = S.pr[5];
  R d1[3], d2[3], dt[3] = {0, 0, 0};
  for(int q = 0; q < 3; ++q) { d1[q] = st.d1(q); d2[q] = st.d2(q); }
  const R u1 = st.f(U1,0,0), u2 = st.f(U2,0,0);

  for(int q = 0; q < 3; ++q) dt[q] -= u1 * d1[q] + u2 * d2[q];   // :37-44

  const R div_u = d1[U1] + d2[U2];                                 // :47-52
  dt[LND] -= div_u;
  dt[U1]  -= cs2 * d1[LND];
  dt[U2]  -= cs2 * d2[LND];

  const R s11 = d1[U1] - div_u / (R)3;                             // :56-63
  const R s12 = (d1[U2] + d2[U1]) / (R)2;
  const R s22 = d2[U2] - div_u / (R)3;
  const R two_nus = (R)2 * nus;
  dt[U1] += two_nus * (s11 * d1[LND] + s12 * d2[LND]);
  dt[U2] += two_nus * (s12 * d1[LND] + s22 * d2[LND]);

  const R d11_u1 = st.d11(U1), d11_u2 = st.d11(U2);                // :67-74
  const R d12_u1 = st.d12(U1), d12_u2 = st.d12(U2);
  const R d22_u1 = st.d22(U1), d22_u2 = st.d22(U2);
  const R mixed  = nus / (R)3 + nub;
  dt[U1] += nus * (d11_u1 + d22_u1) + mixed * (d11_u1 + d12_u2);
  dt[U2] += nus * (d11_u2 + d22_u2) + mixed * (d12_u1 + d22_u2);

  dt[LND] += dd * (st.d11(LND) + d1[LND] * d1[LND] +               // :78-79
                   st.d22(LND) + d2[LND] * d2[LND]);

  if(ff * kk < (R)0) dt[U2] += ff * std::cos(kk * (R)i * S.Delta1); // :82-84

  for(int q = 0; q < 3; ++q) out[q] = dt[q];
}

// Flux limiter of the radiative diffusion, src/disk/eqns.cuh:194-196,216:
// Lambda(R) = (2 + R) / (6 + 3 R + R^2) and its derivative
// Lambda'(R) = -R (4 + R) / (6 + 3 R + R^2)^2, in the reference's order of
// operations.
template<class R> void flux_limiter(R bigR, R &Lambda, R &Lambdap)
{
  const R denom = (R)6 + ((R)3 + bigR) * bigR;
  Lambda = ((R)2 + bigR) / denom;
  Lambdap = -bigR * ((R)4 + bigR) / (denom * denom);
}

// disk: variables (lnd, ur, uz, Omg, lne, lner) on the (log r, theta) plane;
// src/disk/eqns.cuh:38-247
template<class R> void rhs_disk(const Sim<R> &S, const Stencil<R> &st,
                                long i, long j, R *out)
{
  enum { LND, UR, UZ, OMG, LNE, LNER };
  // M rS rin gamma nus nub kappa alpha ar es ff dd nu ed rd
  const R M = S.pr[0], rS = S.pr[1], rin = S.pr[2], g = S.pr[3];
  const R nus = S.pr[4], nub = S.pr[5], kap = S.pr[6];
  const R alpha = S.pr[7], ar = S.pr[8], es = S.pr[9], ff = S.pr[10];
  const R pdd = S.pr[11], pnu = S.pr[12], ped = S.pr[13], prd = S.pr[14];

  const R theta = ((R)j + (R)0.5) * S.Delta2;                       // :48-54
  const R sint = std::sin(theta), cost = std::cos(theta);
  const R sphr = rin * std::exp(((R)i + (R)0.5) * S.Delta1);
  const R r = sphr * sint;
  const R sphr_2 = sphr * sphr;

  const R den = std::exp(st.f(LND,0,0));                            // :59-64
  const R eng = std::exp(st.f(LNE,0,0));
  const R rad = std::exp(st.f(LNER,0,0));
  const R gamma1 = g - (R)1;
  const R temp = gamma1 * eng;
  const R nuSS = alpha * g * temp * r * std::sqrt(r / M);

  R dr[6], dz[6], lap[6];
  for(int q = 0; q < 6; ++q) {                                      // :67-85
    const R a = st.d1(q), b = st.d2(q);
    dr[q] = (sint * a + cost * b) / sphr;
    dz[q] = (cost * a - sint * b) / sphr;
  }

  R d11[6], d22[6];                                                 // :88-118
  for(int q = 0; q < 6; ++q) { d11[q] = st.d11(q); d22[q] = st.d22(q); }
  const R d12_lnd = st.d12(LND), d12_ur = st.d12(UR);
  const R d12_uz = st.d12(UZ), d12_lner = st.d12(LNER);
  const R cr = cost / sphr, sr = sint / sphr;
  const R cc = cr * cr, cs = cr * sr, ss = sr * sr;
  const R c2 = cc - ss, s2 = (R)2 * cs;
  for(int q = 0; q < 6; ++q) lap[q] = (d11[q] + d22[q]) / sphr_2;

  const R drr_lnd  = ss * d11[LND]  + cc * d22[LND]  + s2 * d12_lnd  + cr * dz[LND]  - sr * dr[LND];
  const R drr_ur   = ss * d11[UR]   + cc * d22[UR]   + s2 * d12_ur   + cr * dz[UR]   - sr * dr[UR];
  const R drr_lner = ss * d11[LNER] + cc * d22[LNER] + s2 * d12_lner + cr * dz[LNER] - sr * dr[LNER];
  const R drz_lnd  = cs * (d11[LND]  - d22[LND])  + c2 * d12_lnd  - cr * dr[LND]  - sr * dz[LND];
  const R drz_ur   = cs * (d11[UR]   - d22[UR])   + c2 * d12_ur   - cr * dr[UR]   - sr * dz[UR];
  const R drz_uz   = cs * (d11[UZ]   - d22[UZ])   + c2 * d12_uz   - cr * dr[UZ]   - sr * dz[UZ];
  const R drz_lner = cs * (d11[LNER] - d22[LNER]) + c2 * d12_lner - cr * dr[LNER] - sr * dz[LNER];
  const R dzz_lnd  = cc * d11[LND]  + ss * d22[LND]  - s2 * d12_lnd  - cr * dz[LND]  + sr * dr[LND];
  const R dzz_uz   = cc * d11[UZ]   + ss * d22[UZ]   - s2 * d12_uz   - cr * dz[UZ]   + sr * dr[UZ];
  const R dzz_lner = cc * d11[LNER] + ss * d22[LNER] - s2 * d12_lner - cr * dz[LNER] + sr * dr[LNER];

  R dt[6] = {0, 0, 0, 0, 0, 0};
  const R ur = st.f(UR,0,0), uz = st.f(UZ,0,0), Omg = st.f(OMG,0,0);
  dt[LND]  -= ur * dr[LND]  + uz * dz[LND];                         // :121-132
  dt[UR]   -= ur * dr[UR]   + uz * dz[UR] - Omg * Omg * r;
  dt[UZ]   -= ur * dr[UZ]   + uz * dz[UZ];
  dt[OMG]  -= ur * dr[OMG]  + uz * dz[OMG] + (R)2 * ur * Omg / r;
  dt[LNE]  -= ur * dr[LNE]  + uz * dz[LNE];
  dt[LNER] -= ur * dr[LNER] + uz * dz[LNER];

  {                                                                  // :135-158
    const R ur_r = ur / r;
    const R div_u = dr[UR] + dz[UZ] + ur_r;
    dt[LND] -= div_u;
    dt[UR]  -= temp * (dr[LND] + dr[LNE]);
    dt[UZ]  -= temp * (dz[LND] + dz[LNE]);
    dt[LNE] -= div_u * gamma1;

    const R Srr = dr[UR] - div_u / (R)3;
    const R Szz = dz[UZ] - div_u / (R)3;
    const R Spp = ur_r - div_u / (R)3;
    const R Srz = (dz[UR] + dr[UZ]) / (R)2;
    const R Szp = r * dz[OMG] / (R)2;
    const R Spr = r * dr[OMG] / (R)2;
    const R two_nus = (R)2 * (nus + nuSS);
    dt[UR]  += two_nus * (Srr * dr[LND] + Srz * dz[LND]);
    dt[UZ]  += two_nus * (Srz * dr[LND] + Szz * dz[LND]);
    dt[OMG] += two_nus * (Spr * dr[LND] + Szp * dz[LND]) / r;
    dt[LNE] += (two_nus * (Srr * Srr + Szz * Szz + Spp * Spp +
                (R)2 * (Srz * Srz + Szp * Szp + Spr * Spr)) +
               nub * div_u * div_u) / eng;
  }
  {                                                                  // :161-172
    const R tmp1 = (nus + nuSS) + pnu * sphr_2;
    const R tmp2 = (nus + nuSS) / r;
    dt[UR]  += tmp1 * lap[UR]  + tmp2 * (dr[UR] - ur / r);
    dt[UZ]  += tmp1 * lap[UZ]  + tmp2 * (dr[UZ]);
    dt[OMG] += tmp1 * lap[OMG] + tmp2 * (dr[OMG] * (R)3);
    const R mixed = (nus + nuSS) / (R)3 + nub;
    dt[UR] += mixed * (drr_ur + drz_uz + (dr[UR] - ur / r) / r);
    dt[UZ] += mixed * (drz_ur + dzz_uz + dz[UR] / r);
  }
  {                                                                  // :175-235
    const R lap_d_d = lap[LND] + dr[LND] * dr[LND] + dz[LND] * dz[LND];
    const R lap_e_e = lap[LNE] + dr[LNE] * dr[LNE] + dz[LNE] * dz[LNE];
    dt[LND] += lap_d_d * pdd * sphr_2;
    dt[LNE] += lap_e_e * (ped * sphr_2 + kap * gamma1 + g * nuSS);

    const R kappa = ff * den * den * std::pow(temp, (R)-3.5);
    const R chi = es * den + kappa;
    const R d_lner_2 = dr[LNER] * dr[LNER] + dz[LNER] * dz[LNER];
    const R bigR = std::sqrt(d_lner_2) / chi;
    R Lambda, Lambdap;
    flux_limiter(bigR, Lambda, Lambdap);
    dt[UR] -= Lambda * rad * (dr[LND] + dr[LNE]);
    dt[UZ] -= Lambda * rad * (dz[LND] + dz[LNE]);

    const R temp2 = temp * temp;
    const R cooling = kappa * (ar * temp2 * temp2 / den - rad);
    dt[LNE]  -= cooling / eng;
    dt[LNER] += cooling / rad;

    const R d_lnd_d_lner = dr[LND] * dr[LNER] + dz[LND] * dz[LNER];
    const R lap_er_er = lap[LNER] + d_lner_2;
    const R nu_rad = Lambda / chi;
    const R dr_lnEr = dr[LND] + dr[LNER];
    const R dz_lnEr = dz[LND] + dz[LNER];
    const R drr_lnEr = drr_lnd + drr_lner;
    const R drz_lnEr = drz_lnd + drz_lner;
    const R dzz_lnEr = dzz_lnd + dzz_lner;
    const R dr_chi = chi * dr[LND] - (R)3.5 * kappa * dr[LNE];
    const R dz_chi = chi * dz[LND] - (R)3.5 * kappa * dz[LNE];
    const R dr_bigR = ((dr_lnEr * drr_lnEr + dz_lnEr * drz_lnEr) /
                       ((R)1e-16 + bigR) / chi - bigR * dr_chi) / chi;
    const R dz_bigR = ((dr_lnEr * drz_lnEr + dz_lnEr * dzz_lnEr) /
                       ((R)1e-16 + bigR) / chi - bigR * dz_chi) / chi;
    dt[LNER] += lap_er_er * prd * sphr_2 +
      nu_rad * (lap_d_d + (R)2 * d_lnd_d_lner + lap_er_er) +
      dr_lnEr * (Lambdap * dr_bigR - nu_rad * dr_chi) / chi +
      dz_lnEr * (Lambdap * dz_bigR - nu_rad * dz_chi) / chi;
  }
  {                                                                  // :238-244
    const R tmp = sphr - rS;
    const R eps = S.pr[15];   // added: Plummer softening M d / (d^2 + eps^2)^(3/2)
    const R gr = eps == (R)0 ? M / (tmp * tmp)
                             : M * tmp / ((tmp * tmp + eps * eps) * std::sqrt(tmp * tmp + eps * eps));
    dt[UR] -= sint * gr;
    dt[UZ] -= cost * gr;
  }
  for(int q = 0; q < 6; ++q) out[q] = dt[q];
}

// ---------------------------------------------------------------------------
// Boundary fill, src/bcond.cu:22-86.  Dim 1 covers every column including
// the dim-2 ghosts; dim 2 then covers every row including the dim-1 ghosts,
// which is what sets the corners.
// ---------------------------------------------------------------------------
template<class R> int odd_var(const Sim<R> &S)
{ return S.module == DISK ? 1 : 2; }   // src/*/bcond.cuh: u2 (hydro, isot), ur (disk)

template<class R> void bcond(Sim<R> &S, std::vector<R> &a)
{
  const int nv = S.nv;
  for(long k = 0; k < HALF; ++k)
    for(long j = -HALF; j < S.n2 + HALF; ++j) {
      long lo_dst = -(k + 1), hi_dst = S.n1 + k, lo_src, hi_src;
      if(S.p1) { lo_src = S.n1 - (k + 1); hi_src = k; }            // :31-36
      else     { lo_src = k;              hi_src = S.n1 - 1 - k; } // :45-57
      std::memcpy(S.cell(a, lo_dst, j), S.cell(a, lo_src, j), nv * sizeof(R));
      std::memcpy(S.cell(a, hi_dst, j), S.cell(a, hi_src, j), nv * sizeof(R));
    }
  const int odd = odd_var(S);
  for(long i = -HALF; i < S.n1 + HALF; ++i)
    for(long k = 0; k < HALF; ++k) {
      long lo_dst = -(k + 1), hi_dst = S.n2 + k, lo_src, hi_src;
      if(S.p2) { lo_src = S.n2 - (k + 1); hi_src = k; }            // :38-43
      else     { lo_src = k;              hi_src = S.n2 - 1 - k; } // :59-75
      R *ld = S.cell(a, i, lo_dst), *hd = S.cell(a, i, hi_dst);
      const R *ls = S.cell(a, i, lo_src), *hs = S.cell(a, i, hi_src);
      for(int q = 0; q < nv; ++q) {
        const R sgn = (!S.p2 && q == odd) ? (R)-1 : (R)1;
        ld[q] = S.p2 ? ls[q] : sgn * ls[q];
        hd[q] = S.p2 ? hs[q] : sgn * hs[q];
      }
    }
}

template<class R> void rhs_cell(const Sim<R> &S, const R *c, long i, long j, R *out)
{
  Stencil<R> st{c, S.w * S.nv, S.nv, S.Delta1, S.Delta2};
  if(S.module == HYDRO) rhs_hydro(S, st, out);
  else if(S.module == HYDRO_ISOT) rhs_isot(S, st, i, out);
  else rhs_disk(S, st, i, j, out);
}

// kick: v <- F(u) + beta v over the interior; src/kick.cu:41-57,98.
template<class R> void kick(Sim<R> &S, R beta)
{
  S.refresh();
  const int nv = S.nv;
  #pragma omp parallel for schedule(static)
  for(long i = 0; i < S.n1; ++i) {
    R f[6];
    for(long j = 0; j < S.n2; ++j) {
      const size_t o = S.idx(i, j);
      rhs_cell(S, &S.u[o], i, j, f);
      for(int q = 0; q < nv; ++q) S.v[o + q] = beta * S.v[o + q] + f[q];
    }
  }
}

// drift: u <- u + dt v, with the sticky NaN flag; src/drift.cu:23-32.
template<class R> void drift(Sim<R> &S, R dt)
{
  const int nv = S.nv;
  int flag = 0;
  #pragma omp parallel for schedule(static) reduction(|:flag)
  for(long i = 0; i < S.n1; ++i)
    for(long j = 0; j < S.n2; ++j) {
      const size_t o = S.idx(i, j);
      for(int q = 0; q < nv; ++q) {
        S.u[o + q] += dt * S.v[o + q];
        if(S.u[o + q] != S.u[o + q]) flag = 1;
      }
    }
  if(flag) S.nan_flag = 1;
}

// Point-wise first_order pre-step; src/kick.cu:116-138 and src/*/Euler.cuh.
template<class R> void euler(Sim<R> &S, R dt)
{
  if(S.module == HYDRO) return;                                   // hydro/Euler.cuh:19-22
  if(S.module == HYDRO_ISOT) {                                    // hydro-isot/Euler.cuh:21-44
    const R ff = S.P("f"), kk = S.P("k");
    // RANDOMIZE runs on every call, forcing or not (src/kick.cu:132-135).
    S.rand1 = 6.2831853071795865 * unit_double(splitmix64(S.rng));
    S.rand2 = 6.2831853071795865 * unit_double(splitmix64(S.rng));
    if(!(ff * kk > (R)0)) return;
    const R r1 = (R)S.rand1, r2 = (R)S.rand2;
    const R m1 = kk * std::cos(r1), m2 = kk * std::sin(r1);
    const long k1 = (long)(m1 + (R)0.5), k2 = (long)(m2 + (R)0.5);
    #pragma omp parallel for schedule(static)
    for(long i = 0; i < S.n1; ++i)
      for(long j = 0; j < S.n2; ++j) {
        R *c = S.cell(S.u, i, j);
        const R phi = (R)6.2831853071795865 *
          ((R)(k1 * i) * S.Delta1 + (R)(k2 * j) * S.Delta2);
        const R amp = std::sqrt(dt) * (ff * std::cos(phi + r2)) / (kk * std::exp(c[0]));
        c[1] += amp * m1;
        c[2] += amp * m2;
      }
    return;
  }
  // disk/Euler.cuh:19-69
  const R g = S.P("gamma"), ar = S.P("ar"), ff = S.P("ff");
  #pragma omp parallel for schedule(static)
  for(long i = 0; i < S.n1; ++i)
    for(long j = 0; j < S.n2; ++j) {
      R *c = S.cell(S.u, i, j);
      const R den = std::exp(c[0]), eng = std::exp(c[4]), rad = std::exp(c[5]);
      const R gamma1 = g - (R)1;
      const R g2 = gamma1 * gamma1;
      const R scale = std::pow(ar * g2 * g2 / den, (R)1 / (R)3);
      const R x0 = (eng + rad) * scale;
      const R a = std::pow(std::sqrt((R)0.00390625 + x0 * x0 * x0 / (R)27) + (R)0.0625,
                           (R)1 / (R)3);
      const R b = a - x0 / ((R)3 * a);
      const R cc = std::sqrt((R)2 * b);
      const R eq_eng = (std::sqrt((R)0.5 / cc - (R)0.5 * b) - (R)0.5 * cc) / scale;
      const R dt_eng = (eq_eng - eng) / dt;
      const R temp = gamma1 * eng;
      const R temp2 = temp * temp;
      const R kappa = ff * den * den * std::pow(temp, (R)-3.5);
      const R cooling = kappa * (ar * temp2 * temp2 / den - rad);
      if(cooling * cooling > dt_eng * dt_eng) {
        c[4] = std::log(eq_eng);
        c[5] = std::log(eng + rad - eq_eng);
      }
    }
}

// One low-storage RK3 step; src/step.cpp:22-43.  Coefficients in E (double)
// are rounded to R where the reference passes them to R-typed arguments.
template<class R> int step(Sim<R> &S, double t, double dt)
{
  const double alpha[] = {0.0, 1.0 / 3.0, 3.0 / 4.0};
  const double beta[]  = {0.0, -5.0 / 9.0, -153.0 / 128.0};
  const double gamma[] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
  (void)alpha; (void)t;   // t is passed to the kernel but unused (src/kick.cu:59-106)
  euler(S, (R)dt);
  for(int k = 0; k < 3; ++k) {
    bcond(S, S.u);
    kick(S, (R)beta[k]);
    drift(S, (R)(dt * gamma[k]));
  }
  return S.nan_flag;
}

// Module settings() and the Delta upload; src/config.cpp:24-41,
// src/hydro-isot/config.h:21-26, src/disk/config.h:21-34.
template<class R> void config(Sim<R> &S)
{
  if(S.module == HYDRO_ISOT) { S.l1 = S.l2 = (double)(R)(2.0 * M_PI); }
  if(S.module == DISK) {
    S.l2 = (double)(R)M_PI;
    const R hd2 = (R)(0.5 * S.l2 / S.n2);
    const R hd1 = (R)std::log(hd2 + std::sqrt(hd2 * hd2 + 1.0));
    S.l1 = (double)(R)(2.0 * hd1 * S.n1);
    S.p1 = 0; S.p2 = 0;
  }
  S.Delta1 = (R)(S.l1 / S.n1);
  S.Delta2 = (R)(S.l2 / S.n2);
}

// ---------------------------------------------------------------------------
// Time steps
// ---------------------------------------------------------------------------

// Closed form, src/getdt.cpp:26 (host E arithmetic; n/l in R).
template<class R> double getdt(const Sim<R> &S)
{
  const R a = (R)S.n1 / (R)S.l1, b = (R)S.n2 / (R)S.l2;
  return (double)(R)S.c / (double)(R)(a + b);
}

// Largest modified wavenumbers of the 6th-order D1 (max over kh of
// (45 sin kh - 9 sin 2kh + sin 3kh)/30) and of -D11 (at kh = pi: 1088/180).
constexpr double KMAX1 = 1.5859783962656393;
constexpr double KMAX2 = 1088.0 / 180.0;
// RK3 stability limits on the imaginary and negative real axes:
// |1 + z + z^2/2 + z^3/6| = 1 at z = i sqrt(3) and at z = -2.5127.
constexpr double RK3_IMAG = 1.7320508075688772;
constexpr double RK3_REAL = 2.5127453266183286;

// CFL step: dt = c * min(sqrt3 / lambda_adv, 2.5127 / lambda_diff), minimum
// over cells, with lambda_adv = (|u1| + cs) k1/Delta1 + (|u2| + cs) k1/Delta2
// and lambda_diff = D_max * k2 (1/Delta1^2 + 1/Delta2^2).  New work: the
// reference has only the closed form (src/getdt.cpp:26-28).
template<class R> double getdt_cfl(const Sim<R> &S)
{
  const double k1a = KMAX1 / (double)S.Delta1, k1b = KMAX1 / (double)S.Delta2;
  const double k2s = KMAX2 * (1.0 / ((double)S.Delta1 * S.Delta1) +
                              1.0 / ((double)S.Delta2 * S.Delta2));
  double lam_adv = 0.0, dmax = 0.0;
  if(S.module == HYDRO) {
    const double g = S.get("gamma");
    dmax = std::max({S.get("nus") * 4.0 / 3.0 + S.get("nub"),
                     (g - 1.0) * S.get("kappa"), S.get("dd")});
  } else if(S.module == HYDRO_ISOT) {
    dmax = std::max(S.get("nus") * 4.0 / 3.0 + S.get("nub"), S.get("dd"));
  }
  double dmax_disk = 0.0;
  for(long i = 0; i < S.n1; ++i)
    for(long j = 0; j < S.n2; ++j) {
      const R *c = &S.u[S.idx(i, j)];
      double a1 = std::fabs((double)c[1]), a2 = std::fabs((double)c[2]), cs;
      if(S.module == HYDRO)
        cs = std::sqrt(S.get("gamma") * (S.get("gamma") - 1.0) * std::exp((double)c[3]));
      else if(S.module == HYDRO_ISOT)
        cs = std::sqrt(S.get("cs2"));
      else {
        // disk: velocities (ur, uz) rotated into the log-r / theta directions
        // have |.| <= |ur| + |uz|; metric factor 1/sphr on both axes.
        const double g = S.get("gamma");
        const double sphr = S.get("rin") * std::exp((i + 0.5) * (double)S.Delta1);
        const double theta = (j + 0.5) * (double)S.Delta2;
        const double r = sphr * std::sin(theta);
        const double temp = (g - 1.0) * std::exp((double)c[4]);
        const double vv = a1 + a2;
        cs = std::sqrt(g * temp + std::exp((double)c[5]) / 3.0);
        const double la = (vv + cs) * (k1a + k1b) / sphr;
        lam_adv = std::max(lam_adv, la);
        const double nuSS = S.get("alpha") * g * temp * r * std::sqrt(r / S.get("M"));
        const double nuv = S.get("nus") + nuSS;
        const double dloc = std::max({nuv * 4.0 / 3.0 + S.get("nub") + S.get("nu") * sphr * sphr,
                                      S.get("dd") * sphr * sphr,
                                      S.get("ed") * sphr * sphr + S.get("kappa") * (g - 1.0) + g * nuSS,
                                      S.get("rd") * sphr * sphr + 1.0 / (3.0 * S.get("es") * std::exp((double)c[0]))})
                            / (sphr * sphr);
        dmax_disk = std::max(dmax_disk, dloc);
        continue;
      }
      lam_adv = std::max(lam_adv, (a1 + cs) * k1a + (a2 + cs) * k1b);
    }
  if(S.module == DISK) dmax = dmax_disk;
  double dt = std::numeric_limits<double>::infinity();
  if(lam_adv > 0) dt = std::min(dt, RK3_IMAG / lam_adv);
  if(dmax > 0) dt = std::min(dt, RK3_REAL / (dmax * k2s));
  return S.c * dt;
}

// ---------------------------------------------------------------------------
// Initial conditions, src/init.cpp:26-42 and the modules' init.h.  The
// reference's default build has R = float: the cell centres and the
// parameters its IC functions read are FP32 values, so they are rounded to
// float here too; each function is then evaluated in FP64.  The libc rand()
// draws are replaced by one sequential SplitMix64 stream started at `seed`,
// consumed in the reference's call order (row-major over i then j; one draw
// per cell in KH, two in Kf), each mapped to ((z >> 11) + 1/2) 2^-53.
// ---------------------------------------------------------------------------
double open_unit(uint64_t &state) { return ((double)(splitmix64(state) >> 11) + 0.5) * 0x1.0p-53; }
double fermi_dirac(double a, double d) { return 1.0 / (std::exp(a / d) + 1.0); }

// src/hydro/init.h:26-112 -> (lnd, u1, u2, lne)
void ic_hydro(const std::string &name, double g, double l1, double l2, double x, double y,
              uint64_t &rng, double *c)
{
  x = (float)(x - 0.5 * l1);                       // R-typed shift, init.h:30-31 (each IC)
  y = (float)(y - 0.5 * l2);
  double lnd = 0, u1 = 0, lne = 0;
  if(name == "blast") {                            // init.h:28-36
    lne = std::log(((x * x + y * y < 0.01) ? 1.0 : 0.01) / (g - 1.0));
  } else if(name == "bow") {                       // init.h:38-49
    const double r = std::sqrt(x * x + y * y);
    const double d = 9.9 * fermi_dirac(r - 0.01, 0.001) + 0.1;
    lnd = std::log(d);
    u1 = (10.0 / d - 1.0) / 99.0;
    lne = std::log(0.01 / d / (g - 1.0));
  } else if(name == "implos") {                    // init.h:51-60
    const double d = 0.125 + 0.875 * fermi_dirac(0.15 - std::fabs(x) - std::fabs(y), 0.001);
    const double p = 0.14 + 0.86 * fermi_dirac(0.15 - std::fabs(x) - std::fabs(y), 0.001);
    lnd = std::log(d);
    lne = std::log(p / d / (g - 1.0));
  } else if(name == "KH") {                        // init.h:62-79
    double d;
    if(std::fabs(y) < 0.25) { d = 2.0; u1 = 0.5 + 0.01 * open_unit(rng) - 0.005; }
    else                    { d = 1.0; u1 = -0.5 + 0.01 * open_unit(rng) - 0.005; }
    lnd = std::log(d);
    lne = std::log(2.5 / d / (g - 1.0));
  } else if(name == "Sod") {                       // init.h:81-97
    double d, P;
    if(std::fmod(x - 2 * y + 1.5, 1.0) < 0.5) { d = 1.0; P = 1.0; }
    else                                      { d = 0.125; P = 0.1; }
    lnd = std::log(d);
    lne = std::log(P / d / (g - 1.0));
  }                                                // otherwise zeros, init.h:99-102
  c[0] = lnd; c[1] = u1; c[2] = 0.0; c[3] = lne;
}

// src/hydro-isot/init.h:21-49 -> (lnd, u1, u2)
void ic_isot(const std::string &name, double nus, double fi, double ki, double x,
             uint64_t &rng, double *c)
{
  c[0] = c[1] = c[2] = 0.0;
  if(name != "Kf") return;
  const double r1 = open_unit(rng);
  const double r2 = open_unit(rng);
  const double q1 = std::sqrt(-2.0 * std::log(r1)) * std::sin(2.0 * M_PI * r2);
  const double q2 = std::sqrt(-2.0 * std::log(r1)) * std::cos(2.0 * M_PI * r2);
  const double a = fi / (nus * ki * ki);
  c[1] = 1.0e-9 * q1;
  c[2] = 1.0e-9 * q2 + a * std::cos(ki * x);
}

// src/disk/init.h:25-68 (Hawley 2000 torus) -> (lnd, ur, uz, Omg, lne, lner)
void ic_disk(double M, double rin, double Gamma, double ar, double lnr, double theta, double *c)
{
  const double r = rin * std::exp(lnr);
  const double cyl_r = r * std::sin(theta);
  const double q0 = 2.0, r0 = 16.0, d0 = 1.0, e0 = 0.01, bg = 0.0001;
  const double g1 = Gamma - 1.0;
  const double q1 = 2.0 * q0 - 2.0;
  const double lK = std::sqrt(M * std::pow(r0, q1 - 1.0));
  const double K = (g1 * e0) / std::pow(d0, g1);
  double den = Gamma * e0 - M / r0 + lK * lK / (std::pow(r0, q1) * q1)
             + M / r - lK * lK / (std::pow(cyl_r, q1) * q1);
  if(den > 0.0) den = std::pow(den * g1 / (Gamma * K), 1.0 / g1);
  else          den = -std::pow(-den * g1 / (Gamma * K), 1.0 / g1);
  double Omg = lK * std::pow(cyl_r, -q0);
  if(den < bg) { den = bg; Omg *= std::sin(theta) * std::sin(theta); }
  const double eng = K * std::pow(den, g1) / g1;
  const double tbg = K * std::pow(bg, g1);
  const double rad = ar * tbg * tbg * tbg * tbg / bg;
  c[0] = std::log(den); c[1] = 0.0; c[2] = 0.0; c[3] = Omg; c[4] = std::log(eng); c[5] = std::log(rad);
}

template<class R> void init_ic(Sim<R> &S, const std::string &name, uint64_t seed)
{
  uint64_t rng = seed;
  const float l1 = (float)S.l1, l2 = (float)S.l2;
  auto pf = [&](const char *n) { return (double)(float)S.get(n); };
  for(long i = 0; i < S.n1; ++i) {
    const float x = (float)((double)l1 * (i + 0.5) / S.n1);   // src/init.cpp:31
    for(long j = 0; j < S.n2; ++j) {
      const float y = (float)((double)l2 * (j + 0.5) / S.n2); // src/init.cpp:33
      double c[6];
      if(S.module == HYDRO) ic_hydro(name, pf("gamma"), l1, l2, x, y, rng, c);
      else if(S.module == HYDRO_ISOT) ic_isot(name, pf("nus"), pf("f"), pf("k"), x, rng, c);
      else ic_disk(pf("M"), pf("rin"), pf("gamma"), pf("ar"), x, y, c);
      R *u = S.cell(S.u, i, j);
      for(int q = 0; q < S.nv; ++q) u[q] = (R)c[q];
    }
  }
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI for the ctypes wrapper (oracle/oracle.py)
// ---------------------------------------------------------------------------
struct OrcHandle { int fp64; void *sim; };

template<class F> auto dispatch(OrcHandle *h, F f)
{
  if(h->fp64) return f(*(Sim<double> *)h->sim);
  return f(*(Sim<float> *)h->sim);
}

extern "C" {

void *orc_create(int module, int n1, int n2, int fp64)
{
  if(module < 0 || module > 2 || n1 < 1 || n2 < 1) return nullptr;
  OrcHandle *h = new OrcHandle{fp64, nullptr};
  if(fp64) h->sim = new Sim<double>(module, n1, n2);
  else     h->sim = new Sim<float>(module, n1, n2);
  return h;
}

void orc_destroy(void *p)
{
  OrcHandle *h = (OrcHandle *)p;
  if(!h) return;
  if(h->fp64) delete (Sim<double> *)h->sim; else delete (Sim<float> *)h->sim;
  delete h;
}

int orc_nvar(void *p) { return dispatch((OrcHandle *)p, [](auto &S) { return S.nv; }); }

int orc_set_param(void *p, const char *name, double val)
{
  return dispatch((OrcHandle *)p, [&](auto &S) {
    for(auto &q : S.par) if(!strcmp(q.name, name)) { q.value = val; return 0; }
    return -1;
  });
}

double orc_get_param(void *p, const char *name)
{ return dispatch((OrcHandle *)p, [&](auto &S) { return S.get(name); }); }

void orc_set_geometry(void *p, double l1, double l2, int p1, int p2, double c)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    S.l1 = l1; S.l2 = l2; S.p1 = p1; S.p2 = p2; S.c = c; return 0; });
}

void orc_get_geometry(void *p, double *out)  // l1 l2 p1 p2 c Delta1 Delta2
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    out[0] = S.l1; out[1] = S.l2; out[2] = S.p1; out[3] = S.p2; out[4] = S.c;
    out[5] = (double)S.Delta1; out[6] = (double)S.Delta2; return 0; });
}

void orc_config(void *p) { dispatch((OrcHandle *)p, [](auto &S) { config(S); return 0; }); }

void orc_init(void *p, const char *name, unsigned long long seed)
{ dispatch((OrcHandle *)p, [&](auto &S) { init_ic(S, name ? name : "", seed); return 0; }); }

unsigned long long orc_splitmix64(unsigned long long *state)
{ uint64_t s = *state; const uint64_t z = splitmix64(s); *state = s; return z; }

void orc_seed(void *p, unsigned long long seed)
{ dispatch((OrcHandle *)p, [&](auto &S) { S.rng = seed; return 0; }); }

void orc_get_rand(void *p, double *out)
{ dispatch((OrcHandle *)p, [&](auto &S) { out[0] = S.rand1; out[1] = S.rand2; return 0; }); }

// Interior state in .raw order: n1 x n2 cells x nvar, no ghosts.
void orc_upload(void *p, const double *x)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    using R = typename std::remove_reference<decltype(S.u[0])>::type;
    size_t k = 0;
    for(long i = 0; i < S.n1; ++i)
      for(long j = 0; j < S.n2; ++j)
        for(int q = 0; q < S.nv; ++q) S.u[S.idx(i, j) + q] = (R)x[k++];
    return 0;
  });
}

void orc_download(void *p, double *x)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    size_t k = 0;
    for(long i = 0; i < S.n1; ++i)
      for(long j = 0; j < S.n2; ++j)
        for(int q = 0; q < S.nv; ++q) x[k++] = (double)S.u[S.idx(i, j) + q];
    return 0;
  });
}

// Whole ghosted u array, (n1 + 6) x (n2 + 6) x nvar.
void orc_download_ghosted(void *p, double *x)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    for(size_t k = 0; k < S.u.size(); ++k) x[k] = (double)S.u[k];
    return 0;
  });
}

void orc_upload_v(void *p, const double *x)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    using R = typename std::remove_reference<decltype(S.v[0])>::type;
    size_t k = 0;
    for(long i = 0; i < S.n1; ++i)
      for(long j = 0; j < S.n2; ++j)
        for(int q = 0; q < S.nv; ++q) S.v[S.idx(i, j) + q] = (R)x[k++];
    return 0;
  });
}

void orc_download_v(void *p, double *x)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    size_t k = 0;
    for(long i = 0; i < S.n1; ++i)
      for(long j = 0; j < S.n2; ++j)
        for(int q = 0; q < S.nv; ++q) x[k++] = (double)S.v[S.idx(i, j) + q];
    return 0;
  });
}

void orc_bcond(void *p) { dispatch((OrcHandle *)p, [](auto &S) { bcond(S, S.u); return 0; }); }

void orc_kick(void *p, double beta)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    using R = typename std::remove_reference<decltype(S.u[0])>::type;
    kick(S, (R)beta); return 0; });
}

void orc_drift(void *p, double dt)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    using R = typename std::remove_reference<decltype(S.u[0])>::type;
    drift(S, (R)dt); return 0; });
}

void orc_euler(void *p, double dt)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    using R = typename std::remove_reference<decltype(S.u[0])>::type;
    euler(S, (R)dt); return 0; });
}

// F(u) over the interior after a boundary fill, into out (.raw order).
void orc_rhs(void *p, double *out)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    using R = typename std::remove_reference<decltype(S.u[0])>::type;
    bcond(S, S.u);
    S.refresh();
    #pragma omp parallel for schedule(static)
    for(long i = 0; i < S.n1; ++i)
      for(long j = 0; j < S.n2; ++j) {
        R f[6];
        rhs_cell(S, &S.u[S.idx(i, j)], i, j, f);
        for(int q = 0; q < S.nv; ++q) out[((size_t)i * S.n2 + j) * S.nv + q] = (double)f[q];
      }
    return 0;
  });
}

// One operator (0 D1, 1 D2, 2 D11, 3 D22, 4 D12) of every variable over the
// interior after a boundary fill, into out (.raw order).
void orc_deriv(void *p, int op, double *out)
{
  dispatch((OrcHandle *)p, [&](auto &S) {
    using R = typename std::remove_reference<decltype(S.u[0])>::type;
    bcond(S, S.u);
    S.refresh();
    for(long i = 0; i < S.n1; ++i)
      for(long j = 0; j < S.n2; ++j) {
        Stencil<R> st{&S.u[S.idx(i, j)], S.w * S.nv, S.nv, S.Delta1, S.Delta2};
        for(int q = 0; q < S.nv; ++q) {
          R d = op == 0 ? st.d1(q) : op == 1 ? st.d2(q) : op == 2 ? st.d11(q)
              : op == 3 ? st.d22(q) : st.d12(q);
          out[((size_t)i * S.n2 + j) * S.nv + q] = (double)d;
        }
      }
    return 0;
  });
}

int orc_step(void *p, double t, double dt)
{ return dispatch((OrcHandle *)p, [&](auto &S) { return step(S, t, dt); }); }

// n steps of fixed dt; returns the NaN flag.
int orc_run(void *p, double t, double dt, int n)
{
  return dispatch((OrcHandle *)p, [&](auto &S) {
    int r = 0;
    for(int k = 0; k < n && !r; ++k) r = step(S, t + k * dt, dt);
    return r;
  });
}

// The disk flux limiter Lambda(R) and Lambda'(R) in FP64 (fp64 != 0) or
// FP32, as rhs_disk evaluates them.
void orc_flux_limiter(double bigR, int fp64, double *out)
{
  if(fp64) { double a, b; flux_limiter<double>(bigR, a, b); out[0] = a; out[1] = b; }
  else { float a, b; flux_limiter<float>((float)bigR, a, b); out[0] = a; out[1] = b; }
}

double orc_getdt(void *p) { return dispatch((OrcHandle *)p, [](auto &S) { return getdt(S); }); }
double orc_getdt_cfl(void *p) { return dispatch((OrcHandle *)p, [](auto &S) { return getdt_cfl(S); }); }
int orc_nan_flag(void *p) { return dispatch((OrcHandle *)p, [](auto &S) { return S.nan_flag; }); }

} // extern "C"
