// capi.hip -- the C-ABI host driver (include/fg2_amd.h).
//
// Host counterparts of the reference's setup/config/para/getdt/step/solve
// (src/setup.cpp, src/config.cpp, src/para.cpp, src/getdt.cpp,
// src/step.cpp, src/solve.cpp), over the kernels of kernels.hip.  The global
// state of the reference (namespace global, src/fg2.h:50-58) lives in a run
// handle, so many runs can coexist (batching, serving).
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "fg2_amd.h"
#include "launch.h"
#include "physics.h"

using namespace fg2;

namespace {

thread_local std::string g_error;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if(e_ != hipSuccess)                                                         \
      return fail(FG2_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));          \
  } while(0)

// A module's parameter names and reference defaults, in Params::par order
// (the module's PARAMS: src/<module>/eqns.cuh).
struct ParamTable {
  const ParamDef *p;
  int n;
  const ParamDef *begin() const { return p; }
  const ParamDef *end() const { return p + n; }
  size_t size() const { return (size_t)n; }
  const ParamDef &operator[](size_t k) const { return p[k]; }
};
ParamTable param_defs(int m) { return ParamTable{module_info(m).params, module_info(m).npar}; }

// Largest modified wavenumber of -D11 (at kh = pi), and the RK3 stability
// limits on the imaginary and negative real axes (|1+z+z^2/2+z^3/6| = 1).
// The D1 counterpart, 1.58598, is applied on the device with the rates.
constexpr double KMAX2 = 1088.0 / 180.0;
constexpr double RK3_IMAG = 1.7320508075688772;
constexpr double RK3_REAL = 2.5127453266183286;

// Williamson low-storage RK3 (src/step.cpp:24-26).  The stage times
// t + alpha dt (alpha = 0, 1/3, 3/4) are not needed: kick ignores t.
constexpr double RK_BETA[3]  = {0.0, -5.0 / 9.0, -153.0 / 128.0};
constexpr double RK_GAMMA[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

// SplitMix64 (Steele, Lea, Flood 2014), the seeded stand-in for rand().
uint64_t splitmix64(uint64_t &state)
{
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void reset_diag_kernel(DeviceDiag *d, int lam_only)
{
  if(!lam_only) d->flags = 0;
  d->lam_adv_bits = 0;
  d->lam_diff_bits = 0;
}

// After timed launches on a run's state: NaN flag back to what it was
// before the timing (a NaN of an earlier unchecked step stays reported),
// CFL rates cleared.
__global__ void restore_diag_kernel(DeviceDiag *d, unsigned int flags)
{
  d->flags = flags;
  d->lam_adv_bits = 0;
  d->lam_diff_bits = 0;
}

// Host values of a run's device clock (DeviceDiag::t .. rng), and what a
// batch reads back of it.
struct ClockInit {
  double t, t_final, fixed_dt, c, lam_diff_const, k2s;
  unsigned long long rng;
};
struct ClockOut {
  double t, dt;                       // dt: the last step, or a stalled clock's rule
  long long steps;
  unsigned long long rng;
  unsigned int flags;
};

// Set a run's device clock (a kernel, not a copy, so it needs no staging
// memory); clear: also the NaN flag and the rates (a fresh served slot).
__global__ void set_clock_kernel(DeviceDiag *d, const ClockInit k, const int clear)
{
  if(clear) { d->flags = 0; d->pad = 0; d->lam_adv_bits = 0; d->lam_diff_bits = 0; }
  d->flags &= ~2u;                    // no stalled clock
  d->t = k.t; d->t_final = k.t_final; d->dt = 0.0;
  d->fixed_dt = k.fixed_dt; d->c = k.c; d->lam_diff_const = k.lam_diff_const; d->k2s = k.k2s;
  d->rng = k.rng;
  d->steps = 0;
}

} // namespace

struct fg2_run {
  int module = 0, nv = 0, n1 = 0, n2 = 0;
  long s = 0;
  double l1 = 1.0, l2 = 1.0, c = 0.5;   // src/global.cpp:28-29
  int p1 = 1, p2 = 1;
  float Delta1 = 0, Delta2 = 0;
  bool configured = false;
  std::vector<double> par;
  float *buf[3] = {nullptr, nullptr, nullptr};   // allocations
  float *in[3] = {nullptr, nullptr, nullptr};    // interior pointers
  int cur = 0;                                   // state lives in in[cur]; v is in[2]
  DeviceDiag *diag = nullptr;
  DeviceDiag *host_diag = nullptr;               // pinned readback
  double *sum_partial = nullptr, *sum_out = nullptr;
  float *rowtab = nullptr;       // disk: per-row table (Params::rowtab)
  SpectralWork *spec = nullptr;  // spectral diagnostics workspace, made on first use
  hipStream_t stream = nullptr;
  int path = FG2_PATH_FUSED, dt_mode = FG2_DT_CLOSED;
  FusedShape shape{0, 1};
  bool tile_auto = true;
  int kernel_pick = FG2_KERNEL_AUTO;   // fused-path kernel of a one-pass module
  uint64_t rng = 0;
  float rand1 = 0, rand2 = 0;
  bool lam_valid = false;        // diag holds the CFL rates of the current state
  char para_buf[256];
  size_t bytes = 0;
  int cus = 256;                 // compute units of the run's device
  // slab of a taller periodic grid (slab_n1g = 0: the run is a whole grid)
  int slab_row0 = 0, slab_n1g = 0;
  float *slab_buf = nullptr;     // halo rows: SLAB_HALO above, then SLAB_HALO below, stride s
  hipEvent_t slab_ev = nullptr;  // marks the state the neighbours copy from
};

namespace {

// Resident workgroups per CU of a kernel family (kind 0: substage kernels,
// kind 1: one-pass step kernel), from the hardware occupancy of the
// compiled kernels; cached per (module, kind, waves).
int blocks_per_cu(int module, int kind, int waves)
{
  static std::map<std::tuple<int, int, int>, int> cache;
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_tuple(module, kind, waves);
  auto it = cache.find(key);
  if(it != cache.end()) return it->second;
  int nb = 0;
  if(kernel_occupancy(module, kind, waves, &nb) != hipSuccess || nb < 1) nb = 1;
  cache[key] = nb;
  return nb;
}

// Rows per tile that minimise (rounds of resident tiles) x (cost of a tile),
// a tile costing its rows plus `extra` (the halo rows it loads or
// recomputes).  Filling whole rounds matters: the kernels are VALU bound,
// and a last round that is half full leaves half the SIMDs idle.
int tile_rows(long n1, long nstrips, long slots, long h_min, long extra)
{
  long best_h = std::max(1L, std::min(n1, h_min)), best_cost = -1;
  for(long h = h_min; h <= std::min(n1, 1024L); ++h) {
    const long tiles = nstrips * ((n1 + h - 1) / h);
    const long cost = ((tiles + slots - 1) / slots) * (h + extra);
    if(best_cost < 0 || cost < best_cost) { best_cost = cost; best_h = h; }
  }
  return (int)best_h;
}

// Persistent grid of a batched launch: one resident round of workgroups of
// the batch kernel that runs (its hardware occupancy times the CUs).
int batch_grid(int module, bool step, int waves, int cus, bool cfl = false)
{
  return std::max(1, blocks_per_cu(module, step ? (cfl ? 5 : 3) : 2, waves) * cus);
}

// Substage kernels: a tile is one wavefront's 58-column strip, loading 6
// halo rows besides its own.  A grid that fits one round with one- or
// two-row tiles takes them (hydro 128^2 .. 512^2: 10-18 us per step against
// 24-27 us with 8-row tiles).
int auto_tile_h(const fg2_run *r, int waves)
{
  const long nstrips = (r->n2 + STRIP - 1) / STRIP;
  const long slots = (long)blocks_per_cu(r->module, 0, waves) * r->cus * waves;
  return tile_rows(r->n1, nstrips, slots, 1, 3);
}

Params make_params(const fg2_run *r)
{
  Params P{};
  P.n1 = r->n1; P.n2 = r->n2; P.p1 = r->p1; P.p2 = r->p2;
  P.Delta1 = r->Delta1; P.Delta2 = r->Delta2;
  const double D1 = r->Delta1, D2 = r->Delta2;
  const double c1[3] = {45.0 / 60.0, -9.0 / 60.0, 1.0 / 60.0};
  const double c2[4] = {-490.0 / 180.0, 270.0 / 180.0, -27.0 / 180.0, 2.0 / 180.0};
  const double cm[3] = {270.0 / 720.0, -27.0 / 720.0, 2.0 / 720.0};
  for(int k = 0; k < 3; ++k) {
    P.a1[k] = (float)(c1[k] / D1);
    P.a2[k] = (float)(c1[k] / D2);
    P.m12[k] = (float)(cm[k] / (D1 * D2));
  }
  for(int k = 0; k < 4; ++k) {
    P.b1[k] = (float)(c2[k] / (D1 * D1));
    P.b2[k] = (float)(c2[k] / (D2 * D2));
  }
  P.blap0 = (float)(c2[0] / (D1 * D1) + c2[0] / (D2 * D2));
  for(size_t k = 0; k < r->par.size() && k < 16; ++k) P.par[k] = (float)r->par[k];
  P.rand1 = r->rand1; P.rand2 = r->rand2;
  P.rowtab = r->rowtab;
  return P;
}

static_assert(FG2_SLAB_HALO == SLAB_HALO, "halo rows of a slab step");

// Rows of the grid along dim 1: the run's own, or the whole grid's for a slab.
int rows_of_grid(const fg2_run *r) { return r->slab_n1g ? r->slab_n1g : r->n1; }

// Halo rows of a slab for the step kernel (interior cell 0 of each buffer).
SlabHalo slab_halo_of(const fg2_run *r)
{
  return SlabHalo{r->slab_buf + ROW_PAD, r->slab_buf + (long)SLAB_HALO * r->s + ROW_PAD,
                  r->slab_row0, r->slab_n1g};
}

bool first_order_active(const fg2_run *r)
{
  float par[16] = {};
  for(size_t k = 0; k < r->par.size() && k < 16; ++k) par[k] = (float)r->par[k];
  return module_info(r->module).first_order_on(par);
}

// Draw the two random phases of one Euler() call of a module whose pre-step
// takes them (src/hydro-isot/Euler.cuh:39-44).
void randomize(fg2_run *r)
{
  if(!module_info(r->module).random_phases) return;
  const double u1 = (double)(splitmix64(r->rng) >> 11) * 0x1.0p-53;
  const double u2 = (double)(splitmix64(r->rng) >> 11) * 0x1.0p-53;
  r->rand1 = (float)(6.2831853071795865 * u1);
  r->rand2 = (float)(6.2831853071795865 * u2);
}

int check_run(const fg2_run *r)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  if(!r->configured) return fail(FG2_ERR_ARG, "run not configured: call fg2_config first");
  return FG2_OK;
}

FusedShape shape_of(const fg2_run *r)
{
  FusedShape sh = r->shape;
  if(sh.waves != 1 && sh.waves != 2 && sh.waves != 4) sh.waves = 1;
  if(r->tile_auto || sh.tile_h < 1) sh.tile_h = auto_tile_h(r, sh.waves);
  return sh;
}

FusedShape step_shape_of(const fg2_run *r);

// A grid for which the one-pass kernel takes its small-grid shape, or its
// shortest tiles (12 rows) filling at most three quarters of one round of
// resident workgroups, is latency bound: a one-pass tile takes tile_h + 12
// or more dependent iterations of three stages, while the substage path's
// short tiles take three short launches.  Per step, substage against
// one-pass: hydro 128^2 10.2 / 12.9 us, 512^2 18.1 / 22.0 us, 896^2 (12-row
// tiles, 0.59 of a round) 32.5 / 38.6 us; hydro-isot 1024^2 (0.56) 31.0 /
// 33.9 us.  Fuller rounds go the other way: hydro 1024^2 (0.84) 38.8 / 39.2
// us, hydro-isot 1280^2 (0.84) 44.5 / 40.0 us.
bool latency_bound(const fg2_run *r)
{
  if(!r->tile_auto || r->slab_n1g) return false;
  const FusedShape sh = step_shape_of(r);
  if(sh.waves == 1) return true;
  if(sh.tile_h > 12) return false;
  const int kind = (r->p1 && r->p2) ? 1 : 4;
  const long slots = (long)blocks_per_cu(r->module, kind, sh.waves) * r->cus;
  return 4 * step_launch_tiles(r->n1, r->n2, sh.tile_h, sh.waves, nullptr) <= 3 * slots;
}

// The one-pass step kernel serves the Cartesian modules, periodic or not,
// except latency-bound grids with the automatic shape (the substage path),
// unless the run's kernel is fixed (slabs always take the one-pass kernel).
bool uses_step_kernel(const fg2_run *r)
{
  if(r->path != FG2_PATH_FUSED || !module_info(r->module).one_pass) return false;
  if(r->slab_n1g || r->kernel_pick == FG2_KERNEL_ONE_PASS) return true;
  if(r->kernel_pick == FG2_KERNEL_SUBSTAGE) return false;
  return !latency_bound(r);
}

FusedShape step_shape_of(const fg2_run *r)
{
  FusedShape sh = r->shape;
  if(sh.waves != 1 && sh.waves != 2 && sh.waves != 4) sh.waves = 4;
  if(r->tile_auto || sh.tile_h < 1) {
    // a workgroup tile of 64 NW - 18 output columns; stage 0 recomputes 12
    // halo rows and stage 1 six, so a tile costs about its rows + 6.  Tiles
    // of at least 12 rows: below that the recomputed rows dominate (on a
    // 448 x 3584 slab, 14-row tiles fill one round exactly and were the
    // fastest measured, 7% ahead of 16)
    // (tiles counted as launched: a narrow last strip takes one workgroup
    // slot per four row chunks)
    sh.waves = 4;
    const int kind = (r->p1 && r->p2) ? 1 : 4;
    const long slots = (long)blocks_per_cu(r->module, kind, sh.waves) * r->cus;
    long best_cost = -1;
    sh.tile_h = std::max(1, std::min(r->n1, 12));
    for(int h = 12; h <= std::min(r->n1, 1024); ++h) {
      const long cost = ((step_launch_tiles(r->n1, r->n2, h, 4, nullptr) + slots - 1) / slots) *
                        (h + 6);
      if(best_cost < 0 || cost < best_cost) { best_cost = cost; sh.tile_h = h; }
    }
    // A grid that fills under half of one round is latency bound: a tile
    // takes tile_h + 12 dependent iterations whatever the load.  Such grids
    // take one-wavefront workgroups (46 output columns, no waits on other
    // wavefronts) and the shortest tiles for which the launch keeps at most
    // 1.5 wavefronts per SIMD.  Measured on hydro 128^2 .. 512^2 this runs
    // a step in 13-21 us against 32-34 us for 12-row tiles of 4 wavefronts.
    const long tiles = step_launch_tiles(r->n1, r->n2, sh.tile_h, 4, nullptr);
    if(2 * tiles <= slots) {
      const long strips1 = (r->n2 + 63 - 6 * HALF) / (64 - 6 * HALF);
      const long cap = 6L * r->cus;                 // 1.5 wavefronts x 4 SIMDs per CU
      long h = 1;
      while(h < r->n1 && strips1 * ((r->n1 + h - 1) / h) > cap) ++h;
      sh.waves = 1;
      sh.tile_h = (int)h;
    }
  }
  return sh;
}

// Diffusive rate of a module with a constant D_max (its diffusivity()):
// D_max (k2/D1^2 + k2/D2^2); a negative D_max means a per-cell rate.
double lam_diff_const(const fg2_run *r)
{
  const double k2s = KMAX2 * (1.0 / ((double)r->Delta1 * r->Delta1) +
                              1.0 / ((double)r->Delta2 * r->Delta2));
  return module_info(r->module).diffusivity(r->par.data()) * k2s;
}

bool cell_diffusivity(const fg2_run *r) { return module_info(r->module).diffusivity(r->par.data()) < 0.0; }

double dt_from_rates(const fg2_run *r, double lam_adv, double lam_diff_cell)
{
  const double k2s = KMAX2 * (1.0 / ((double)r->Delta1 * r->Delta1) +
                              1.0 / ((double)r->Delta2 * r->Delta2));
  const double lam_diff = cell_diffusivity(r) ? lam_diff_cell * k2s : lam_diff_const(r);
  double dt = INFINITY;
  if(lam_adv > 0) dt = std::min(dt, RK3_IMAG / lam_adv);
  if(lam_diff > 0) dt = std::min(dt, RK3_REAL / lam_diff);
  return r->c * dt;
}

// A run's device clock from t to t_final: a fixed step (fixed_dt > 0) or the
// CFL rule of dt_from_rates, with the run's forcing stream.
ClockInit clock_of(const fg2_run *r, double t, double t_final, double fixed_dt)
{
  ClockInit k;
  k.t = t;
  k.t_final = t_final;
  k.fixed_dt = fixed_dt;
  k.c = r->c;
  k.lam_diff_const = cell_diffusivity(r) ? -1.0 : lam_diff_const(r);
  k.k2s = KMAX2 * (1.0 / ((double)r->Delta1 * r->Delta1) + 1.0 / ((double)r->Delta2 * r->Delta2));
  k.rng = r->rng;
  return k;
}

double bits_to_double(unsigned long long b)
{
  double d;
  memcpy(&d, &b, sizeof d);
  return d;
}

// Refill the disk row table after a change of Delta1, M, rS or rin; the
// stream is drained so batches on other streams see the table.
int refresh_rows(fg2_run *r)
{
  if(!r->rowtab || !r->configured) return FG2_OK;
  HIP_TRY(launch_rows(r->module, r->rowtab, make_params(r), r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return FG2_OK;
}

int substages(fg2_run *r, double t, double dt, int cfl)
{
  const Params P = make_params(r);
  (void)t;  // the reference passes t + alpha dt to kick, which ignores it
  if(first_order_active(r))
    HIP_TRY(launch_euler(r->module, r->in[r->cur], r->s, P, (float)dt, r->stream));
  if(r->path == FG2_PATH_PLAIN) {
    float *u = r->in[r->cur], *v = r->in[2];
    for(int k = 0; k < 3; ++k) {
      HIP_TRY(launch_bcond(r->module, u, r->s, P, r->p1, r->p2, r->stream));
      HIP_TRY(launch_kick(r->module, v, u, r->s, P, (float)RK_BETA[k], r->stream));
      HIP_TRY(launch_drift(r->module, u, v, r->s, P, (float)(dt * RK_GAMMA[k]), r->diag, r->stream));
    }
    r->lam_valid = false;
    return FG2_OK;
  }
  if(cfl) reset_diag_kernel<<<1, 1, 0, r->stream>>>(r->diag, 1);
  if(uses_step_kernel(r)) {
    HIP_TRY(launch_step(r->module, r->in[r->cur], r->in[r->cur ^ 1], r->s, P, dt,
                        step_shape_of(r), cfl, r->diag, r->stream));
    r->cur ^= 1;
    r->lam_valid = cfl != 0;
    return FG2_OK;
  }
  const FusedShape sh = shape_of(r);
  for(int k = 0; k < 3; ++k) {
    HIP_TRY(launch_substage(r->module, k, r->in[r->cur], r->in[r->cur ^ 1], r->in[2], r->s, P,
                            (float)RK_BETA[k], (float)(dt * RK_GAMMA[k]), sh, cfl, r->diag,
                            r->stream));
    r->cur ^= 1;
  }
  r->lam_valid = cfl != 0;
  return FG2_OK;
}

} // namespace

extern "C" {

const char *fg2_last_error(void) { return g_error.c_str(); }

int fg2_module_count(void) { return module_count(); }

int fg2_module(int k, const char **name, int *nvar, const char **fields)
{
  if(k < 0 || k >= module_count()) return fail(FG2_ERR_ARG, "no module %d", k);
  const ModuleInfo &m = module_info(k);
  if(name) *name = m.name;
  if(nvar) *nvar = m.nv;
  if(fields) *fields = m.fields;
  return FG2_OK;
}

int fg2_setup(fg2_run **out, const char *module, int n1, int n2, size_t *bytes)
{
  if(!out) return fail(FG2_ERR_ARG, "NULL output handle");
  *out = nullptr;
  const int m = module_id(module);
  if(m < 0) return fail(FG2_ERR_ARG, "unknown module \"%s\"", module ? module : "(null)");
  if(n1 < HALF || n2 < HALF) return fail(FG2_ERR_SHAPE, "grid %d x %d: both sides must be >= %d", n1, n2, HALF);
  fg2_run *r = new(std::nothrow) fg2_run;
  if(!r) return fail(FG2_ERR_NOMEM, "host allocation failed");
  r->module = m; r->nv = nvar_of(m); r->n1 = n1; r->n2 = n2;
  r->s = row_stride(n2, r->nv);
  for(auto &p : param_defs(m)) r->par.push_back(p.value);
  const long nf = array_floats(n1, n2, r->nv);
  const long off = interior_offset(n2, r->nv);
  hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
  for(int k = 0; k < 3 && e == hipSuccess; ++k) {
    e = hipMalloc((void **)&r->buf[k], nf * sizeof(float));
    if(e == hipSuccess) {
      r->in[k] = r->buf[k] + off;
      // -FLT_MAX everywhere, as src/setup.cpp:80-83
      const float neg = -FLT_MAX;
      unsigned int bits;
      memcpy(&bits, &neg, 4);
      e = hipMemsetD32Async((hipDeviceptr_t)r->buf[k], (int)bits, nf, r->stream);
    }
  }
  if(e == hipSuccess) e = hipMalloc((void **)&r->diag, sizeof(DeviceDiag));
  if(e == hipSuccess) e = hipMemsetAsync(r->diag, 0, sizeof(DeviceDiag), r->stream);
  if(e == hipSuccess) e = hipHostMalloc((void **)&r->host_diag, sizeof(DeviceDiag), hipHostMallocDefault);
  if(e == hipSuccess) e = hipMalloc((void **)&r->sum_partial, 4 * (size_t)n1 * sizeof(double));
  if(e == hipSuccess) e = hipMalloc((void **)&r->sum_out, 4 * sizeof(double));
  if(e == hipSuccess && module_info(m).row_table)
    e = hipMalloc((void **)&r->rowtab, 4 * (size_t)n1 * sizeof(float));
  int dev = 0;
  if(e == hipSuccess) e = hipGetDevice(&dev);
  if(e == hipSuccess) e = hipDeviceGetAttribute(&r->cus, hipDeviceAttributeMultiprocessorCount, dev);
  if(e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if(e != hipSuccess) {
    const int code = (e == hipErrorOutOfMemory) ? FG2_ERR_NOMEM : FG2_ERR_HIP;
    fg2_free(r);
    return fail(code, "setup %d x %d: %s", n1, n2, hipGetErrorString(e));
  }
  r->bytes = 3 * nf * sizeof(float);
  if(bytes) *bytes = r->bytes;
  *out = r;
  return FG2_OK;
}

void fg2_free(fg2_run *r)
{
  if(!r) return;
  if(r->stream) (void)hipStreamSynchronize(r->stream);
  for(int k = 0; k < 3; ++k) if(r->buf[k]) (void)hipFree(r->buf[k]);
  if(r->diag) (void)hipFree(r->diag);
  if(r->host_diag) (void)hipHostFree(r->host_diag);
  if(r->sum_partial) (void)hipFree(r->sum_partial);
  if(r->sum_out) (void)hipFree(r->sum_out);
  if(r->rowtab) (void)hipFree(r->rowtab);
  if(r->spec) spectral_free(r->spec);
  if(r->slab_buf) (void)hipFree(r->slab_buf);
  if(r->slab_ev) (void)hipEventDestroy(r->slab_ev);
  if(r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
}

int fg2_set_geometry(fg2_run *r, double l1, double l2, int p1, int p2, double c)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  if(!(l1 > 0) || !(l2 > 0) || !(c > 0)) return fail(FG2_ERR_ARG, "l1, l2, c must be > 0");
  // global::l1, l2, c are R (float) in the reference
  r->l1 = (float)l1; r->l2 = (float)l2; r->c = (float)c;
  r->p1 = p1 ? 1 : 0; r->p2 = p2 ? 1 : 0;
  if(r->configured) {
    r->Delta1 = (float)r->l1 / (float)r->n1;
    r->Delta2 = (float)r->l2 / (float)r->n2;
  }
  return refresh_rows(r);
}

int fg2_get_geometry(const fg2_run *r, double *out)
{
  if(!r || !out) return fail(FG2_ERR_ARG, "NULL argument");
  out[0] = r->l1; out[1] = r->l2; out[2] = r->p1; out[3] = r->p2; out[4] = r->c;
  out[5] = r->Delta1; out[6] = r->Delta2;
  return FG2_OK;
}

int fg2_config(fg2_run *r)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  // the module's settings() (src/<module>/config.h)
  Geometry g{r->l1, r->l2, r->p1, r->p2};
  module_info(r->module).settings(g, r->n1, r->n2);
  r->l1 = g.l1; r->l2 = g.l2; r->p1 = g.p1; r->p2 = g.p2;
  // Delta = l / n in R (src/config.cpp:39); a slab takes the whole grid's
  r->Delta1 = (float)r->l1 / (float)rows_of_grid(r);
  r->Delta2 = (float)r->l2 / (float)r->n2;
  r->configured = true;
  return refresh_rows(r);
}

const char *fg2_para(fg2_run *r, const char *in)
{
  if(!r || !in) { fail(FG2_ERR_ARG, "NULL argument"); return nullptr; }
  while(*in == '-') ++in;                       // src/para.cpp:31
  const char *eq = strchr(in, '=');
  if(!eq) { fail(FG2_ERR_ARG, "\"%s\" is not name=val", in); return nullptr; }
  const std::string name(in, eq - in);
  const auto &defs = param_defs(r->module);
  for(size_t k = 0; k < defs.size(); ++k)
    if(name == defs[k].name) {
      const float val = (float)atof(eq + 1);     // stored as R (src/para.cpp:33)
      r->par[k] = val;
      if(refresh_rows(r)) return nullptr;
      snprintf(r->para_buf, sizeof r->para_buf, "%s = %g", name.c_str(), (double)val);
      return r->para_buf;
    }
  fail(FG2_ERR_ARG, "module has no parameter \"%s\"", name.c_str());
  return nullptr;
}

int fg2_get_param(const fg2_run *r, const char *name, double *value)
{
  if(!r || !name || !value) return fail(FG2_ERR_ARG, "NULL argument");
  const auto &defs = param_defs(r->module);
  for(size_t k = 0; k < defs.size(); ++k)
    if(!strcmp(name, defs[k].name)) { *value = r->par[k]; return FG2_OK; }
  return fail(FG2_ERR_ARG, "module has no parameter \"%s\"", name);
}

int fg2_seed(fg2_run *r, unsigned long long seed)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  r->rng = seed;
  return FG2_OK;
}

int fg2_layout(const fg2_run *r, int *nvar, long *stride)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  if(nvar) *nvar = r->nv;
  if(stride) *stride = r->s;
  return FG2_OK;
}

float *fg2_u(fg2_run *r) { return r ? r->in[r->cur] : nullptr; }
float *fg2_v(fg2_run *r) { return r ? r->in[2] : nullptr; }
void *fg2_stream(fg2_run *r) { return r ? (void *)r->stream : nullptr; }

static int copy_interior(fg2_run *r, float *dev, float *host, bool to_dev)
{
  if(!r || !host) return fail(FG2_ERR_ARG, "NULL argument");
  const size_t hp = (size_t)r->n2 * r->nv * sizeof(float);
  const size_t dp = (size_t)r->s * sizeof(float);
  if(to_dev) HIP_TRY(hipMemcpy2DAsync(dev, dp, host, hp, hp, r->n1, hipMemcpyHostToDevice, r->stream));
  else       HIP_TRY(hipMemcpy2DAsync(host, hp, dev, dp, hp, r->n1, hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return FG2_OK;
}

int fg2_upload(fg2_run *r, const float *h)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  r->lam_valid = false;
  return copy_interior(r, r->in[r->cur], (float *)h, true);
}
int fg2_download(fg2_run *r, float *h) { return r ? copy_interior(r, r->in[r->cur], h, false) : fail(FG2_ERR_ARG, "NULL run"); }

int fg2_init(fg2_run *r, const char *name, unsigned long long seed)
{
  if(int e = check_run(r)) return e;
  const int ic = module_info(r->module).ic_of(name ? name : "");   // the module's pick() (init.h)
  const Params P = make_params(r);
  HIP_TRY(launch_ic(r->module, ic, r->in[r->cur], r->s, r->n1, r->n2, (float)r->l1, (float)r->l2,
                    P.par, seed, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  r->lam_valid = false;
  return FG2_OK;
}

// .raw snapshot header (src/io.cpp:62-67): int32 size[6], double time, R info[3].
int fg2_dump(fg2_run *r, const char *path, double time)
{
  if(!r || !path) return fail(FG2_ERR_ARG, "NULL argument");
  std::vector<float> x((size_t)r->n1 * r->n2 * r->nv);
  if(int e = copy_interior(r, r->in[r->cur], x.data(), false)) return e;
  FILE *f = fopen(path, "wb");
  if(!f) return fail(FG2_ERR_ARG, "cannot open \"%s\" for writing", path);
  const int32_t size[6] = {r->n1, r->n2, r->nv, (int32_t)sizeof(float), r->p1, r->p2};
  const float info[3] = {(float)r->l1, (float)r->l2, (float)r->c};
  bool ok = fwrite(size, sizeof size, 1, f) == 1 && fwrite(&time, sizeof time, 1, f) == 1 &&
            fwrite(info, sizeof info, 1, f) == 1 &&
            fwrite(x.data(), sizeof(float), x.size(), f) == x.size();
  ok = (fclose(f) == 0) && ok;
  return ok ? FG2_OK : fail(FG2_ERR_ARG, "short write to \"%s\"", path);
}

int fg2_load(fg2_run *r, const char *path, double *time)
{
  if(!r || !path) return fail(FG2_ERR_ARG, "NULL argument");
  FILE *f = fopen(path, "rb");
  if(!f) return fail(FG2_ERR_ARG, "cannot open \"%s\"", path);
  int32_t size[6];
  double t = 0;
  float info[3];
  if(fread(size, sizeof size, 1, f) != 1 || fread(&t, sizeof t, 1, f) != 1 ||
     fread(info, sizeof info, 1, f) != 1) {
    fclose(f);
    return fail(FG2_ERR_ARG, "\"%s\": short header", path);
  }
  if(size[0] != r->n1 || size[1] != r->n2 || size[2] != r->nv || size[3] != (int32_t)sizeof(float)) {
    fclose(f);
    return fail(FG2_ERR_SHAPE, "input data is not compatible with the setup "
                "(file %d x %d x %d of %d-byte words, run %d x %d x %d)",
                size[0], size[1], size[2], size[3], r->n1, r->n2, r->nv);
  }
  std::vector<float> x((size_t)r->n1 * r->n2 * r->nv);
  const bool ok = fread(x.data(), sizeof(float), x.size(), f) == x.size();
  fclose(f);
  if(!ok) return fail(FG2_ERR_ARG, "\"%s\": short data", path);
  if(!(info[0] > 0) || !(info[1] > 0) || !(info[2] > 0))
    return fail(FG2_ERR_ARG, "\"%s\": l1, l2, c must be > 0", path);
  if(int e = fg2_upload(r, x.data())) return e;
  fg2_set_geometry(r, info[0], info[1], size[4], size[5], info[2]);   // src/io.cpp:49-53
  if(time) *time = t;
  return FG2_OK;
}
int fg2_upload_v(fg2_run *r, const float *h) { return r ? copy_interior(r, r->in[2], (float *)h, true) : fail(FG2_ERR_ARG, "NULL run"); }
int fg2_download_v(fg2_run *r, float *h) { return r ? copy_interior(r, r->in[2], h, false) : fail(FG2_ERR_ARG, "NULL run"); }

int fg2_euler(fg2_run *r, float *u, float dt)
{
  if(int e = check_run(r)) return e;
  if(!u) return fail(FG2_ERR_ARG, "NULL u");
  randomize(r);
  if(!first_order_active(r)) return FG2_OK;
  HIP_TRY(launch_euler(r->module, u, r->s, make_params(r), dt, r->stream));
  return FG2_OK;
}

int fg2_bcond(fg2_run *r, float *u, int p1, int p2)
{
  if(int e = check_run(r)) return e;
  if(!u) return fail(FG2_ERR_ARG, "NULL u");
  HIP_TRY(launch_bcond(r->module, u, r->s, make_params(r), p1 ? 1 : 0, p2 ? 1 : 0, r->stream));
  return FG2_OK;
}

int fg2_kick(fg2_run *r, float *v, const float *u, float t, float beta)
{
  (void)t;
  if(int e = check_run(r)) return e;
  if(!u || !v) return fail(FG2_ERR_ARG, "NULL array");
  HIP_TRY(launch_kick(r->module, v, u, r->s, make_params(r), beta, r->stream));
  return FG2_OK;
}

int fg2_drift(fg2_run *r, float *u, const float *v, float dt)
{
  if(int e = check_run(r)) return e;
  if(!u || !v) return fail(FG2_ERR_ARG, "NULL array");
  HIP_TRY(launch_drift(r->module, u, v, r->s, make_params(r), dt, r->diag, r->stream));
  return FG2_OK;
}

double fg2_getdt(const fg2_run *r)
{
  if(!r) return 0.0;
  // (E)c / (n1 / l1 + n2 / l2) with l, c in R (src/getdt.cpp:26)
  const float a = (float)rows_of_grid(r) / (float)r->l1, b = (float)r->n2 / (float)r->l2;
  return (double)(float)r->c / (double)(a + b);
}

int fg2_getdt_cfl(fg2_run *r, double *dt)
{
  if(int e = check_run(r)) return e;
  if(!dt) return fail(FG2_ERR_ARG, "NULL dt");
  if(!r->lam_valid) {
    reset_diag_kernel<<<1, 1, 0, r->stream>>>(r->diag, 1);
    HIP_TRY(launch_cfl(r->module, r->in[r->cur], r->s, make_params(r), r->diag, r->stream));
    r->lam_valid = true;
  }
  HIP_TRY(hipMemcpyAsync(r->host_diag, r->diag, sizeof(DeviceDiag), hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  *dt = dt_from_rates(r, bits_to_double(r->host_diag->lam_adv_bits),
                      bits_to_double(r->host_diag->lam_diff_bits));
  return FG2_OK;
}

int fg2_set_path(fg2_run *r, int path)
{
  if(!r || (path != FG2_PATH_FUSED && path != FG2_PATH_PLAIN && path != FG2_PATH_SUBSTAGE))
    return fail(FG2_ERR_ARG, "bad path");
  r->path = path;
  return FG2_OK;
}

int fg2_set_dt_mode(fg2_run *r, int mode)
{
  if(!r || (mode != FG2_DT_CLOSED && mode != FG2_DT_CFL)) return fail(FG2_ERR_ARG, "bad dt mode");
  r->dt_mode = mode;
  return FG2_OK;
}

int fg2_set_tile(fg2_run *r, int tile_h, int waves)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  if(waves != 1 && waves != 2 && waves != 4) return fail(FG2_ERR_SHAPE, "waves must be 1, 2 or 4");
  if(tile_h < 0) return fail(FG2_ERR_SHAPE, "tile_h must be >= 0");
  r->tile_auto = tile_h == 0;
  r->shape.tile_h = tile_h;
  r->shape.waves = waves;
  return FG2_OK;
}

int fg2_set_kernel(fg2_run *r, int kernel)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  if(kernel != FG2_KERNEL_AUTO && kernel != FG2_KERNEL_ONE_PASS && kernel != FG2_KERNEL_SUBSTAGE)
    return fail(FG2_ERR_ARG, "bad kernel choice %d", kernel);
  if(kernel == FG2_KERNEL_ONE_PASS && !module_info(r->module).one_pass)
    return fail(FG2_ERR_ARG, "this module has no one-pass step kernel");
  r->kernel_pick = kernel;
  return FG2_OK;
}

int fg2_get_kernel(const fg2_run *r, int *kernel)
{
  if(!r || !kernel) return fail(FG2_ERR_ARG, "NULL argument");
  *kernel = uses_step_kernel(r) ? FG2_KERNEL_ONE_PASS
                                : r->path == FG2_PATH_PLAIN ? FG2_KERNEL_AUTO : FG2_KERNEL_SUBSTAGE;
  return FG2_OK;
}

int fg2_get_tile(const fg2_run *r, int *tile_h, int *waves)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  const FusedShape sh = uses_step_kernel(r) ? step_shape_of(r) : shape_of(r);
  if(tile_h) *tile_h = sh.tile_h;
  if(waves) *waves = sh.waves;
  return FG2_OK;
}

// ---------------------------------------------------------------------------
// Autotuner (replaces the device-property heuristic of src/setup.cpp:39-58):
// times the run's own kernel on its own state for candidate shapes and keeps
// the fastest.  The timed launches read the state and write only the other
// state buffer and v, so the state is unchanged.  Results are cached per
// (module, n1, n2, kernel) for the process.
// ---------------------------------------------------------------------------
namespace {

struct TuneKey {
  int module, n1, n2, step;
  bool operator<(const TuneKey &o) const
  { return std::tie(module, n1, n2, step) < std::tie(o.module, o.n1, o.n2, o.step); }
};
struct TunePick {
  FusedShape sh{0, 0};
  bool step = false;
};
std::map<TuneKey, TunePick> g_tune_cache;
std::mutex g_tune_mutex;

// Time of one step with shape `sh`: the one-pass kernel, or the three
// substage kernels each reading the state and writing the other state buffer
// and v (v is dead at the start of a step), so the state is unchanged.  The
// least of three timed repetitions after one warm-up.
int time_launch(fg2_run *r, const FusedShape &sh, bool step, float *ms)
{
  const Params P = make_params(r);
  const double dt = 1e-7;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const SlabHalo H = slab_halo_of(r);
  auto once = [&]() -> hipError_t {
    if(step) return launch_step(r->module, r->in[r->cur], r->in[r->cur ^ 1], r->s, P, dt, sh, 0,
                                r->diag, r->stream, r->slab_n1g ? &H : nullptr);
    for(int k = 0; k < 3; ++k) {
      const hipError_t e = launch_substage(r->module, k, r->in[r->cur], r->in[r->cur ^ 1], r->in[2],
                                           r->s, P, (float)RK_BETA[k], (float)(dt * RK_GAMMA[k]),
                                           sh, 0, r->diag, r->stream);
      if(e != hipSuccess) return e;
    }
    return hipSuccess;
  };
  HIP_TRY(once());
  *ms = 1e30f;
  for(int rep = 0; rep < 3; ++rep) {
    HIP_TRY(hipEventRecord(e0, r->stream));
    HIP_TRY(once());
    HIP_TRY(hipEventRecord(e1, r->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    *ms = std::min(*ms, t);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return FG2_OK;
}

} // namespace

extern "C" int fg2_autotune(fg2_run *r, int *tile_h, int *waves)
{
  if(int e = check_run(r)) return e;
  if(r->path == FG2_PATH_PLAIN) return fail(FG2_ERR_ARG, "the plain path has no tile shape");
  // a latency-bound grid of a one-pass module with the automatic kernel
  // choice times both kernels
  const bool one_pass = module_info(r->module).one_pass;
  const bool both = one_pass && r->kernel_pick == FG2_KERNEL_AUTO && latency_bound(r);
  const bool step = uses_step_kernel(r);
  const TuneKey key{r->module, r->n1, r->n2, both ? 3 : step ? (r->p1 && r->p2 ? 1 : 2) : 0};
  TunePick best;
  {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    auto it = g_tune_cache.find(key);
    if(it != g_tune_cache.end()) best = it->second;
  }
  if(best.sh.tile_h == 0) {
    // the NaN flag of steps queued before the tuning is kept
    HIP_TRY(hipMemcpyAsync(r->host_diag, r->diag, sizeof(DeviceDiag), hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    const unsigned int flags = r->host_diag->flags;
    std::vector<TunePick> cand;
    if(both) cand.push_back({step_shape_of(r), true});
    cand.push_back({step ? step_shape_of(r) : shape_of(r), step});
    const int hs_step[] = {32, 48, 64, 96, 112, 128, 160, 224};
    const int hs_sub[] = {1, 2, 4, 8, 16, 24, 32, 48, 64, 128};
    if(step) {
      for(int w : {4, 2})
        for(int h : hs_step) if(h <= r->n1) cand.push_back({{h, w}, true});
    } else {
      for(int w : {4, 2, 1})
        for(int h : hs_sub) if(h <= r->n1) cand.push_back({{h, w}, false});
    }
    float best_ms = 1e30f;
    for(const TunePick &c : cand) {
      float ms;
      if(int e = time_launch(r, c.sh, c.step, &ms)) return e;
      if(ms < best_ms) { best_ms = ms; best = c; }
    }
    restore_diag_kernel<<<1, 1, 0, r->stream>>>(r->diag, flags);
    HIP_TRY(hipStreamSynchronize(r->stream));
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    g_tune_cache[key] = best;
  }
  r->shape = best.sh;
  r->tile_auto = false;
  if(one_pass) r->kernel_pick = best.step ? FG2_KERNEL_ONE_PASS : FG2_KERNEL_SUBSTAGE;
  r->lam_valid = false;
  if(tile_h) *tile_h = best.sh.tile_h;
  if(waves) *waves = best.sh.waves;
  return FG2_OK;
}

int fg2_step_async(fg2_run *r, double t, double dt)
{
  if(int e = check_run(r)) return e;
  if(r->slab_n1g) return fail(FG2_ERR_ARG, "a slab steps with fg2_slab_step");
  randomize(r);
  return substages(r, t, dt, r->dt_mode == FG2_DT_CFL);
}

int fg2_substage(fg2_run *r, int stage, double dt)
{
  if(int e = check_run(r)) return e;
  if(stage < 0 || stage > 2) return fail(FG2_ERR_ARG, "stage must be 0, 1 or 2");
  if(r->slab_n1g) return fail(FG2_ERR_ARG, "a slab steps with fg2_slab_step");
  HIP_TRY(launch_substage(r->module, stage, r->in[r->cur], r->in[r->cur ^ 1], r->in[2], r->s,
                          make_params(r), (float)RK_BETA[stage], (float)(dt * RK_GAMMA[stage]),
                          shape_of(r), 0, r->diag, r->stream));
  r->cur ^= 1;
  r->lam_valid = false;
  return FG2_OK;
}

int fg2_check(fg2_run *r)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  HIP_TRY(hipMemcpyAsync(r->host_diag, r->diag, sizeof(DeviceDiag), hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  if(r->host_diag->flags & 1u) {
    reset_diag_kernel<<<1, 1, 0, r->stream>>>(r->diag, 0);
    (void)hipStreamSynchronize(r->stream);
    r->lam_valid = false;
    return fail(FG2_ERR_NAN, "the state contains NaN");
  }
  return FG2_OK;
}

int fg2_step(fg2_run *r, double t, double dt)
{
  if(int e = fg2_step_async(r, t, dt)) return e;
  return fg2_check(r);
}

// -- slabs of a grid cut along dim 1 -----------------------------------------

int fg2_slab(fg2_run *r, int row0, int n1_global)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  if(n1_global == 0) {
    r->slab_row0 = r->slab_n1g = 0;
    return r->configured ? fg2_config(r) : FG2_OK;
  }
  if(!module_info(r->module).one_pass || !r->p1)
    return fail(FG2_ERR_ARG, "%s grid not periodic along dim 1 or stepped by substages: no slabs",
                module_info(r->module).name);
  if(r->n1 < SLAB_HALO)
    return fail(FG2_ERR_SHAPE, "a slab needs at least %d rows, this run has %d", SLAB_HALO, r->n1);
  if(n1_global < r->n1 || row0 < 0 || row0 >= n1_global)
    return fail(FG2_ERR_SHAPE, "slab rows [%d, %d) do not fit a grid of %d rows", row0,
                row0 + r->n1, n1_global);
  if(!r->slab_buf) {
    const size_t bytes = 2 * (size_t)SLAB_HALO * r->s * sizeof(float);
    HIP_TRY(hipMalloc((void **)&r->slab_buf, bytes));
    HIP_TRY(hipMemsetAsync(r->slab_buf, 0, bytes, r->stream));
    HIP_TRY(hipEventCreateWithFlags(&r->slab_ev, hipEventDisableTiming));
  }
  r->slab_row0 = row0;
  r->slab_n1g = n1_global;
  r->lam_valid = false;
  return r->configured ? fg2_config(r) : FG2_OK;
}

int fg2_slab_halo(fg2_run *r, float **recv_lo, float **recv_hi, const float **send_first,
                  const float **send_last, size_t *bytes)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  if(!r->slab_n1g) return fail(FG2_ERR_ARG, "not a slab: call fg2_slab first");
  const long rows = (long)SLAB_HALO * r->s;
  // a row's storage starts ROW_PAD floats before its interior cell 0
  if(recv_lo) *recv_lo = r->slab_buf;
  if(recv_hi) *recv_hi = r->slab_buf + rows;
  if(send_first) *send_first = r->in[r->cur] - ROW_PAD;
  if(send_last) *send_last = r->in[r->cur] + (long)(r->n1 - SLAB_HALO) * r->s - ROW_PAD;
  if(bytes) *bytes = (size_t)rows * sizeof(float);
  return FG2_OK;
}

int fg2_slab_exchange(fg2_run *const *runs, int nslabs)
{
  if(!runs || nslabs < 1) return fail(FG2_ERR_ARG, "bad slab list");
  const fg2_run *a = runs[0];
  if(!a || !a->slab_n1g) return fail(FG2_ERR_ARG, "slab 0 is not a slab");
  long rows = 0;
  for(int k = 0; k < nslabs; ++k) {
    const fg2_run *r = runs[k], *nx = runs[(k + 1) % nslabs];
    if(!r || !r->slab_n1g) return fail(FG2_ERR_ARG, "slab %d is not a slab", k);
    if(r->module != a->module || r->n2 != a->n2 || r->slab_n1g != a->slab_n1g)
      return fail(FG2_ERR_ARG, "slab %d: module, n2 or grid rows differ from slab 0", k);
    if((r->slab_row0 + r->n1) % r->slab_n1g != nx->slab_row0)
      return fail(FG2_ERR_SHAPE, "slab %d ends at row %d but the next one starts at %d", k,
                  r->slab_row0 + r->n1, nx->slab_row0);
    rows += r->n1;
  }
  if(rows != a->slab_n1g)
    return fail(FG2_ERR_SHAPE, "the slabs hold %ld rows, the grid %d", rows, a->slab_n1g);
  for(int k = 0; k < nslabs; ++k) HIP_TRY(hipEventRecord(runs[k]->slab_ev, runs[k]->stream));
  // Each slab pulls its halo rows on its own stream, after its own last step
  // (which read the old halos) and after the neighbour's last step.
  for(int k = 0; k < nslabs; ++k) {
    fg2_run *r = runs[k];
    const fg2_run *pv = runs[(k + nslabs - 1) % nslabs], *nx = runs[(k + 1) % nslabs];
    float *lo = nullptr, *hi = nullptr;
    const float *pv_last = nullptr, *nx_first = nullptr;
    size_t bytes = 0;
    fg2_slab_halo(r, &lo, &hi, nullptr, nullptr, &bytes);
    fg2_slab_halo(const_cast<fg2_run *>(pv), nullptr, nullptr, nullptr, &pv_last, nullptr);
    fg2_slab_halo(const_cast<fg2_run *>(nx), nullptr, nullptr, &nx_first, nullptr, nullptr);
    HIP_TRY(hipStreamWaitEvent(r->stream, pv->slab_ev, 0));
    HIP_TRY(hipMemcpyAsync(lo, pv_last, bytes, hipMemcpyDeviceToDevice, r->stream));
    HIP_TRY(hipStreamWaitEvent(r->stream, nx->slab_ev, 0));
    HIP_TRY(hipMemcpyAsync(hi, nx_first, bytes, hipMemcpyDeviceToDevice, r->stream));
  }
  return FG2_OK;
}

int fg2_slab_step(fg2_run *r, double dt)
{
  if(int e = check_run(r)) return e;
  if(!r->slab_n1g) return fail(FG2_ERR_ARG, "not a slab: call fg2_slab first");
  if(r->path != FG2_PATH_FUSED || !r->p1 || !r->p2)
    return fail(FG2_ERR_ARG, "a slab steps with the one-pass kernel: fused path, periodic grid");
  if(first_order_active(r))
    return fail(FG2_ERR_ARG, "slabs with an active first_order are not supported");
  randomize(r);
  const int cfl = r->dt_mode == FG2_DT_CFL;
  if(cfl) reset_diag_kernel<<<1, 1, 0, r->stream>>>(r->diag, 1);
  const SlabHalo H = slab_halo_of(r);
  HIP_TRY(launch_step(r->module, r->in[r->cur], r->in[r->cur ^ 1], r->s, make_params(r), dt,
                      step_shape_of(r), cfl, r->diag, r->stream, &H));
  r->cur ^= 1;
  r->lam_valid = cfl != 0;
  return FG2_OK;
}

int fg2_solve(fg2_run *r, double *t, double t_final, int *steps)
{
  if(int e = check_run(r)) return e;
  if(!t) return fail(FG2_ERR_ARG, "NULL t");
  int m = 0;
  double tt = *t;
  int status = FG2_OK;
  while(tt < t_final) {                                       // src/solve.cpp:37-49
    double dt0;
    if(r->dt_mode == FG2_DT_CFL) {
      if(int e = fg2_getdt_cfl(r, &dt0)) { status = e; break; }
    } else dt0 = fg2_getdt(r);
    const double dt = std::min(t_final - tt, dt0);
    if(!(dt > 0)) { status = fail(FG2_ERR_ARG, "non-positive time step %g", dt); break; }
    ++m;
    if(int e = fg2_step(r, tt, dt)) { status = e; break; }
    tt = std::min(tt + dt, t_final);
  }
  *t = tt;
  if(steps) *steps = m;
  return status;
}

int fg2_sums(fg2_run *r, double *out)
{
  if(int e = check_run(r)) return e;
  if(!out) return fail(FG2_ERR_ARG, "NULL out");
  HIP_TRY(launch_sums(r->module, r->in[r->cur], r->s, make_params(r), r->sum_partial, r->sum_out, r->stream));
  HIP_TRY(hipMemcpyAsync(out, r->sum_out, 4 * sizeof(double), hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return FG2_OK;
}

// Spectral diagnostics of tools/spec.pro, tools/vis.pro (see spectra.hip).
static int spectral_status(int code)
{
  switch(code) {
  case SPECTRA_OK: return FG2_OK;
  case SPECTRA_FFT: return fail(FG2_ERR_HIP, "hipFFT plan or execution failed");
  case SPECTRA_NOMEM: return fail(FG2_ERR_NOMEM, "spectral workspace allocation failed");
  case SPECTRA_SHAPE: return fail(FG2_ERR_SHAPE, "too many spectral shells for one workgroup's LDS");
  default: return fail(FG2_ERR_HIP, "spectral kernels: %s", hipGetErrorString(hipGetLastError()));
  }
}

static int spectral_work(fg2_run *r)
{
  if(r->spec) return FG2_OK;
  return spectral_status(spectral_create(&r->spec, r->n1, r->n2));
}

int fg2_spectrum(fg2_run *r, double *E, double *Z, unsigned long long *counts, int *nbins)
{
  if(!r || !nbins) return fail(FG2_ERR_ARG, "NULL argument");
  const int need = spectral_bins(r->n1, r->n2);
  if((E || Z || counts) && *nbins < need) {
    *nbins = need;
    return fail(FG2_ERR_ARG, "spectrum has %d shells, arrays hold fewer", need);
  }
  *nbins = need;
  if(!E && !Z && !counts) return FG2_OK;
  if(int e = spectral_work(r)) return e;
  return spectral_status(spectral_shells(r->spec, r->in[r->cur], r->s, r->nv, E, Z, counts,
                                         r->stream));
}

int fg2_vorticity(fg2_run *r, float *host)
{
  if(!r || !host) return fail(FG2_ERR_ARG, "NULL argument");
  if(int e = spectral_work(r)) return e;
  return spectral_status(spectral_vorticity(r->spec, r->in[r->cur], r->s, r->nv, host, r->stream));
}

int fg2_perf_model(const fg2_run *r, double *flops, double *bits, double *alg)
{
  if(!r) return fail(FG2_ERR_ARG, "NULL run");
  const double n = (double)r->n1 * r->n2;
  const double m1 = r->n1 + ORDER, m2 = r->n2 + ORDER;
  // eqns_flop of the module (src/*/config.h), src/config.cpp:33
  if(flops) *flops = 3.0 * n * (module_info(r->module).flop + r->nv * 2.0);
  if(bits) *bits = 3.0 * (m1 * m2 + n * 5.0 + (m1 + m2) * 2.0 * ORDER) * r->nv * sizeof(float) * 8;
  if(alg) *alg = 10.0 * r->nv * sizeof(float);
  return FG2_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------
struct fg2_batch {
  int module = 0, nruns = 0, waves = 1, grid = 0, total = 0;
  bool step = false;          // one-pass step kernel (all runs periodic) or substages
  int cus = 256;
  std::vector<fg2_run *> runs;
  std::vector<int> tile_h;
  BatchRun *dev[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};  // [stage][parity]
  std::vector<BatchRun> host[3][2];
  int *counters = nullptr;   // one per stage
  hipStream_t stream = nullptr;
  int parity = 0;
  int fo_blocks = 0;         // blocks of the batched Euler pre-step (0: no run has one)
  bool clocked = false;      // each step takes its dt from the runs' device clocks
  bool any_cfl = false;      // some record reduces CFL rates (clocked CFL runs)
  // hipGraph of the last fg2_batch_graph_run / fg2_batch_solve, kept for
  // replays while its launch shape (steps, waves, grid, tiles, pre-steps,
  // clock) is unchanged; the launch records it reads are refilled before
  // every call
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_steps = -1, g_waves = 0, g_grid = 0, g_total = 0, g_fo = 0;
  bool g_clocked = false, g_cfl = false;
  // fg2_batch_solve: each run's clock, gathered after every replay
  ClockOut *clk_dev = nullptr, *clk_host = nullptr;
};

namespace {

__global__ void zero_ints_kernel(int *p, int n)
{
  if((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void gather_clock_kernel(const BatchRun *set0, const int nruns, ClockOut *out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= nruns) return;
  const DeviceDiag *d = set0[i].diag;
  out[i] = ClockOut{d->t, d->dt, d->steps, d->rng, d->flags};
}

// One run's launch record: the one-pass step (stage < 0) or substage
// `stage`, reading the state buffer in[ub] and writing in[ub ^ 1] (and v).
// dt: the host's step (NULL: the device clock writes it before each launch,
// and then a run in CFL mode also reduces its new state's rates).  The
// caller sets tile_begin and fo_begin.
BatchRun make_record(fg2_run *r, int ub, int stage, int tile_h, int waves, const double *dt,
                     bool clocked)
{
  BatchRun R{};
  R.uin = r->in[ub];
  R.uout = r->in[ub ^ 1];
  R.v = r->in[2];
  R.s = r->s;
  R.P = make_params(r);
  R.tile_h = tile_h;
  if(stage < 0) {
    R.ntiles = (int)step_launch_tiles(r->n1, r->n2, tile_h, waves, &R.nstrips);
    for(int k = 0; k < 3; ++k) R.g[k] = dt ? (float)(*dt * RK_GAMMA[k]) : 0.0f;
  } else {
    R.nstrips = (r->n2 + STRIP - 1) / STRIP;
    R.ntiles = R.nstrips * ((r->n1 + tile_h - 1) / tile_h);
    R.beta = (float)RK_BETA[stage];
    R.gdt = dt ? (float)(*dt * RK_GAMMA[stage]) : 0.0f;
  }
  R.cfl = clocked && r->dt_mode == FG2_DT_CFL;
  R.fo = first_order_active(r);
  R.edt = dt ? (float)*dt : 0.0f;
  R.diag = r->diag;
  return R;
}

// Running sums of the records' tiles and Euler blocks; returns the tiles,
// *fo_blocks the Euler blocks.
int number_records(BatchRun *R, int n, int *fo_blocks)
{
  int tiles = 0, blocks = 0;
  for(int i = 0; i < n; ++i) {
    R[i].tile_begin = tiles;
    tiles += R[i].ntiles;
    R[i].fo_begin = blocks;
    if(R[i].fo) blocks += (int)(((long)R[i].P.n1 * R[i].P.n2 + 255) / 256);
  }
  if(fo_blocks) *fo_blocks = blocks;
  return tiles;
}

// Launch records of one batched step of every run, for both parities of the
// state buffers (parity p: the step starting from in[cur ^ p]).  timing:
// records for the autotuner that leave every run's state unchanged -- each
// substage reads the state u[cur] and writes only the other state buffer
// and v (which is dead at the start of a step); the one-pass step already
// writes only the other state buffer.  dts = NULL: the device clock.
int batch_fill(fg2_batch *b, const double *dts, bool timing = false)
{
  const int nstage = b->step ? 1 : 3;
  for(int par = 0; par < 2; ++par)
    for(int k = 0; k < nstage; ++k) {
      auto &H = b->host[k][par];
      H.resize(b->nruns);
      for(int i = 0; i < b->nruns; ++i) {
        fg2_run *r = b->runs[i];
        // stage k reads u[cur ^ (k & 1)] and writes u[cur ^ 1 ^ (k & 1)]
        const int flip = (b->step || timing) ? 0 : (k & 1);
        H[i] = make_record(r, (r->cur ^ par) ^ flip, b->step ? -1 : k, b->tile_h[i], b->waves,
                           dts ? dts + i : nullptr, b->clocked);
      }
      b->total = number_records(H.data(), b->nruns, &b->fo_blocks);
      HIP_TRY(hipMemcpyAsync(b->dev[k][par], H.data(), b->nruns * sizeof(BatchRun),
                             hipMemcpyHostToDevice, b->stream));
    }
  b->any_cfl = false;
  for(const BatchRun &R : b->host[0][0]) b->any_cfl = b->any_cfl || R.cfl;
  return FG2_OK;
}

// One batched step on the batch's stream from the records of parity par:
// the device clock (clocked batches), the Euler pre-step of the runs that
// have one, then the step or its three substages.  timing: the step alone.
int batch_launch(fg2_batch *b, int par, bool timing = false)
{
  if(b->clocked && !timing)
    HIP_TRY(launch_clock(b->module, b->dev[0][par], b->dev[1][par], b->dev[2][par],
                         b->step ? 1 : 3, b->step, b->nruns, b->stream));
  if(b->fo_blocks && !timing)
    HIP_TRY(launch_batch_euler(b->module, b->dev[0][par], b->nruns, b->fo_blocks, b->stream));
  zero_ints_kernel<<<1, 64, 0, b->stream>>>(b->counters, 3);
  if(b->step) {
    // the CFL variant of the kernel has its own occupancy, so its own grid
    const bool cfl = b->any_cfl && !timing;
    HIP_TRY(launch_batch_step(b->module, b->dev[0][par], b->nruns, b->total, b->waves,
                              cfl ? batch_grid(b->module, true, b->waves, b->cus, true) : b->grid,
                              b->counters, cfl, b->stream));
    return FG2_OK;
  }
  for(int k = 0; k < 3; ++k)
    HIP_TRY(launch_batch_substage(b->module, k, b->dev[k][par], b->nruns, b->total, b->waves,
                                  b->grid, b->counters + k, b->stream));
  return FG2_OK;
}

// Capture `steps` batched steps (parity alternating from 0) into the
// batch's graph, or keep the one captured last when its shape is the same.
int batch_graph(fg2_batch *b, int steps)
{
  const bool reuse = b->gexec && b->g_steps == steps && b->g_waves == b->waves &&
                     b->g_grid == b->grid && b->g_total == b->total &&
                     b->g_fo == b->fo_blocks && b->g_clocked == b->clocked &&
                     b->g_cfl == b->any_cfl;
  if(reuse) return FG2_OK;
  if(b->gexec) { HIP_TRY(hipGraphExecDestroy(b->gexec)); b->gexec = nullptr; }
  if(b->graph) { HIP_TRY(hipGraphDestroy(b->graph)); b->graph = nullptr; }
  b->g_steps = -1;
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int err = FG2_OK;
  for(int k = 0; k < steps && !err; ++k) err = batch_launch(b, k & 1);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(b->stream, &g);
  if(err) { if(g) (void)hipGraphDestroy(g); return err; }
  if(ce != hipSuccess) return fail(FG2_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
  b->graph = g;
  HIP_TRY(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  b->g_steps = steps; b->g_waves = b->waves; b->g_grid = b->grid; b->g_total = b->total;
  b->g_fo = b->fo_blocks; b->g_clocked = b->clocked; b->g_cfl = b->any_cfl;
  return FG2_OK;
}

// Order the batch's stream after every run's queued work (uploads).
int batch_join_runs(fg2_batch *b)
{
  HIP_TRY(hipStreamSynchronize(b->stream));
  for(fg2_run *r : b->runs) HIP_TRY(hipStreamSynchronize(r->stream));
  return FG2_OK;
}

// Set every run's device clock (t[i] .. t_final[i], or with t = NULL a
// fixed step dts[i] and no end) and, for runs in CFL mode, the rates of the
// current state; on the batch's stream, which the caller then drains.
int batch_set_clocks(fg2_batch *b, const double *t, const double *t_final, const double *dts)
{
  for(int i = 0; i < b->nruns; ++i) {
    fg2_run *r = b->runs[i];
    const bool cfl = t && r->dt_mode == FG2_DT_CFL;
    const ClockInit k = t ? clock_of(r, t[i], t_final[i], cfl ? 0.0 : fg2_getdt(r))
                          : clock_of(r, 0.0, INFINITY, dts[i]);
    set_clock_kernel<<<1, 1, 0, b->stream>>>(r->diag, k, 0);
    if(cfl && !r->lam_valid) {
      reset_diag_kernel<<<1, 1, 0, b->stream>>>(r->diag, 1);
      HIP_TRY(launch_cfl(r->module, r->in[r->cur], r->s, make_params(r), r->diag, b->stream));
    }
  }
  HIP_TRY(hipGetLastError());
  return FG2_OK;
}

} // namespace

extern "C" {

int fg2_batch_create(fg2_batch **out, fg2_run *const *runs, int nruns, int waves, int tile_h)
{
  if(!out || !runs || nruns < 1) return fail(FG2_ERR_ARG, "bad batch arguments");
  *out = nullptr;
  if(waves != 1 && waves != 2 && waves != 4) return fail(FG2_ERR_SHAPE, "waves must be 1, 2 or 4");
  fg2_batch *b = new(std::nothrow) fg2_batch;
  if(!b) return fail(FG2_ERR_NOMEM, "host allocation failed");
  b->module = runs[0]->module;
  b->nruns = nruns;
  b->waves = waves;
  b->step = true;
  for(int i = 0; i < nruns; ++i) {
    if(int e = check_run(runs[i])) { delete b; return e; }
    if(runs[i]->module != b->module) { delete b; return fail(FG2_ERR_ARG, "a batch holds runs of one module"); }
    if(runs[i]->slab_n1g) { delete b; return fail(FG2_ERR_ARG, "a batch holds whole grids, not slabs"); }
    b->step = b->step && runs[i]->p1 && runs[i]->p2 && module_info(runs[i]->module).one_pass;
  }
  for(int i = 0; i < nruns; ++i) {
    b->runs.push_back(runs[i]);
    const int auto_h = b->step ? 64 : shape_of(runs[i]).tile_h;
    b->tile_h.push_back(std::min(tile_h > 0 ? tile_h : auto_h, runs[i]->n1));
  }
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  for(int k = 0; k < 3 && e == hipSuccess; ++k)
    for(int p = 0; p < 2 && e == hipSuccess; ++p)
      e = hipMalloc((void **)&b->dev[k][p], nruns * sizeof(BatchRun));
  if(e == hipSuccess) e = hipMalloc((void **)&b->counters, 4 * sizeof(int));
  int dev = 0, cus = 256;
  if(e == hipSuccess) e = hipGetDevice(&dev);
  if(e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if(e != hipSuccess) { fg2_batch_free(b); return fail(FG2_ERR_HIP, "batch setup: %s", hipGetErrorString(e)); }
  b->cus = cus;
  b->grid = batch_grid(b->module, b->step, waves, cus);
  *out = b;
  return FG2_OK;
}

void fg2_batch_free(fg2_batch *b)
{
  if(!b) return;
  if(b->stream) (void)hipStreamSynchronize(b->stream);
  for(int k = 0; k < 3; ++k)
    for(int p = 0; p < 2; ++p) if(b->dev[k][p]) (void)hipFree(b->dev[k][p]);
  if(b->counters) (void)hipFree(b->counters);
  if(b->clk_dev) (void)hipFree(b->clk_dev);
  if(b->clk_host) (void)hipHostFree(b->clk_host);
  if(b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if(b->graph) (void)hipGraphDestroy(b->graph);
  if(b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

void *fg2_batch_stream(fg2_batch *b) { return b ? (void *)b->stream : nullptr; }

// Time one batched step (the records write every run's other state buffer,
// which the next real step overwrites) for candidate (tile_h, waves) and
// keep the fastest; the same tile height is used for every run.
int fg2_batch_autotune(fg2_batch *b, const double *dts, int *tile_h, int *waves)
{
  if(!b || !dts) return fail(FG2_ERR_ARG, "NULL argument");
  // the NaN flags of steps queued before the tuning are kept
  std::vector<unsigned int> flags;
  for(fg2_run *r : b->runs) {
    HIP_TRY(hipMemcpyAsync(r->host_diag, r->diag, sizeof(DeviceDiag), hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    flags.push_back(r->host_diag->flags);
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->clocked = false;
  // A candidate is a tile height h and a width; each is timed twice: every
  // run cut into h-row tiles (the last one shorter), and every run cut into
  // the same number of tiles of equal height (at most h rows), which keeps
  // short last tiles from idling their slots for most of the launch.
  std::vector<FusedShape> cand;
  // short tiles serve batches that fill under one round (latency bound)
  const int hs[] = {1, 2, 4, 8, 16, 32, 48, 64, 96, 128};
  for(int w : {4, 2, 1}) {
    for(int h : hs) cand.push_back({h, w});
    // and the two heights the rounds model ranks first for this shape: the
    // batch's tiles (every run's strips x chunks) over the resident slots
    const int out = b->step ? 64 * w - 6 * HALF : STRIP;
    const long slots = (long)batch_grid(b->module, b->step, w, b->cus) * (b->step ? 1 : w);
    const long extra = b->step ? 6 : 3;
    long c1 = -1, c2 = -1, h1 = 0, h2 = 0;
    for(long h = 1; h <= 256; ++h) {
      long tiles = 0;
      for(fg2_run *r : b->runs)
        tiles += b->step ? step_launch_tiles(r->n1, r->n2, (int)std::min<long>(h, r->n1), w, nullptr)
                         : (long)((r->n2 + out - 1) / out) * ((r->n1 + h - 1) / h);
      const long cost = ((tiles + slots - 1) / slots) * (h + extra);
      if(c1 < 0 || cost < c1) { c2 = c1; h2 = h1; c1 = cost; h1 = h; }
      else if(c2 < 0 || cost < c2) { c2 = cost; h2 = h; }
    }
    for(long h : {h1, h2})
      if(h > 0 && std::find_if(cand.begin(), cand.end(), [&](const FusedShape &c) {
           return c.tile_h == h && c.waves == w; }) == cand.end())
        cand.push_back({(int)h, w});
  }
  auto height = [&](int h, bool even, int i) {
    const int n1 = b->runs[i]->n1;
    if(!even) return std::min(h, n1);
    const int chunks = (n1 + h - 1) / h;
    return (n1 + chunks - 1) / chunks;
  };
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  float best_ms = 1e30f;
  FusedShape best{0, 0};
  bool best_even = false;
  for(const FusedShape &c : cand)
  for(bool even : {false, true}) {
    b->waves = c.waves;
    b->grid = batch_grid(b->module, b->step, c.waves, b->cus);
    for(int i = 0; i < b->nruns; ++i) b->tile_h[i] = height(c.tile_h, even, i);
    HIP_TRY(hipStreamSynchronize(b->stream));
    if(int e = batch_fill(b, dts, true)) return e;
    if(int e = batch_launch(b, 0, true)) return e;
    float ms = 1e30f;                       // the least of three timed launches
    for(int rep = 0; rep < 3; ++rep) {
      HIP_TRY(hipEventRecord(e0, b->stream));
      if(int e = batch_launch(b, 0, true)) return e;
      HIP_TRY(hipEventRecord(e1, b->stream));
      HIP_TRY(hipEventSynchronize(e1));
      float t = 0;
      HIP_TRY(hipEventElapsedTime(&t, e0, e1));
      ms = std::min(ms, t);
    }
    if(ms < best_ms) { best_ms = ms; best = c; best_even = even; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->waves = best.waves;
  b->grid = batch_grid(b->module, b->step, best.waves, b->cus);
  for(int i = 0; i < b->nruns; ++i) b->tile_h[i] = height(best.tile_h, best_even, i);
  for(int i = 0; i < b->nruns; ++i) {
    restore_diag_kernel<<<1, 1, 0, b->stream>>>(b->runs[i]->diag, flags[i]);
    b->runs[i]->lam_valid = false;          // the CFL rates were cleared
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  if(tile_h) *tile_h = best.tile_h;
  if(waves) *waves = best.waves;
  return FG2_OK;
}

int fg2_batch_step_async(fg2_batch *b, const double *dts)
{
  if(!b || !dts) return fail(FG2_ERR_ARG, "NULL argument");
  // the host copies of the launch records are rewritten below, after the
  // runs' own queued work (uploads)
  if(int e = batch_join_runs(b)) return e;
  for(fg2_run *r : b->runs) randomize(r);
  b->clocked = false;
  if(int e = batch_fill(b, dts)) return e;
  if(int e = batch_launch(b, 0)) return e;
  for(fg2_run *r : b->runs) { r->cur ^= 1; r->lam_valid = false; }
  return FG2_OK;
}

int fg2_batch_step(fg2_batch *b, const double *dts)
{
  if(int e = fg2_batch_step_async(b, dts)) return e;
  HIP_TRY(hipStreamSynchronize(b->stream));
  int status = FG2_OK;
  for(fg2_run *r : b->runs) {
    const int e = fg2_check(r);
    if(e && !status) status = e;
  }
  return status;
}

int fg2_batch_graph_run(fg2_batch *b, const double *dts, int steps, int replays)
{
  if(!b || !dts || steps < 1 || replays < 0) return fail(FG2_ERR_ARG, "bad graph arguments");
  if((steps & 1) && replays > 1)
    return fail(FG2_ERR_ARG, "replaying an odd step count more than once is not supported");
  if(int e = batch_join_runs(b)) return e;
  // runs with a hydro-isot forcing draw new phases every step: the device
  // clock draws them (from each run's own stream) with the fixed dts
  bool forced = false;
  for(fg2_run *r : b->runs)
    forced = forced || (module_info(r->module).random_phases && first_order_active(r));
  b->clocked = forced;
  if(forced) {
    // the clock draws the phases of every step of the call; the host's
    // stream steps past them below
    if(int e = batch_set_clocks(b, nullptr, nullptr, dts)) return e;
  }
  if(int e = batch_fill(b, dts)) return e;
  HIP_TRY(hipStreamSynchronize(b->stream));
  if(int e = batch_graph(b, steps)) return e;
  for(int k = 0; k < replays; ++k) HIP_TRY(hipGraphLaunch(b->gexec, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for(fg2_run *r : b->runs) {
    // a replay of an odd number of steps flips every run's state buffer
    if(steps & 1) r->cur ^= 1;
    for(long k = 0; k < (long)steps * replays; ++k) randomize(r);
    r->lam_valid = false;
  }
  b->clocked = false;
  return FG2_OK;
}

int fg2_batch_solve(fg2_batch *b, double *t, const double *t_final, int max_steps, int *steps)
{
  if(!b || !t || !t_final || max_steps < 1) return fail(FG2_ERR_ARG, "bad solve arguments");
  for(int i = 0; i < b->nruns; ++i)
    if(!(t[i] <= t_final[i])) return fail(FG2_ERR_ARG, "run %d: t = %g past t_final = %g", i, t[i], t_final[i]);
  if(int e = batch_join_runs(b)) return e;
  if(!b->clk_dev) {
    HIP_TRY(hipMalloc((void **)&b->clk_dev, b->nruns * sizeof(ClockOut)));
    HIP_TRY(hipHostMalloc((void **)&b->clk_host, b->nruns * sizeof(ClockOut), hipHostMallocDefault));
  }
  b->clocked = true;
  if(int e = batch_set_clocks(b, t, t_final, nullptr)) return e;
  if(int e = batch_fill(b, nullptr)) return e;
  HIP_TRY(hipStreamSynchronize(b->stream));
  // whole graphs of CHUNK steps (even: every replay ends on the records of
  // parity 0), then single launches for the remainder of max_steps; the
  // clocks are read back after each replay
  constexpr int CHUNK = 16;
  int done = 0, status = FG2_OK;
  auto gather = [&]() -> int {
    gather_clock_kernel<<<(b->nruns + 255) / 256, 256, 0, b->stream>>>(b->dev[0][0], b->nruns, b->clk_dev);
    HIP_TRY(hipMemcpyAsync(b->clk_host, b->clk_dev, b->nruns * sizeof(ClockOut),
                           hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return FG2_OK;
  };
  // a run is through at its end time, with a NaN, or with a stalled clock
  auto finished = [&]() {
    for(int i = 0; i < b->nruns; ++i)
      if(b->clk_host[i].t < t_final[i] && !(b->clk_host[i].flags & 3u)) return false;
    return true;
  };
  bool all = false;
  if(max_steps >= CHUNK) {
    if(int e = batch_graph(b, CHUNK)) { status = e; goto out; }
    while(done + CHUNK <= max_steps) {
      HIP_TRY(hipGraphLaunch(b->gexec, b->stream));
      done += CHUNK;
      if(int e = gather()) { status = e; goto out; }
      if((all = finished())) break;
    }
  }
  for(int par = 0; !all && done < max_steps; ++done, par ^= 1)
    if(int e = batch_launch(b, par)) { status = e; goto out; }
  if(int e = gather()) { status = e; goto out; }
  // an odd count of single launches leaves the state in the other buffer
  if((done % CHUNK) & 1)
    for(fg2_run *r : b->runs) r->cur ^= 1;
  for(int i = 0; i < b->nruns; ++i) {
    fg2_run *r = b->runs[i];
    t[i] = b->clk_host[i].t;
    if(steps) steps[i] = (int)b->clk_host[i].steps;
    r->rng = b->clk_host[i].rng;            // past the phases of the steps taken
    r->lam_valid = r->dt_mode == FG2_DT_CFL;  // the last step's epilogue reduced them
  }
  for(fg2_run *r : b->runs) {
    const int e = fg2_check(r);
    if(e && !status) status = e;
  }
  for(int i = 0; i < b->nruns && !status; ++i)
    if(b->clk_host[i].flags & 2u)            // as fg2_solve refuses such a step
      status = fail(FG2_ERR_ARG, "run %d: non-positive time step %g at t = %.17g < t_final = %.17g "
                    "after %lld steps", i, b->clk_host[i].dt, b->clk_host[i].t, t_final[i],
                    b->clk_host[i].steps);
out:
  b->clocked = false;
  return status;
}

} // extern "C"

// ---------------------------------------------------------------------------
// The batched step without any allocation (fg2_batch_step_static).
//
// fg2_batch_step's kernels read per-run launch records and tile-queue
// counters that fg2_batch_create allocates.  Here they live in `scratch`, a
// file-scope device array of fixed size that every GPU the library runs on
// holds until the process ends: 3 x FG2_STATIC_MAX_RUNS records (one set
// per substage) and the counters.  Its device address is taken once per
// device with hipGetSymbolAddress on the first call there (the array's name
// in host code is not a device address), and every later use goes through
// that pointer.  All calls on one GPU share the array, so calls that
// overlap on one GPU race; a mutex serialises calls from one process.
// ---------------------------------------------------------------------------
namespace {

constexpr int STATIC_MAX_RUNS = FG2_STATIC_MAX_RUNS;
constexpr size_t STATIC_REC_BYTES = 3 * (size_t)STATIC_MAX_RUNS * sizeof(BatchRun);
constexpr size_t STATIC_BYTES = STATIC_REC_BYTES + 256;      // + the tile counters

static __device__ __align__(256) unsigned char scratch[STATIC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_static_mu;
unsigned char *g_scratch_ptr[MAX_DEVICES] = {};              // scratch's address per device
BatchRun g_static_host[3][STATIC_MAX_RUNS];                  // host side of the records

// scratch's device address on the current device (looked up on the first
// call there); the caller holds g_static_mu
int static_scratch(unsigned char **out)
{
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if(dev < 0 || dev >= MAX_DEVICES) return fail(FG2_ERR_ARG, "device %d out of range", dev);
  if(!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (unsigned char *)p;
  }
  *out = g_scratch_ptr[dev];
  return FG2_OK;
}

} // namespace

extern "C" int fg2_batch_step_static(fg2_run *const *runs, int nruns, const double *dts,
                                     int waves, int tile_h)
{
  if(!runs || !dts || nruns < 1) return fail(FG2_ERR_ARG, "bad batch arguments");
  if(nruns > STATIC_MAX_RUNS)
    return fail(FG2_ERR_SHAPE, "%d runs: fg2_batch_step_static takes at most %d", nruns,
                STATIC_MAX_RUNS);
  if(waves != 1 && waves != 2 && waves != 4) return fail(FG2_ERR_SHAPE, "waves must be 1, 2 or 4");
  const int module = runs[0] ? runs[0]->module : 0;
  bool step = true;
  for(int i = 0; i < nruns; ++i) {
    fg2_run *r = runs[i];
    if(int e = check_run(r)) return e;
    if(r->module != module) return fail(FG2_ERR_ARG, "a batch holds runs of one module");
    if(r->slab_n1g) return fail(FG2_ERR_ARG, "a batch holds whole grids, not slabs");
    step = step && r->p1 && r->p2 && module_info(r->module).one_pass;
  }
  std::lock_guard<std::mutex> lk(g_static_mu);
  unsigned char *base = nullptr;
  if(int e = static_scratch(&base)) return e;
  int dev = 0, cus = 256;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // order after every run's queued work (uploads), as fg2_batch_step does
  for(int i = 0; i < nruns; ++i) {
    HIP_TRY(hipStreamSynchronize(runs[i]->stream));
    randomize(runs[i]);
  }
  // launch records, cut as fg2_batch_create + batch_fill cut them
  const int nstage = step ? 1 : 3;
  int total = 0, fo_blocks = 0;
  for(int k = 0; k < nstage; ++k) {
    for(int i = 0; i < nruns; ++i) {
      fg2_run *r = runs[i];
      const int h = std::min(tile_h > 0 ? tile_h : (step ? 64 : shape_of(r).tile_h), r->n1);
      g_static_host[k][i] = make_record(r, r->cur ^ (step ? 0 : (k & 1)), step ? -1 : k, h, waves,
                                        dts + i, false);
    }
    total = number_records(g_static_host[k], nruns, &fo_blocks);
  }
  hipStream_t st = runs[0]->stream;
  BatchRun *rec = (BatchRun *)base;
  int *counters = (int *)(base + STATIC_REC_BYTES);
  for(int k = 0; k < nstage; ++k)
    HIP_TRY(hipMemcpyAsync(rec + (size_t)k * STATIC_MAX_RUNS, g_static_host[k],
                           nruns * sizeof(BatchRun), hipMemcpyHostToDevice, st));
  if(fo_blocks) HIP_TRY(launch_batch_euler(module, rec, nruns, fo_blocks, st));
  zero_ints_kernel<<<1, 64, 0, st>>>(counters, 3);      // the queues start at tile 0
  const int grid = batch_grid(module, step, waves, cus);
  if(step) {
    HIP_TRY(launch_batch_step(module, rec, nruns, total, waves, grid, counters, 0, st));
  } else {
    for(int k = 0; k < 3; ++k)
      HIP_TRY(launch_batch_substage(module, k, rec + (size_t)k * STATIC_MAX_RUNS, nruns, total,
                                    waves, grid, counters + k, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  int status = FG2_OK;
  for(int i = 0; i < nruns; ++i) {
    runs[i]->cur ^= 1;
    runs[i]->lam_valid = false;
    const int e = fg2_check(runs[i]);
    if(e && !status) status = e;
  }
  return status;
}

// ---------------------------------------------------------------------------
// Serving: continuous batching of many runs of different sizes.
//
// Requests (module, grid, step count, initial state) queue on submit.  Each
// tick retires runs whose copy-back has completed, admits queued requests in
// arrival order while their state fits the HBM budget, and advances every
// active run by one RK3 step with one batched one-pass launch per module.
// State stays on the device from admission to retirement, in an arena
// allocated once (first-fit with coalescing), so serving makes no hipMalloc
// or hipFree calls; a served run holds only its two state arrays (the one-
// pass step needs no v).  Launch records are quadruple-buffered in pinned
// memory behind events, so the host prepares the next ticks while one runs.
// States move between the caller's host arrays and the arena through pinned
// staging buffers; the host-side copies into and out of them are cut into
// row blocks of at most a per-tick byte budget, so admitting or retiring a
// large run never holds up the ticks that keep the device busy.
// Each launch's tile height and workgroup shape come from the active runs
// (server_shape) unless fixed at creation.
// A run's arithmetic is that of fg2_step (the same step tile), so a result
// is the same bits alone and in any mix.
// ---------------------------------------------------------------------------
namespace {

struct Request {
  long long id;
  int module, n1, n2, nv, steps, done_steps = 0;
  long s = 0;
  std::vector<float> host;             // initial state, then the result
  double t_arrive = 0, t_admit = -1, t_done = -1;
  size_t off = 0, bytes = 0;           // arena region: two state arrays
  float *in[2] = {nullptr, nullptr};
  int cur = 0, diag_slot = -1;
  Params P{};
  double dt = 0;
  bool cfl = false, fo = false;        // CFL steps; an active first_order pre-step
  ClockInit clk{};                     // the run's device clock from admission on
  FusedShape shape{0, 4};              // per-run mode: the run's own launch shape
  int sidx = 0;                        // per-run mode: the pool stream it steps on
  int status = FG2_OK;
  bool active = false, retiring = false;
  int up_slot = -1, up_rows = 0;       // staging slot and rows sent while uploading
  int down_slot = -1, down_rows = 0;   // staging slot and rows copied out while retiring
  bool down_landed = false;            // the state and flags are in the staging slot
};

} // namespace

struct fg2_server {
  size_t budget = 0, in_use = 0;
  int waves = 0, tile_h = 0, cus = 256;        // 0: chosen per launch (server_shape)
  hipStream_t stream = nullptr;
  // one compute stream per module (mstream[0] is `stream`): the batched
  // launches of the modules of a tick run side by side, each filling the
  // slots the other's last round leaves idle
  static constexpr int MAXM = 8;
  hipStream_t mstream[MAXM] = {};
  char *arena = nullptr;
  std::map<size_t, size_t> free_list;          // offset -> size
  DeviceDiag *diags = nullptr;
  DeviceDiag *diags_host = nullptr;            // pinned readback of retired slots
  int ndiag = 0;
  std::vector<int> free_diag;
  long long next_id = 0;
  std::vector<Request *> queue, uploading, active, finished;
  std::map<long long, Request *> all;
  static constexpr int SLOTS = 4;              // launch records in flight
  BatchRun *rec_dev[SLOTS] = {};
  BatchRun *rec_pinned[SLOTS] = {};
  size_t rec_cap = 0;
  hipEvent_t rec_done[SLOTS] = {};
  int rec_slot = 0;
  int *counters = nullptr;                     // one per slot
  std::vector<BatchRun> rec_host;
  long long cell_steps = 0;
  // host <-> device copies of admitted / retired states run on their own
  // stream through pinned staging buffers, so they overlap the steps
  hipStream_t copy = nullptr;
  static constexpr int STAGES = 4;
  float *up[STAGES] = {}, *down[STAGES] = {};
  size_t up_cap[STAGES] = {}, down_cap[STAGES] = {};
  hipEvent_t up_done[STAGES] = {}, down_done[STAGES] = {};
  Request *up_owner[STAGES] = {}, *down_owner[STAGES] = {};
  // host bytes copied into or out of the staging buffers per tick
  static constexpr size_t COPY_BUDGET = (size_t)8 << 20;
  // per-run mode (FG2_SERVE_PER_RUN): each active run launches its own
  // one-pass step on one of NPOOL streams; at most SLOTS ticks in flight
  int mode = FG2_SERVE_BATCHED;
  static constexpr int NPOOL = 8;
  hipStream_t pool[NPOOL] = {};
  hipEvent_t tick_ev[SLOTS][NPOOL] = {};
  int tick_slot = 0, next_sidx = 0;
};

static_assert(Modules::size <= fg2_server::MAXM, "one compute stream per module");

namespace {

// Grow a pinned staging buffer to hold `need` bytes: at least double its
// size, so a stream of growing requests reallocates (a slow pinned
// allocation that holds up the tick) a few times, not at every new maximum.
int stage_buffer(float **buf, size_t *cap, size_t need)
{
  if(need <= *cap) return FG2_OK;
  if(*buf) HIP_TRY(hipHostFree(*buf));
  *buf = nullptr;
  const size_t c = std::max({need, 2 * *cap, (size_t)1 << 24});
  HIP_TRY(hipHostMalloc((void **)buf, c, hipHostMallocDefault));
  *cap = c;
  return FG2_OK;
}

bool arena_alloc(fg2_server *S, size_t bytes, size_t *off)
{
  bytes = (bytes + 255) / 256 * 256;
  for(auto it = S->free_list.begin(); it != S->free_list.end(); ++it)
    if(it->second >= bytes) {
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      S->free_list.erase(it);
      if(rest) S->free_list[at] = rest;
      return true;
    }
  return false;
}

void arena_free(fg2_server *S, size_t off, size_t bytes)
{
  bytes = (bytes + 255) / 256 * 256;
  auto it = S->free_list.emplace(off, bytes).first;
  auto nx = std::next(it);
  if(nx != S->free_list.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    S->free_list.erase(nx);
  }
  if(it != S->free_list.begin()) {
    auto pv = std::prev(it);
    if(pv->first + pv->second == it->first) {
      pv->second += it->second;
      S->free_list.erase(it);
    }
  }
}

int server_grow(fg2_server *S, size_t need)
{
  if(need <= S->rec_cap) return FG2_OK;
  size_t cap = std::max<size_t>(need, 2 * S->rec_cap + 64);
  for(int m = 0; m < module_count(); ++m) HIP_TRY(hipStreamSynchronize(S->mstream[m]));
  for(int k = 0; k < fg2_server::SLOTS; ++k) {
    if(S->rec_dev[k]) HIP_TRY(hipFree(S->rec_dev[k]));
    if(S->rec_pinned[k]) HIP_TRY(hipHostFree(S->rec_pinned[k]));
    HIP_TRY(hipMalloc((void **)&S->rec_dev[k], cap * sizeof(BatchRun)));
    HIP_TRY(hipHostMalloc((void **)&S->rec_pinned[k], cap * sizeof(BatchRun), hipHostMallocDefault));
  }
  S->rec_cap = cap;
  return FG2_OK;
}

// The stream a request's steps run on.
hipStream_t run_stream(const fg2_server *S, const Request *q)
{
  return S->mode != FG2_SERVE_BATCHED ? S->pool[q->sidx] : S->mstream[q->module];
}

// Bytes of one row of a request's state in its host array.
size_t host_row_bytes(const Request *q) { return (size_t)q->n2 * q->nv * sizeof(float); }

// Copy rows [*rows, n1) of a state between a host array and a pinned
// staging buffer (to_stage: host -> staging) within *budget bytes, at least
// one row when any budget is left; returns the first row copied.  For an
// upload, each block is then queued to the device by the caller.
int stage_rows(const Request *q, char *stage, char *host, bool to_stage, int *rows,
               size_t *budget)
{
  const size_t hp = host_row_bytes(q);
  const int r0 = *rows;
  if(r0 >= q->n1 || *budget == 0) return r0;
  const int r1 = (int)std::min<size_t>((size_t)q->n1, (size_t)r0 + std::max<size_t>(1, *budget / hp));
  const size_t bytes = (size_t)(r1 - r0) * hp, at = (size_t)r0 * hp;
  if(to_stage) memcpy(stage + at, host + at, bytes);
  else memcpy(host + at, stage + at, bytes);
  *budget -= std::min(*budget, bytes);
  *rows = r1;
  return r0;
}

// Admission.  Queued requests, in arrival order, take arena space, a flag
// slot and a staging buffer whose previous upload has reached the device.
// Uploading runs then send rows within *budget host bytes: each block is
// copied into the staging buffer and queued to the device on the copy
// stream.  A run whose rows are all sent becomes active; its steps wait for
// the upload on its compute stream.
int server_admit(fg2_server *S, double now, size_t *budget)
{
  while(!S->queue.empty() && !S->free_diag.empty()) {
    int u = -1;
    for(int k = 0; k < fg2_server::STAGES && u < 0; ++k)
      if(!S->up_owner[k] && hipEventQuery(S->up_done[k]) == hipSuccess) u = k;
    if(u < 0) break;                                   // every staging buffer is in use
    Request *q = S->queue.front();
    const size_t one = (size_t)array_floats(q->n1, q->n2, q->nv) * sizeof(float);
    size_t off;
    if(!arena_alloc(S, 2 * one, &off)) break;          // wait for retirements
    if(int e = stage_buffer(&S->up[u], &S->up_cap[u], host_row_bytes(q) * q->n1)) {
      arena_free(S, off, 2 * one);
      return e;
    }
    q->off = off;
    q->bytes = 2 * one;
    q->s = row_stride(q->n2, q->nv);
    const long io = interior_offset(q->n2, q->nv);
    q->in[0] = (float *)(S->arena + off) + io;
    q->in[1] = (float *)(S->arena + off + one) + io;
    q->diag_slot = S->free_diag.back();
    S->free_diag.pop_back();
    set_clock_kernel<<<1, 1, 0, S->copy>>>(S->diags + q->diag_slot, q->clk, 1);
    q->up_slot = u;
    q->up_rows = 0;
    S->up_owner[u] = q;
    S->in_use += q->bytes;
    S->queue.erase(S->queue.begin());
    S->uploading.push_back(q);
  }
  while(!S->uploading.empty()) {
    Request *q = S->uploading.front();
    const int u = q->up_slot;
    const size_t hp = host_row_bytes(q);
    char *stage = (char *)S->up[u];
    while(q->up_rows < q->n1 && *budget > 0) {
      const int r0 = stage_rows(q, stage, (char *)q->host.data(), true, &q->up_rows, budget);
      HIP_TRY(hipMemcpy2DAsync(q->in[0] + (long)r0 * q->s, q->s * sizeof(float),
                               stage + (size_t)r0 * hp, hp, hp, q->up_rows - r0,
                               hipMemcpyHostToDevice, S->copy));
    }
    if(q->up_rows < q->n1) break;                      // the budget is spent
    HIP_TRY(hipEventRecord(S->up_done[u], S->copy));
    S->up_owner[u] = nullptr;
    q->up_slot = -1;
    if(S->mode != FG2_SERVE_BATCHED) {
      q->sidx = S->next_sidx;
      S->next_sidx = (S->next_sidx + 1) % fg2_server::NPOOL;
    }
    // the plain path's first kick multiplies the old v by beta_0 = 0: start
    // v at zero so that whatever the arena held cannot make it NaN
    if(S->mode == FG2_SERVE_PER_RUN_PLAIN)
      HIP_TRY(hipMemsetAsync(q->in[1] - interior_offset(q->n2, q->nv), 0, q->bytes / 2,
                             run_stream(S, q)));
    HIP_TRY(hipStreamWaitEvent(run_stream(S, q), S->up_done[u], 0));  // steps after the upload
    if(q->cfl)                           // the first step's rates: those of the uploaded state
      HIP_TRY(launch_cfl(q->module, q->in[0], q->s, q->P, S->diags + q->diag_slot, run_stream(S, q)));
    q->t_admit = now;
    q->active = true;
    S->uploading.erase(S->uploading.begin());
    S->active.push_back(q);
  }
  return FG2_OK;
}

// Workgroup shape and tile height of one batched launch over the active runs
// of module m: the server's fixed values, or (0) two-wavefront workgroups
// (on a saturated stream of the serving workload they served 46 G
// cell-updates/s against 45 G for the best four-wavefront and 43 G for the
// best one-wavefront shape) with the tile height that minimises the rounds
// of resident workgroups times the cost of a tile (its rows + 6).
FusedShape server_shape(const fg2_server *S, int m)
{
  FusedShape best{S->tile_h > 0 ? S->tile_h : 64, S->waves > 0 ? S->waves : 2};
  if(S->tile_h > 0 && S->waves > 0) return best;
  long best_cost = -1;
  for(int w : {4, 2, 1}) {
    if(w != (S->waves > 0 ? S->waves : 2)) continue;
    const long slots = batch_grid(m, true, w, S->cus);
    for(int h : {16, 24, 32, 40, 48, 56, 64, 72, 84, 96, 112, 128}) {
      if(S->tile_h > 0 && h != S->tile_h) continue;
      long tiles = 0;
      for(const Request *q : S->active)
        if(q->module == m && q->done_steps < q->steps)
          tiles += step_launch_tiles(q->n1, q->n2, std::min(h, q->n1), w, nullptr);
      const long cost = ((tiles + slots - 1) / slots) * (h + 6);
      if(best_cost < 0 || cost < best_cost) { best_cost = cost; best = FusedShape{h, w}; }
    }
  }
  return best;
}

// Per-run modes: every active run's own step on its pool stream, either the
// one-pass step with the one-pass kernel's automatic shape for a single run
// of its size (fg2_step's with FG2_KERNEL_ONE_PASS) or the plain path
// (fg2_step's FG2_PATH_PLAIN: u in place, v in the run's second array).
int server_launch_per_run(fg2_server *S)
{
  const int slot = S->tick_slot;
  S->tick_slot = (S->tick_slot + 1) % fg2_server::SLOTS;
  for(int k = 0; k < fg2_server::NPOOL; ++k)
    HIP_TRY(hipEventSynchronize(S->tick_ev[slot][k]));     // at most SLOTS ticks queued
  for(Request *q : S->active) {
    if(q->done_steps >= q->steps) continue;
    hipStream_t st = S->pool[q->sidx];
    DeviceDiag *diag = S->diags + q->diag_slot;
    if(S->mode == FG2_SERVE_PER_RUN_PLAIN) {
      float *u = q->in[q->cur], *v = q->in[q->cur ^ 1];
      for(int k = 0; k < 3; ++k) {
        HIP_TRY(launch_bcond(q->module, u, q->s, q->P, q->P.p1, q->P.p2, st));
        HIP_TRY(launch_kick(q->module, v, u, q->s, q->P, (float)RK_BETA[k], st));
        HIP_TRY(launch_drift(q->module, u, v, q->s, q->P, (float)(q->dt * RK_GAMMA[k]), diag, st));
      }
    } else {
      HIP_TRY(launch_step(q->module, q->in[q->cur], q->in[q->cur ^ 1], q->s, q->P, q->dt, q->shape,
                          0, diag, st));
      q->cur ^= 1;
    }
    ++q->done_steps;
    S->cell_steps += (long long)q->n1 * q->n2;
  }
  for(int k = 0; k < fg2_server::NPOOL; ++k) HIP_TRY(hipEventRecord(S->tick_ev[slot][k], S->pool[k]));
  return FG2_OK;
}

int server_launch(fg2_server *S)
{
  if(S->mode != FG2_SERVE_BATCHED) return server_launch_per_run(S);
  for(int m = 0; m < module_count(); ++m) {        // one launch per module with active runs
    hipStream_t ms = S->mstream[m];
    S->rec_host.clear();
    const FusedShape sh = server_shape(S, m);
    const int th = sh.tile_h;
    bool clocked = false, cfl = false;
    for(Request *q : S->active) {
      if(q->module != m || q->done_steps >= q->steps) continue;
      BatchRun R{};
      R.uin = q->in[q->cur];
      R.uout = q->in[q->cur ^ 1];
      R.s = q->s;
      R.P = q->P;
      R.tile_h = std::min(th, q->n1);
      R.ntiles = (int)step_launch_tiles(q->n1, q->n2, R.tile_h, sh.waves, &R.nstrips);
      for(int k = 0; k < 3; ++k) R.g[k] = (float)(q->dt * RK_GAMMA[k]);
      R.cfl = q->cfl;
      R.fo = q->fo;
      R.edt = (float)q->dt;
      R.diag = S->diags + q->diag_slot;
      S->rec_host.push_back(R);
      // CFL steps, and forcing phases drawn every step, come from the clock
      clocked = clocked || q->cfl || (q->fo && module_info(m).random_phases);
      cfl = cfl || q->cfl;
    }
    if(S->rec_host.empty()) continue;
    int fo_blocks = 0;
    const int begin = number_records(S->rec_host.data(), (int)S->rec_host.size(), &fo_blocks);
    if(int e = server_grow(S, S->rec_host.size())) return e;
    const int slot = S->rec_slot;
    S->rec_slot = (S->rec_slot + 1) % fg2_server::SLOTS;
    HIP_TRY(hipEventSynchronize(S->rec_done[slot]));   // the slot's previous launch is done
    memcpy(S->rec_pinned[slot], S->rec_host.data(), S->rec_host.size() * sizeof(BatchRun));
    HIP_TRY(hipMemcpyAsync(S->rec_dev[slot], S->rec_pinned[slot], S->rec_host.size() * sizeof(BatchRun),
                           hipMemcpyHostToDevice, ms));
    const int n = (int)S->rec_host.size();
    if(clocked)
      HIP_TRY(launch_clock(m, S->rec_dev[slot], nullptr, nullptr, 1, 1, n, ms));
    if(fo_blocks) HIP_TRY(launch_batch_euler(m, S->rec_dev[slot], n, fo_blocks, ms));
    zero_ints_kernel<<<1, 64, 0, ms>>>(S->counters + slot, 1);
    HIP_TRY(launch_batch_step(m, S->rec_dev[slot], n, begin, sh.waves,
                              batch_grid(m, true, sh.waves, S->cus, cfl), S->counters + slot,
                              cfl, ms));
    HIP_TRY(hipEventRecord(S->rec_done[slot], ms));
  }
  for(Request *q : S->active) {
    if(q->done_steps >= q->steps) continue;
    q->cur ^= 1;
    ++q->done_steps;
    S->cell_steps += (long long)q->n1 * q->n2;
  }
  return FG2_OK;
}

int server_retire(fg2_server *S, double now, size_t *budget)
{
  // queue the copy-back (on the copy stream, after the run's last step) of
  // runs that took their last step, one pinned staging buffer each
  for(Request *q : S->active)
    if(!q->retiring && q->done_steps >= q->steps) {
      int d = -1;
      for(int k = 0; k < fg2_server::STAGES && d < 0; ++k) if(!S->down_owner[k]) d = k;
      if(d < 0) break;                                          // retry next tick
      const size_t hp = host_row_bytes(q);
      if(int e = stage_buffer(&S->down[d], &S->down_cap[d], hp * q->n1)) return e;
      HIP_TRY(hipEventRecord(S->down_done[d], run_stream(S, q)));  // the run's last step
      HIP_TRY(hipStreamWaitEvent(S->copy, S->down_done[d], 0));
      HIP_TRY(hipMemcpy2DAsync(S->down[d], hp, q->in[q->cur], q->s * sizeof(float), hp, q->n1,
                               hipMemcpyDeviceToHost, S->copy));
      HIP_TRY(hipMemcpyAsync(S->diags_host + q->diag_slot, S->diags + q->diag_slot,
                             sizeof(DeviceDiag), hipMemcpyDeviceToHost, S->copy));
      HIP_TRY(hipEventRecord(S->down_done[d], S->copy));
      S->down_owner[d] = q;
      q->down_slot = d;
      q->down_rows = 0;
      q->down_landed = false;
      q->retiring = true;
    }
  // a run whose copy-back has landed releases its arena space and flag slot
  // at once, then copies its rows out of the staging buffer within *budget
  // host bytes per tick; it is finished when the last row is out
  for(size_t k = 0; k < S->active.size();) {
    Request *q = S->active[k];
    if(!q->retiring) { ++k; continue; }
    const int d = q->down_slot;
    if(!q->down_landed) {
      if(hipEventQuery(S->down_done[d]) != hipSuccess) { ++k; continue; }
      q->down_landed = true;
      const unsigned int fl = S->diags_host[q->diag_slot].flags;
      q->status = (fl & 1u) ? FG2_ERR_NAN : (fl & 2u) ? FG2_ERR_ARG : FG2_OK;
      arena_free(S, q->off, q->bytes);
      S->in_use -= q->bytes;
      S->free_diag.push_back(q->diag_slot);
    }
    while(q->down_rows < q->n1 && *budget > 0)
      stage_rows(q, (char *)S->down[d], (char *)q->host.data(), false, &q->down_rows, budget);
    if(q->down_rows < q->n1) { ++k; continue; }
    S->down_owner[d] = nullptr;
    q->down_slot = -1;
    q->t_done = now;
    q->active = false;
    S->active.erase(S->active.begin() + k);
    S->finished.push_back(q);
  }
  return FG2_OK;
}

} // namespace

extern "C" {

int fg2_server_create(fg2_server **out, size_t hbm_budget, int waves, int tile_h)
{
  if(!out || hbm_budget < (1u << 20)) return fail(FG2_ERR_ARG, "bad server arguments");
  if(waves != 0 && waves != 1 && waves != 2 && waves != 4)
    return fail(FG2_ERR_SHAPE, "waves must be 0 (chosen per launch), 1, 2 or 4");
  fg2_server *S = new(std::nothrow) fg2_server;
  if(!S) return fail(FG2_ERR_NOMEM, "host allocation failed");
  S->budget = hbm_budget / 256 * 256;
  S->waves = waves;
  S->tile_h = tile_h > 0 ? tile_h : 0;     // 0: chosen per launch (server_launch)
  S->ndiag = 8192;
  int dev = 0;
  hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
  S->mstream[0] = S->stream;
  for(int m = 1; m < module_count() && m < fg2_server::MAXM && e == hipSuccess; ++m)
    e = hipStreamCreateWithFlags(&S->mstream[m], hipStreamNonBlocking);
  if(e == hipSuccess) e = hipGetDevice(&dev);
  if(e == hipSuccess) e = hipDeviceGetAttribute(&S->cus, hipDeviceAttributeMultiprocessorCount, dev);
  if(e == hipSuccess) e = hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking);
  for(int k = 0; k < fg2_server::STAGES && e == hipSuccess; ++k) {
    e = hipEventCreateWithFlags(&S->up_done[k], hipEventDisableTiming);
    if(e == hipSuccess) e = hipEventCreateWithFlags(&S->down_done[k], hipEventDisableTiming);
    if(e == hipSuccess) e = hipEventRecord(S->up_done[k], S->copy);
  }
  if(e == hipSuccess) e = hipMalloc((void **)&S->arena, S->budget);
  if(e == hipSuccess) e = hipMalloc((void **)&S->diags, S->ndiag * sizeof(DeviceDiag));
  if(e == hipSuccess) e = hipHostMalloc((void **)&S->diags_host, S->ndiag * sizeof(DeviceDiag), hipHostMallocDefault);
  if(e == hipSuccess) e = hipMalloc((void **)&S->counters, fg2_server::SLOTS * sizeof(int));
  for(int k = 0; k < fg2_server::SLOTS && e == hipSuccess; ++k)
    e = hipEventCreateWithFlags(&S->rec_done[k], hipEventDisableTiming);
  for(int k = 0; k < fg2_server::SLOTS && e == hipSuccess; ++k) e = hipEventRecord(S->rec_done[k], S->stream);
  for(int j = 0; j < fg2_server::NPOOL && e == hipSuccess; ++j) {
    e = hipStreamCreateWithFlags(&S->pool[j], hipStreamNonBlocking);
    for(int k = 0; k < fg2_server::SLOTS && e == hipSuccess; ++k) {
      e = hipEventCreateWithFlags(&S->tick_ev[k][j], hipEventDisableTiming);
      if(e == hipSuccess) e = hipEventRecord(S->tick_ev[k][j], S->pool[j]);
    }
  }
  if(e != hipSuccess) {
    const int code = e == hipErrorOutOfMemory ? FG2_ERR_NOMEM : FG2_ERR_HIP;
    fg2_server_free(S);
    return fail(code, "server setup: %s", hipGetErrorString(e));
  }
  S->free_list[0] = S->budget;
  for(int k = S->ndiag - 1; k >= 0; --k) S->free_diag.push_back(k);
  *out = S;
  return FG2_OK;
}

void fg2_server_free(fg2_server *S)
{
  if(!S) return;
  for(int m = 0; m < fg2_server::MAXM; ++m) if(S->mstream[m]) (void)hipStreamSynchronize(S->mstream[m]);
  if(S->copy) (void)hipStreamSynchronize(S->copy);
  for(int j = 0; j < fg2_server::NPOOL; ++j) {
    if(S->pool[j]) (void)hipStreamSynchronize(S->pool[j]);
    for(int k = 0; k < fg2_server::SLOTS; ++k) if(S->tick_ev[k][j]) (void)hipEventDestroy(S->tick_ev[k][j]);
    if(S->pool[j]) (void)hipStreamDestroy(S->pool[j]);
  }
  for(auto &kv : S->all) delete kv.second;
  for(int k = 0; k < fg2_server::STAGES; ++k) {
    if(S->up[k]) (void)hipHostFree(S->up[k]);
    if(S->down[k]) (void)hipHostFree(S->down[k]);
    if(S->up_done[k]) (void)hipEventDestroy(S->up_done[k]);
    if(S->down_done[k]) (void)hipEventDestroy(S->down_done[k]);
  }
  if(S->copy) (void)hipStreamDestroy(S->copy);
  for(int k = 0; k < fg2_server::SLOTS; ++k) {
    if(S->rec_dev[k]) (void)hipFree(S->rec_dev[k]);
    if(S->rec_pinned[k]) (void)hipHostFree(S->rec_pinned[k]);
    if(S->rec_done[k]) (void)hipEventDestroy(S->rec_done[k]);
  }
  if(S->arena) (void)hipFree(S->arena);
  if(S->diags) (void)hipFree(S->diags);
  if(S->diags_host) (void)hipHostFree(S->diags_host);
  if(S->counters) (void)hipFree(S->counters);
  for(int m = 1; m < fg2_server::MAXM; ++m) if(S->mstream[m]) (void)hipStreamDestroy(S->mstream[m]);
  if(S->stream) (void)hipStreamDestroy(S->stream);
  delete S;
}

int fg2_server_submit(fg2_server *S, const char *module, int n1, int n2, int steps,
                      const float *init, double now, long long *id)
{
  if(!S || !init || steps < 1) return fail(FG2_ERR_ARG, "bad request");
  const int m = module_id(module);
  if(m < 0 || !module_info(m).one_pass)
    return fail(FG2_ERR_ARG, "served runs take the one-pass step (module %s)", module ? module : "(null)");
  if(n1 < HALF || n2 < HALF) return fail(FG2_ERR_SHAPE, "grid %d x %d too small", n1, n2);
  // geometry, parameters and dt of a default run of this module (periodic,
  // src/global.cpp:28-29 and the module's settings()), no device memory
  fg2_run tmp;
  tmp.module = m; tmp.nv = nvar_of(m); tmp.n1 = n1; tmp.n2 = n2; tmp.cus = S->cus;
  for(auto &p : param_defs(m)) tmp.par.push_back(p.value);
  fg2_config(&tmp);
  return fg2_server_submit_run(S, &tmp, steps, init, now, id);
}

int fg2_server_submit_run(fg2_server *S, const fg2_run *like, int steps, const float *init,
                          double now, long long *id)
{
  if(!S || !like || !init || steps < 1) return fail(FG2_ERR_ARG, "bad request");
  if(int e = check_run(like)) return e;
  const int m = like->module;
  if(!module_info(m).one_pass)
    return fail(FG2_ERR_ARG, "served runs take the one-pass step (module %s)", module_info(m).name);
  if(!like->p1 || !like->p2) return fail(FG2_ERR_ARG, "served runs are periodic");
  if(like->slab_n1g) return fail(FG2_ERR_ARG, "served runs are whole grids, not slabs");
  const int n1 = like->n1, n2 = like->n2;
  const size_t need = 2 * (size_t)array_floats(n1, n2, nvar_of(m)) * sizeof(float);
  if(need > S->budget) return fail(FG2_ERR_NOMEM, "request needs %zu bytes, budget %zu", need, S->budget);
  const bool cfl = like->dt_mode == FG2_DT_CFL, fo = first_order_active(like);
  if(S->mode != FG2_SERVE_BATCHED && (cfl || fo))
    return fail(FG2_ERR_ARG, "per-run mode serves fixed-dt runs without a first_order pre-step");
  Request *q = new(std::nothrow) Request;
  if(!q) return fail(FG2_ERR_NOMEM, "host allocation failed");
  q->id = S->next_id++;
  q->module = m; q->n1 = n1; q->n2 = n2; q->nv = nvar_of(m); q->steps = steps;
  q->host.assign(init, init + (size_t)n1 * n2 * q->nv);
  // the run's parameters, geometry, dt rule and forcing stream
  q->P = make_params(like);
  q->cfl = cfl;
  q->fo = fo;
  q->dt = fg2_getdt(like);
  q->clk = clock_of(like, 0.0, INFINITY, cfl ? 0.0 : q->dt);
  fg2_run shape_run;
  shape_run.module = m; shape_run.nv = q->nv; shape_run.n1 = n1; shape_run.n2 = n2;
  shape_run.cus = S->cus;
  q->shape = step_shape_of(&shape_run);
  q->t_arrive = now;
  S->queue.push_back(q);
  S->all[q->id] = q;
  if(id) *id = q->id;
  return FG2_OK;
}

int fg2_server_tick(fg2_server *S, double now, int *active, int *queued, int *finished)
{
  if(!S) return fail(FG2_ERR_ARG, "NULL server");
  size_t budget = fg2_server::COPY_BUDGET;
  if(int e = server_retire(S, now, &budget)) return e;
  if(int e = server_admit(S, now, &budget)) return e;
  if(int e = server_launch(S)) return e;
  if(int e = server_retire(S, now, &budget)) return e;
  if(S->active.empty() && S->queue.empty() && S->uploading.empty()) {
    for(int m = 0; m < module_count(); ++m) HIP_TRY(hipStreamSynchronize(S->mstream[m]));
    HIP_TRY(hipStreamSynchronize(S->copy));
    for(int j = 0; j < fg2_server::NPOOL; ++j) HIP_TRY(hipStreamSynchronize(S->pool[j]));
  }
  if(active) *active = (int)S->active.size();
  if(queued) *queued = (int)(S->queue.size() + S->uploading.size());
  if(finished) *finished = (int)S->finished.size();
  return FG2_OK;
}

int fg2_server_result(fg2_server *S, long long id, float *out, double *times, int *status)
{
  if(!S) return fail(FG2_ERR_ARG, "NULL server");
  auto it = S->all.find(id);
  if(it == S->all.end()) return fail(FG2_ERR_ARG, "no request %lld", id);
  Request *q = it->second;
  if(times) { times[0] = q->t_arrive; times[1] = q->t_admit; times[2] = q->t_done; }
  if(status) *status = q->t_done < 0 ? -1 : q->status;
  if(out && q->t_done >= 0) memcpy(out, q->host.data(), q->host.size() * sizeof(float));
  return FG2_OK;
}

int fg2_server_set_mode(fg2_server *S, int mode)
{
  if(!S || (mode != FG2_SERVE_BATCHED && mode != FG2_SERVE_PER_RUN &&
            mode != FG2_SERVE_PER_RUN_PLAIN))
    return fail(FG2_ERR_ARG, "bad server mode");
  if(!S->active.empty() || !S->queue.empty() || !S->uploading.empty())
    return fail(FG2_ERR_ARG, "the mode changes only while the server is idle");
  S->mode = mode;
  return FG2_OK;
}

int fg2_server_stats(fg2_server *S, long long *cell_steps, size_t *bytes_in_use)
{
  if(!S) return fail(FG2_ERR_ARG, "NULL server");
  if(cell_steps) *cell_steps = S->cell_steps;
  if(bytes_in_use) *bytes_in_use = S->in_use;
  return FG2_OK;
}

} // extern "C"
