// kernels.hip -- HIP kernels for gfx950: the plain per-operation path
// (kick, drift, boundary fill, first_order), the fused RK3 substage, the
// one-pass RK3 step and their batched forms.
//
// Fused substage, one wavefront per tile of 64 columns x tile_h rows:
//   - lane j owns column j0 + j; the wave marches down the slow axis (n1)
//     keeping rows i-3 .. i+3 of its column, and the next one or two rows
//     in flight, in a register ring (8 or 9 slots) whose slot index is a
//     compile-time constant in the unrolled loop, so the ring never moves
//     data;
//   - fast-axis neighbours of the centre row come from DPP wave shifts
//     (j+-1) and ds_bpermute lane shifts (j+-2, j+-3): no LDS allocation,
//     no barrier;
//   - the mixed derivative uses g_k(x) = f(i+k,x) - f(i-k,x), shifted by
//     +-k lanes, so D12 = sum_k m_k (g_k(j+k) - g_k(j-k));
//   - boundaries are index maps (periodic wrap, even mirror on dim 1,
//     mirror with the module's odd-variable sign on dim 2), which give the
//     values the reference's ghost fill would (src/bcond.cu:22-86);
//   - the epilogue does v <- F + beta v and u_out <- u_in + gamma dt v,
//     skipping the v read in stage 0 (beta = 0) and the v write in stage 2
//     (the next step's stage 0 ignores v).
// Wavefronts of one workgroup never share data, so they synchronise only at
// wavefront scope: no s_barrier and no vmcnt drain inside the march.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <type_traits>
#include "launch.h"
#include "physics.h"

namespace fg2 {

// Compiler barrier for LDS traffic between lanes of one wavefront.  A
// wavefront's DS instructions execute in order, so no hardware wait is needed.
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template<int NV> __device__ __forceinline__ void load_cell(const float *p, float (&x)[NV])
{
  if constexpr(NV == 4) {
    const float4 a = *(const float4 *)p;
    x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
  } else if constexpr(NV == 6) {
    const float2 a = *(const float2 *)p, b = *(const float2 *)(p + 2), c = *(const float2 *)(p + 4);
    x[0] = a.x; x[1] = a.y; x[2] = b.x; x[3] = b.y; x[4] = c.x; x[5] = c.y;
  } else {
#pragma unroll
    for(int q = 0; q < NV; ++q) x[q] = p[q];
  }
}

template<int NV> __device__ __forceinline__ void store_cell(float *p, const float (&x)[NV])
{
  if constexpr(NV == 4) {
    *(float4 *)p = make_float4(x[0], x[1], x[2], x[3]);
  } else if constexpr(NV == 6) {
    *(float2 *)p = make_float2(x[0], x[1]);
    *(float2 *)(p + 2) = make_float2(x[2], x[3]);
    *(float2 *)(p + 4) = make_float2(x[4], x[5]);
  } else {
#pragma unroll
    for(int q = 0; q < NV; ++q) p[q] = x[q];
  }
}

// Row index of a possibly-ghost row r (|overflow| <= HALF, n >= HALF):
// periodic wrap, or the even mirror of Neumann1 (src/bcond.cu:45-57).
__device__ __forceinline__ int map_row(int r, int n, int periodic)
{
  if(r < 0) return periodic ? r + n : -1 - r;
  if(r >= n) return periodic ? r - n : 2 * n - 1 - r;
  return r;
}
// Column index of a possibly-ghost column c, and the sign the reflective
// boundary applies to the module's odd variable (src/bcond.cu:59-75).
__device__ __forceinline__ int map_col(int c, int n, int periodic, float &sign)
{
  sign = 1.0f;
  if(c < 0) { if(periodic) return c + n; sign = -1.0f; return -1 - c; }
  if(c >= n) { if(periodic) return c - n; sign = -1.0f; return 2 * n - 1 - c; }
  return c;
}

// Map a linear workgroup id to a tile id so that the workgroups one XCD
// receives (ids x, x+8, x+16, ... under round-robin dispatch) cover a
// contiguous range of tiles, whose neighbouring strips share halo columns
// in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
  constexpr int NX = 8;
  const int per = nb / NX, rem = nb % NX;
  const int x = bid % NX, k = bid / NX;
  return (x < rem) ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

// ---------------------------------------------------------------------------
// Fused substage, one tile.  Shared by the single-run and the batched kernels.
// ---------------------------------------------------------------------------
struct TileOut { int bad; float lam_adv, lam_diff; };

// Value of x at lane (lane + k) of the wavefront (ds_bpermute; addr holds
// ((lane + k) & 63) * 4).  Lanes 3..60 never wrap for |k| <= 3.
__device__ __forceinline__ float lane_shift(float x, int addr)
{
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x)));
}

// Global-memory (address space 1) byte pointers.  A uniform row address
// held in scalar registers plus a per-lane 32-bit byte offset makes one
// global load or store with a scalar base and a single offset register
// (global_load ... vN, s[..]), instead of a 64-bit address pair per lane;
// keeping the global address space keeps these loads off the LDS wait
// counter (a flat access counts on both).
#define FG2_GLOBAL __attribute__((address_space(1)))
using gbytes = FG2_GLOBAL char *;
using gcbytes = FG2_GLOBAL const char *;

// p (the same in every lane) as a global byte pointer in scalar registers
__device__ __forceinline__ unsigned long long uniform_addr(const void *p)
{
  const unsigned long long a = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ gcbytes uniform_gptr(const float *p) { return (gcbytes)uniform_addr(p); }
__device__ __forceinline__ gbytes uniform_gptr(float *p) { return (gbytes)uniform_addr(p); }

// One cell (NV consecutive floats) at a global byte address, as load_cell
// and store_cell do it at a generic one.
typedef float fvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
template<int NV> __device__ __forceinline__ void gload_cell(gcbytes p, float (&x)[NV])
{
  if constexpr(NV == 4) {
    const fvec4 a = *(FG2_GLOBAL const fvec4 *)p;
    x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
  } else if constexpr(NV == 6) {
    const fvec2 a = *(FG2_GLOBAL const fvec2 *)p, b = *(FG2_GLOBAL const fvec2 *)(p + 8),
                c = *(FG2_GLOBAL const fvec2 *)(p + 16);
    x[0] = a.x; x[1] = a.y; x[2] = b.x; x[3] = b.y; x[4] = c.x; x[5] = c.y;
  } else {
#pragma unroll
    for(int q = 0; q < NV; ++q) x[q] = ((FG2_GLOBAL const float *)p)[q];
  }
}
template<int NV> __device__ __forceinline__ void gstore_cell(gbytes p, const float (&x)[NV])
{
  if constexpr(NV == 4) {
    *(FG2_GLOBAL fvec4 *)p = fvec4{x[0], x[1], x[2], x[3]};
  } else if constexpr(NV == 6) {
    *(FG2_GLOBAL fvec2 *)p = fvec2{x[0], x[1]};
    *(FG2_GLOBAL fvec2 *)(p + 8) = fvec2{x[2], x[3]};
    *(FG2_GLOBAL fvec2 *)(p + 16) = fvec2{x[4], x[5]};
  } else {
#pragma unroll
    for(int q = 0; q < NV; ++q) ((FG2_GLOBAL float *)p)[q] = x[q];
  }
}

// A pointer known to point into global memory, as an address-space-1
// pointer: loads and stores through it are global (on the vector-memory
// wait counter only) even where the compiler cannot see where the pointer
// came from (a batch's run records hold the runs' pointers).
__device__ __forceinline__ gcbytes gptr(const float *p) { return (gcbytes)p; }
__device__ __forceinline__ gbytes gptr(float *p) { return (gbytes)p; }

// Value of x at lane (lane + 1) / (lane - 1) of the wavefront, by a DPP
// wave shift (a VALU move, no LDS round trip; lanes 63 / 0 read zero, and
// produce no output).
__device__ __forceinline__ float lane_next(float x)
{ return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, false)); }
__device__ __forceinline__ float lane_prev(float x)
{ return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, false)); }

// Stencil primitives of one variable q of the centre row (ring window w at
// offset p) into x.  Horizontal neighbours: +-1 by DPP wave shifts, +-2 and
// +-3 by ds_bpermute.
template<class M, int RING>
__device__ __forceinline__ void primitives(const float (&w)[RING][M::NV], const int p, const int q,
                                           const int (&addr)[2][3], typename M::P_t &x)
{
  const float cq = w[(p + 3) % RING][q];
#pragma unroll
  for(int k = 1; k <= 3; ++k) {
    const float a = w[(p + 3 + k) % RING][q], b = w[(p + 3 - k) % RING][q];
    x.dv[k - 1][q] = a - b;
    x.sv[k - 1][q] = a + b;
    const float e = k == 1 ? lane_next(cq) : lane_shift(cq, addr[0][k - 1]);
    const float f = k == 1 ? lane_prev(cq) : lane_shift(cq, addr[1][k - 1]);
    x.dh[k - 1][q] = e - f;
    x.sh[k - 1][q] = e + f;
  }
}

// The mixed-derivative differences of x (its dv already filled).
template<class M>
__device__ __forceinline__ void mixed(const int (&addr)[2][3], typename M::P_t &x)
{
#pragma unroll
  for(int k = 1; k <= 3; ++k)
#pragma unroll
    for(int d = 0; d < M::ND; ++d) {
      const float g = x.dv[k - 1][M::D12V[d]];    // f(i+k, .) - f(i-k, .)
      x.mx[k - 1][d] = k == 1 ? lane_next(g) - lane_prev(g)
                              : lane_shift(g, addr[0][k - 1]) - lane_shift(g, addr[1][k - 1]);
    }
}

template<class M, int STAGE, bool CFL>
__device__ __forceinline__ void process_tile(const float *__restrict__ uin, float *__restrict__ uout,
                                             float *__restrict__ v, const long s, const Params &P,
                                             const float beta, const float gdt,
                                             const int j0, const int i0, const int i1,
                                             const int cfl, TileOut &res)
{
  constexpr int NV = M::NV;
  // u rows in the register ring: rows i-3..i+3 in use and the next one or
  // two in flight.  A module whose right-hand side takes its derivatives
  // first keeps one row in flight (8 slots), which with the derivatives in
  // place of the primitives fits three wavefronts per SIMD (168 VGPRs).
  constexpr int RING = M::DERIV_FIRST ? 8 : 9;
  const int lane = threadIdx.x & 63;
  const int n1 = P.n1, n2 = P.n2;
  // lane holds column c; lanes 3..60 with c < n2 produce output
  const int c = j0 - HALF + lane;
  const bool out = lane >= HALF && lane < 64 - HALF && c < n2;
  float sgn;
  const int cc = (c < n2 + HALF) ? map_col(c, n2, P.p2, sgn) : (sgn = 1.0f, n2 - 1);
  // byte offset of the lane's column in a row: every load and store of the
  // tile is a row's uniform address plus this register.  Lanes with output
  // have cc == c; the others read a valid column, so that every load is
  // unconditional (a predicated load would end in a register join that
  // waits for it).
  const unsigned coff = (unsigned)cc * NV * 4u;
  int addr[2][3];                      // bpermute addresses for lane +k / -k
#pragma unroll
  for(int k = 1; k <= 3; ++k) {
    addr[0][k - 1] = ((lane + k) & 63) << 2;
    addr[1][k - 1] = ((lane - k) & 63) << 2;
  }
  const float k1a = 1.5859783962656393f / P.Delta1, k1b = 1.5859783962656393f / P.Delta2;

  // register ring: row r lives in slot (r - (i0 - 3)) % RING; row i + RING
  // - 4 is requested while row i is computed
  float w[RING][NV];
  // rows past i1 + 2 are never used; clamping keeps every load in bounds
  auto load_row = [&](int r, float (&dst)[NV]) {
    gload_cell<NV>(uniform_gptr(uin + (long)map_row(min(r, i1 + 2), n1, P.p1) * s) + coff, dst);
  };
#pragma unroll
  for(int k = 0; k < RING - 1; ++k) load_row(i0 - HALF + k, w[k]);
#pragma unroll
  for(int k = 0; k < 2 * HALF + 1; ++k) w[k][M::ODD] *= sgn;
  // v rows: a module whose right-hand side takes its derivatives first
  // requests row i's v as its iteration starts (the stencil and the
  // right-hand side cover the latency; one slot), the others keep a 3-slot
  // ring with row i+1 in flight (their substage kernels are HBM bound)
  constexpr int VR = M::DERIV_FIRST ? 1 : 3;
  static_assert(RING % VR == 0, "the v ring's slot is the iteration modulo VR");
  float vr[VR][NV];
  if constexpr(STAGE > 0 && VR > 1) gload_cell<NV>(uniform_gptr(v + (long)i0 * s) + coff, vr[0]);

  for(int base = i0; base < i1; base += RING) {
#pragma unroll
    for(int p = 0; p < RING; ++p) {
      const int i = base + p;
      if(i >= i1) break;
      // row i+3 enters use: apply the reflective sign of this lane's column
      if(i > i0) w[(p + 6) % RING][M::ODD] *= sgn;
      // loads complete in issue order for the wait counter: row i's v goes
      // first, so waiting for it leaves the u row requested next in flight
      if constexpr(STAGE > 0 && VR == 1)
        gload_cell<NV>(uniform_gptr(v + (long)i * s) + coff, vr[0]);
      load_row(i + RING - 4, w[(p + RING - 1) % RING]);
      if constexpr(STAGE > 0 && VR > 1)
        gload_cell<NV>(uniform_gptr(v + (long)min(i + 1, i1 - 1) * s) + coff, vr[(p + 1) % VR]);
      const float (&vo)[NV] = vr[p % VR];

      const float (&cen)[NV] = w[(p + 3) % RING];
      typename M::P_t x;
#pragma unroll
      for(int q = 0; q < NV; ++q) x.c[q] = cen[q];
      float f[NV];
      if constexpr(M::DERIV_FIRST) {
        // each variable's derivatives as soon as its stencil is read
        typename M::Deriv D;
#pragma unroll
        for(int q = 0; q < NV; ++q) {
          primitives<M, RING>(w, p, q, addr, x);
          M::deriv_var(P, x, q, D);
        }
        mixed<M>(addr, x);
        M::deriv_mixed(P, x, D);
        if(out) M::rhs_d(P, x.c, D, i, c, f);
      } else {
#pragma unroll
        for(int q = 0; q < NV; ++q) primitives<M, RING>(w, p, q, addr, x);
        mixed<M>(addr, x);
        if(out) M::rhs(P, x, i, c, f);
      }
      if(out) {
        float vn[NV], uo[NV];
#pragma unroll
        for(int q = 0; q < NV; ++q) {
          vn[q] = (STAGE == 0) ? f[q] : __builtin_fmaf(beta, vo[q], f[q]);
          uo[q] = __builtin_fmaf(gdt, vn[q], cen[q]);
          res.bad |= (uo[q] != uo[q]);
        }
        if constexpr(STAGE < 2) gstore_cell<NV>(uniform_gptr(v + (long)i * s) + coff, vn);
        gstore_cell<NV>(uniform_gptr(uout + (long)i * s) + coff, uo);
        if constexpr(STAGE == 2 && CFL) {
          if(cfl) {
            res.lam_adv = fmaxf(res.lam_adv, M::lam_adv(P, uo, i, c, k1a, k1b));
            if constexpr(M::CELL_DIFF) res.lam_diff = fmaxf(res.lam_diff, M::lam_diff(P, uo, i, c));
          }
        }
      }
    }
  }
}

__device__ __forceinline__ float wave_max(float x)
{
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o, 64));
  return x;
}

// Fold one tile's NaN flag and CFL rates into the run's diagnostics.
// (device-scope relaxed atomics, as atomicOr / atomicMax, through global
// pointers: the diagnostics live in device memory)
// Fold a non-negative double's bits into a running maximum in device
// memory (non-negative doubles order like their bit patterns): a relaxed
// load, and the atomic only when the value is larger, so that the blocks
// whose maximum is not (most of them) do not queue atomics on one address.
__device__ __forceinline__ void max_bits(unsigned long long *p, const double x)
{
  FG2_GLOBAL unsigned long long *g = (FG2_GLOBAL unsigned long long *)p;
  const unsigned long long v = (unsigned long long)__double_as_longlong(x);
  if(v > __hip_atomic_load(g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    __hip_atomic_fetch_max(g, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void commit_tile(const TileOut &r, const int cfl, DeviceDiag *diag)
{
  const unsigned long long badmask = __ballot(r.bad);
  const int lane = threadIdx.x & 63;
  if(badmask && lane == 0)
    __hip_atomic_fetch_or((FG2_GLOBAL unsigned int *)&diag->flags, 1u, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
  if(cfl) {
    const float la = wave_max(r.lam_adv), ld = wave_max(r.lam_diff);
    if(lane == 0) {
      max_bits(&diag->lam_adv_bits, (double)la);
      max_bits(&diag->lam_diff_bits, (double)ld);
    }
  }
}

// Wavefronts per SIMD the substage kernel is compiled for: three (at most
// 168 VGPRs) for a module whose right-hand side takes its derivatives first,
// except in the stage-2 variant that reduces CFL rates.
template<class M, bool CFL> constexpr int substage_waves() { return M::DERIV_FIRST && !CFL ? 3 : 1; }

// CFL: the stage-2 variant that reduces the CFL rates of its output when
// cfl is set (the others never do).
template<class M, int STAGE, int WPB, bool CFL>
__global__ __launch_bounds__(64 * WPB, (substage_waves<M, CFL>()))
void substage_kernel(const float *__restrict__ uin, float *__restrict__ uout, float *__restrict__ v,
                     const long s, const Params P, const float beta, const float gdt,
                     const int tile_h, const int nstrips, const int ntiles, const int cfl,
                     DeviceDiag *diag)
{
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tile = xcd_remap(blockIdx.x, gridDim.x) * WPB + wv;
  if(tile >= ntiles) return;
  const int strip = tile % nstrips, chunk = tile / nstrips;
  const int i0 = chunk * tile_h, i1 = min(P.n1, i0 + tile_h);
  TileOut r{0, 0.0f, 0.0f};
  process_tile<M, STAGE, CFL>(uin, uout, v, s, P, beta, gdt, strip * STRIP, i0, i1, cfl, r);
  commit_tile(r, CFL ? cfl : 0, diag);
}

// ---------------------------------------------------------------------------
// Whole RK3 step in one pass (hydro, hydro-isot): temporal blocking of the
// three substages.  A workgroup of NW wavefronts owns 64 NW consecutive columns
// (3 x 3 halo columns each side are recomputed, so 64 NW - 18 are output)
// and marches down rows [i0, i1).  At iteration it the workgroup runs
//   stage 0 on row r0 = i0 - 6 + it   (input u, from HBM, ring W0),
//   stage 1 on row r0 - 3              (input a = stage-0 output, ring W1),
//   stage 2 on row r0 - 6              (input b = stage-1 output, ring W2),
// each stage exactly the per-cell arithmetic of the substage kernel.  The
// stage v's travel in small register rings (V0, V1); only u is read and the
// new u written, so HBM traffic per step is 2 NVAR floats per cell.  (The
// one-barrier schedule, below, lags stages 1 and 2 by four rows each.)  Rows
// and columns outside the grid wrap (periodic), and the right-hand side sees the
// wrapped indices, so every halo value equals the value the substage path
// computes at the wrapped cell, bit for bit.  Horizontal neighbours of a
// stage's centre row come from LDS: one s_barrier per stage, or one per
// iteration for all three stages (step_one_barrier).  A last strip
// of at most 46 columns is packed: each wavefront of its workgroups marches
// its own row chunk on its own 64 lanes (step_tile_shape).
//
// Bounded runs (Neumann rows, p1 = 0, and / or reflecting columns, p2 = 0)
// take the same schedule with the reference's ghost fill applied to every
// stage's output instead of wrapping:
//   rows: a stage output at row -1 - r (r = 0, 1, 2) is a copy of its output
//     at row r, and at row n1 + k (k = 0, 1, 2) a copy of row n1 - 1 - k.
//     The copies are made in the register rings the moment the source row
//     is produced, which is before any stage reads them for a row that is
//     kept (rows outside the grid feed only rows that are themselves
//     replaced);
//   columns: a lane reads its horizontal neighbours (and their g_k) at the
//     mirrored column inside the grid, with the module's odd variable
//     negated, so lanes outside the grid are never read.
// Every value a kept cell sees is then the value of the substage path's
// index-mapped read, bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int pmod(int a, int n) { const int r = a % n; return r < 0 ? r + n : r; }
// Mirror of an index about the faces of [0, n), clamped into [0, n) for
// indices more than n outside (those feed only discarded values).
__device__ __forceinline__ int mirror(int a, int n)
{
  const int m = a < 0 ? -1 - a : a >= n ? 2 * n - 1 - a : a;
  return min(max(m, 0), n - 1);
}

enum StepMode { STEP_PERIODIC = 0, STEP_SLAB = 1, STEP_BOUNDED = 2 };

// Neumann ghost rows of a stage output in a register ring of 8 rows where
// row r (just produced) sits in slot S & 7 and row r - m in slot (S - m) & 7:
// row r = 0, 1, 2 is copied to row -1 - r (2r + 1 slots back), and row
// r = n1 + k (k = 0, 1, 2) is replaced by row n1 - 1 - k (2k + 1 slots back).
// Called from the unrolled row loop, where S is a constant, so every ring
// index folds to a register; the branches are uniform.
template<int NV>
__device__ __forceinline__ void ghost_rows(float (&W)[8][NV], const int S, const int r, const int n1)
{
  auto copy = [&](float (&dst)[NV], const float (&src)[NV]) {
#pragma unroll
    for(int q = 0; q < NV; ++q) dst[q] = src[q];
  };
  if(r == 0) copy(W[(S + 7) & 7], W[S & 7]);
  else if(r == 1) copy(W[(S + 5) & 7], W[S & 7]);
  else if(r == 2) copy(W[(S + 3) & 7], W[S & 7]);
  if(r == n1) copy(W[S & 7], W[(S + 7) & 7]);
  else if(r == n1 + 1) copy(W[S & 7], W[(S + 5) & 7]);
  else if(r == n1 + 2) copy(W[S & 7], W[(S + 3) & 7]);
}

// LDS rows of the one-pass step: each stage's centre row X and the
// differences g_k = f(i+k) - f(i-k) of its ND mixed-derivative variables G.
// For NV = 4, ND = 2 they are lane-major, so a lane reads a neighbour's
// four values (or two g_k) in one 16-byte (8-byte) LDS access: fewer LDS
// instructions, 2.6% on the 3584^2 hydro step.  For NV = 3 they stay
// variable-major: the wider reads would keep more registers live and cost
// hydro-isot its third wavefront per SIMD.
// NB = 2 buffers for the one-barrier schedule (iteration it writes buffer
// it & 1 while slower wavefronts may still read the other).
template<class M, int NW, int NB = 1> struct StepLds {
  static constexpr bool LANE_MAJOR = M::NV == 4 && M::ND == 2;
  struct Lane { float4 X[NB][3][64 * NW]; float2 G[NB][3][3][64 * NW]; };
  struct Var { float X[NB][3][M::NV][64 * NW]; float G[NB][3][3][M::ND][64 * NW]; };
  std::conditional_t<LANE_MAJOR, Lane, Var> r;
};

// Schedule of the one-pass step of module M.  One barrier per row
// iteration: stage 1 lags stage 0 and stage 2 lags stage 1 by four rows, so
// each stage reads only rows produced in earlier iterations, and all three
// publish their LDS rows (double-buffered) before one barrier; their LDS
// reads then overlap each other's arithmetic.  It needs more registers
// (longer-lived rows) and twice the LDS: the hydro step (two wavefronts per
// SIMD either way) gains 8% (214.8 against 233.8 us per 3584^2 step);
// hydro-isot would drop from three wavefronts per SIMD to two and lose 17%,
// and the batched kernels, whose run records take registers of their own,
// would lose occupancy, so they keep three barriers (one per stage).
template<class M> constexpr bool step_one_barrier() { return M::STEP_ONE_BARRIER; }

// One tile of the one-pass step: output rows [i0, i1), output columns
// [j0, j0 + WL - 18) on the WL = 64 NW lanes of the workgroup, or (a packed
// narrow strip: lane0 = 64 w, WL = 64) on wavefront w's own 64 lanes.  All
// threads of the workgroup call it together with the same nit0 >= i1 - i0 +
// 12 (iterations past a wavefront's own rows store nothing; the one-barrier
// schedule, ONE, runs two more).
// STEP_SLAB: the run is rows [H.row0, H.row0 + n1) of a grid of H.n1g rows;
// rows above and below it come from the halo buffers H.lo and H.hi instead
// of wrapping, and the right-hand side sees the global row index.
// STEP_BOUNDED: P.p1 / P.p2 select wrapped or mirrored rows / columns.
template<class M, int NW, int MODE = STEP_PERIODIC, bool ONE = false>
__device__ __forceinline__ void step_tile(const float *__restrict__ uin, float *__restrict__ uout,
                                          const long s, const Params &P, const float g0,
                                          const float g1, const float g2, const float b1,
                                          const float b2, const int j0, const int i0,
                                          const int i1, const int lane0, const int WL,
                                          const int nit0, const int cfl,
                                          StepLds<M, NW, ONE ? 2 : 1> &L, TileOut &res,
                                          const SlabHalo &H = SlabHalo{})
{
  constexpr int NV = M::NV, ND = M::ND, H3 = 3 * HALF;
  constexpr bool LM = StepLds<M, NW>::LANE_MAJOR;
  const int nit = nit0 + (ONE ? 2 : 0);
  auto &X = L.r.X;
  auto &G = L.r.G;
  constexpr bool SLAB = MODE == STEP_SLAB, BOUNDED = MODE == STEP_BOUNDED;
  const int t = threadIdx.x, tl = t - lane0;          // LDS lane, lane in the strip
  const int n1 = P.n1, n2 = P.n2;
  const int c = j0 - H3 + tl;
  const bool mirror_cols = BOUNDED && !P.p2, mirror_rows = BOUNDED && !P.p1;
  // this lane's column (lanes outside a reflecting grid load a mirrored
  // column; no kept value reads them)
  const int cc = mirror_cols ? mirror(c, n2) : pmod(c, n2);
  const bool out = tl >= H3 && tl < WL - H3 && c < n2;  // lanes of stage-2 output
  const float *col = uin + (long)cc * NV;
  const float k1a = 1.5859783962656393f / P.Delta1, k1b = 1.5859783962656393f / P.Delta2;
  // neighbour indices in LDS rows, inside the strip's lanes [lane0, lane0 +
  // WL) (clamped: lanes near the strip's edge compute garbage that no output
  // depends on); with reflecting columns the lane of the mirrored column,
  // and the sign of the odd variable there
  const int lo_lane = lane0, hi_lane = lane0 + WL - 1;
  int nb[2][3];
  float sg[2][3];
#pragma unroll
  for(int k = 1; k <= 3; ++k) {
    nb[0][k - 1] = min(t + k, hi_lane);
    nb[1][k - 1] = max(t - k, lo_lane);
    sg[0][k - 1] = sg[1][k - 1] = 1.0f;
    if(mirror_cols) {
      const int cp = c + k, cm = c - k;
      nb[0][k - 1] = min(max(mirror(cp, n2) - (j0 - H3) + lane0, lo_lane), hi_lane);
      nb[1][k - 1] = min(max(mirror(cm, n2) - (j0 - H3) + lane0, lo_lane), hi_lane);
      sg[0][k - 1] = (cp < 0 || cp >= n2) ? -1.0f : 1.0f;
      sg[1][k - 1] = (cm < 0 || cm >= n2) ? -1.0f : 1.0f;
    }
  }

  // The right-most strip of a grid that the strips do not divide needs only
  // lanes tl <= n2 - j0 + 17 (its last output column plus the 9-column halo).
  // Wavefronts wholly past that, and wavefronts of a packed narrow strip
  // without rows, skip the work and keep only the workgroup's barriers,
  // three (or one) per iteration as below.  The lanes next to them read
  // stale LDS, which only reaches lanes more than 9 columns past the last
  // output column.
  if(64 * __builtin_amdgcn_readfirstlane(t >> 6) - lane0 > n2 - j0 + 2 * H3 - 1 || i1 <= i0) {
    for(int it = 0; it < nit; ++it) {
      __syncthreads();
      if constexpr(!ONE) { __syncthreads(); __syncthreads(); }
    }
    return;
  }

  float W0[8][NV], W1[8][NV], W2[8][NV], V0[4][NV], V1[4][NV];
  auto load_row = [&](int r, float (&dst)[NV]) {
    if constexpr(SLAB) {
      // r >= i0 - 9 >= -9 always; the last prefetches ask for rows up to
      // i1 + 9 (i1 + 11 with one barrier), which no stage reads, so they
      // are clamped into the buffer
      const float *base = r < 0 ? H.lo : r >= n1 ? H.hi : uin;
      const int rr = r < 0 ? r + SLAB_HALO : r >= n1 ? min(r - n1, SLAB_HALO - 1) : r;
      gload_cell<NV>(gptr(base + (long)rr * s + (long)cc * NV), dst);
    } else if constexpr(BOUNDED) {
      gload_cell<NV>(gptr(col + (long)(mirror_rows ? mirror(r, n1) : pmod(r, n1)) * s), dst);
    } else {
      gload_cell<NV>(gptr(col + (long)pmod(r, n1) * s), dst);
    }
  };
#pragma unroll
  for(int k = 0; k < 7; ++k) load_row(i0 - H3 + k, W0[k]);

  // One stage on a centre row whose window rows are win(-3..3), in buffer
  // bf: publish writes the centre row and its g_k to LDS; after a barrier,
  // compute reads the neighbours and evaluates the right-hand side.
  auto publish = [&](const int bf, const int k, auto &&win) {
    const float (&cen)[NV] = win(0);
    if constexpr(LM) {
      X[bf][k][t] = make_float4(cen[0], cen[1], cen[2], cen[3]);
#pragma unroll
      for(int kk = 1; kk <= 3; ++kk)
        G[bf][k][kk - 1][t] = make_float2(win(kk)[M::D12V[0]] - win(-kk)[M::D12V[0]],
                                          win(kk)[M::D12V[1]] - win(-kk)[M::D12V[1]]);
    } else {
#pragma unroll
      for(int q = 0; q < NV; ++q) X[bf][k][q][t] = cen[q];
#pragma unroll
      for(int kk = 1; kk <= 3; ++kk)
#pragma unroll
        for(int d = 0; d < ND; ++d)
          G[bf][k][kk - 1][d][t] = win(kk)[M::D12V[d]] - win(-kk)[M::D12V[d]];
    }
  };
  auto compute = [&](const int bf, const int k, auto &&win, const int i,
                     const float (&vprev)[NV], const float beta, const float gdt,
                     float (&vn)[NV], float (&uo)[NV]) {
    const float (&cen)[NV] = win(0);
    typename M::P_t x;
#pragma unroll
    for(int q = 0; q < NV; ++q) x.c[q] = cen[q];
#pragma unroll
    for(int kk = 1; kk <= 3; ++kk) {
      // neighbour values: centre rows and g_k at lanes nb (read per
      // variable, or per lane from lane-major rows)
      auto xr = [&](int side, int q) -> float {
        if constexpr(LM) {
          const float4 v = X[bf][k][nb[side][kk - 1]];
          return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w;
        } else {
          return X[bf][k][q][nb[side][kk - 1]];
        }
      };
      auto gr = [&](int side, int d) -> float {
        if constexpr(LM) {
          const float2 v = G[bf][k][kk - 1][nb[side][kk - 1]];
          return d == 0 ? v.x : v.y;
        } else {
          return G[bf][k][kk - 1][d][nb[side][kk - 1]];
        }
      };
#pragma unroll
      for(int q = 0; q < NV; ++q) {
        const float a = win(kk)[q], b = win(-kk)[q];
        x.dv[kk - 1][q] = a - b;
        x.sv[kk - 1][q] = a + b;
        float e = xr(0, q), f = xr(1, q);
        if constexpr(BOUNDED) if(q == M::ODD) { e *= sg[0][kk - 1]; f *= sg[1][kk - 1]; }
        x.dh[kk - 1][q] = e - f;
        x.sh[kk - 1][q] = e + f;
      }
#pragma unroll
      for(int d = 0; d < ND; ++d) {
        float e = gr(0, d), f = gr(1, d);
        if constexpr(BOUNDED) if(M::D12V[d] == M::ODD) { e *= sg[0][kk - 1]; f *= sg[1][kk - 1]; }
        x.mx[kk - 1][d] = e - f;
      }
    }
    float f[NV];
    M::rhs(P, x, SLAB ? pmod(H.row0 + i, H.n1g) : pmod(i, n1), cc, f);
#pragma unroll
    for(int q = 0; q < NV; ++q) {
      vn[q] = (k == 0) ? f[q] : __builtin_fmaf(beta, vprev[q], f[q]);
      uo[q] = __builtin_fmaf(gdt, vn[q], cen[q]);
    }
  };
  // stage 2's output row r2: the new state
  auto store = [&](const int r2, const float (&uo)[NV]) {
    if(out && r2 < i1) {
#pragma unroll
      for(int q = 0; q < NV; ++q) res.bad |= (uo[q] != uo[q]);
      gstore_cell<NV>(gptr(uout + (long)r2 * s + (long)c * NV), uo);
      if(cfl) {
        res.lam_adv = fmaxf(res.lam_adv, M::lam_adv(P, uo, r2, c, k1a, k1b));
        if constexpr(M::CELL_DIFF) res.lam_diff = fmaxf(res.lam_diff, M::lam_diff(P, uo, r2, c));
      }
    }
  };

  // Row r of stage 0's output (W1) lives in slot (r - i0 + 6) & 7, of stage
  // 1's (W2) in slot (r - i0 + 11) & 7 (three barriers) or (r - i0 + 12) & 7
  // (one barrier); the v rings take one row per iteration, 4 slots.
  // Issue priority: wavefronts past a barrier, with right-hand sides ahead
  // of them, go before wavefronts still publishing their rows (measured
  // 0.6-0.9% on the 3584^2 hydro step with three barriers; the reverse order
  // costs 2%).
  for(int base = 0; base < nit; base += 8) {
#pragma unroll
    for(int p = 0; p < 8; ++p) {
      const int it = base + p;
      if(it >= nit) break;
      const int r0 = i0 - 2 * HALF + it;
      load_row(r0 + 4, W0[(p + 7) & 7]);
      auto w0 = [&](int d) -> float (&)[NV] { return W0[(p + 3 + d) & 7]; };
      // stage 0 on row r0, its output into W1 and V0
      auto stage0 = [&](const int bf) {
        float vn[NV], uo[NV];
        compute(bf, 0, w0, r0, V0[0], 0.0f, g0, vn, uo);
#pragma unroll
        for(int q = 0; q < NV; ++q) { W1[p & 7][q] = uo[q]; V0[p & 3][q] = vn[q]; }
        // Neumann rows: the output of row r0 lives in slot p & 7, that of
        // row r0 - m in slot (p - m) & 7
        if constexpr(BOUNDED) if(mirror_rows) ghost_rows(W1, p, r0, n1);
      };
      if constexpr(ONE) {
        // stage 1 on row r0 - 4, stage 2 on row r0 - 8: each reads rows the
        // stage before it produced in earlier iterations, so all three
        // publish into buffer p & 1 before the one barrier.  A wavefront
        // writes that buffer again two iterations on, after a barrier that
        // every wavefront passes only once its reads of it are done.  Each
        // stage reads its v slot before the stage before it refills the slot
        // (stage 2 first).
        const int bf = p & 1;
        auto w1 = [&](int d) -> float (&)[NV] { return W1[(p + 4 + d) & 7]; };
        auto w2 = [&](int d) -> float (&)[NV] { return W2[(p + 6 + d) & 7]; };
        const bool s1 = it >= 2 * HALF + 1, s2 = it >= 4 * HALF + 2;
        __builtin_amdgcn_s_setprio(0);
        publish(bf, 0, w0);
        if(s1) publish(bf, 1, w1);
        if(s2) publish(bf, 2, w2);
        __syncthreads();
        __builtin_amdgcn_s_setprio(1);
        if(s2) {
          float vn[NV], uo[NV];
          const int r2 = r0 - 2 * HALF - 2;
          compute(bf, 2, w2, r2, V1[(p + 2) & 3], b2, g2, vn, uo);
          store(r2, uo);
        }
        if(s1) {
          float vn[NV], uo[NV];
          compute(bf, 1, w1, r0 - HALF - 1, V0[p & 3], b1, g1, vn, uo);
#pragma unroll
          for(int q = 0; q < NV; ++q) { W2[(p + 2) & 7][q] = uo[q]; V1[(p + 2) & 3][q] = vn[q]; }
          // row r0 - 4 lives in slot (p + 2) & 7
          if constexpr(BOUNDED) if(mirror_rows) ghost_rows(W2, p + 2, r0 - HALF - 1, n1);
        }
        stage0(bf);
      } else {
        // stage 1 on row r0 - 3 and stage 2 on row r0 - 6 read the rows the
        // stage before them has just produced: a barrier per stage
        auto w1 = [&](int d) -> float (&)[NV] { return W1[(p + 5 + d) & 7]; };
        auto w2 = [&](int d) -> float (&)[NV] { return W2[(p + 7 + d) & 7]; };
        __builtin_amdgcn_s_setprio(0);
        publish(0, 0, w0);
        __syncthreads();
        __builtin_amdgcn_s_setprio(1);
        stage0(0);
        __builtin_amdgcn_s_setprio(0);
        if(it >= 2 * HALF) publish(0, 1, w1);
        __syncthreads();
        __builtin_amdgcn_s_setprio(1);
        if(it >= 2 * HALF) {
          float vn[NV], uo[NV];
          compute(0, 1, w1, r0 - HALF, V0[(p + 1) & 3], b1, g1, vn, uo);
#pragma unroll
          for(int q = 0; q < NV; ++q) { W2[(p + 2) & 7][q] = uo[q]; V1[(p + 2) & 3][q] = vn[q]; }
          // row r0 - 3 lives in slot (p + 2) & 7
          if constexpr(BOUNDED) if(mirror_rows) ghost_rows(W2, p + 2, r0 - HALF, n1);
        }
        __builtin_amdgcn_s_setprio(0);
        if(it >= 4 * HALF) publish(0, 2, w2);
        __syncthreads();
        __builtin_amdgcn_s_setprio(1);
        if(it >= 4 * HALF) {
          float vn[NV], uo[NV];
          const int r2 = r0 - 2 * HALF;
          compute(0, 2, w2, r2, V1[(p + 3) & 3], b2, g2, vn, uo);
          store(r2, uo);
        }
      }
    }
  }
}

// Tile `tile` of a single-run step launch (step_launch_tiles): nfull strips
// of 64 NW lanes times the row chunks, then the narrow last strip (at most
// 46 columns, when the launch packs it) with NW row chunks per workgroup,
// one to each wavefront, so it holds NW times fewer workgroup slots.
struct StepTileShape { int j0, i0, i1, lane0, wl, nit; };
template<int NW>
__device__ __forceinline__ StepTileShape step_tile_shape(const int tile, const int n1,
                                                         const int tile_h, const int nfull)
{
  constexpr int OUT = 64 * NW - 6 * HALF;
  const int nchunks = (n1 + tile_h - 1) / tile_h;
  StepTileShape T;
  if(tile < nfull * nchunks) {
    const int strip = tile % nfull, chunk = tile / nfull;
    T.j0 = strip * OUT;
    T.i0 = chunk * tile_h;
    T.i1 = min(n1, T.i0 + tile_h);
    T.lane0 = 0;
    T.wl = 64 * NW;
    T.nit = T.i1 - T.i0 + 4 * HALF;
  } else {
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int chunk = (tile - nfull * nchunks) * NW + w;
    T.j0 = nfull * OUT;
    T.i0 = chunk * tile_h;
    T.i1 = min(n1, T.i0 + tile_h);          // i1 <= i0: a wavefront without rows
    T.lane0 = 64 * w;
    T.wl = 64;
    T.nit = tile_h + 4 * HALF;              // the same for every wavefront
  }
  return T;
}

template<class M, int NW>
__global__ __launch_bounds__(64 * NW)
void step_kernel(const float *__restrict__ uin, float *__restrict__ uout, const long s,
                 const Params P, const float g0, const float g1, const float g2,
                 const float b1, const float b2, const int tile_h, const int nfull,
                 const int cfl, DeviceDiag *diag)
{
  constexpr bool ONE = step_one_barrier<M>();
  __shared__ StepLds<M, NW, ONE ? 2 : 1> L;
  const StepTileShape T = step_tile_shape<NW>(xcd_remap(blockIdx.x, gridDim.x), P.n1, tile_h,
                                              nfull);
  TileOut res{0, 0.0f, 0.0f};
  step_tile<M, NW, STEP_PERIODIC, ONE>(uin, uout, s, P, g0, g1, g2, b1, b2, T.j0, T.i0, T.i1,
                                       T.lane0, T.wl, T.nit, cfl, L, res);
  commit_tile(res, cfl, diag);
}

// The one-pass step of a slab: as step_kernel, with the rows outside the
// slab read from its halo buffers (same per-cell arithmetic, so a grid cut
// into slabs steps to the same bits as the whole grid).
template<class M, int NW>
__global__ __launch_bounds__(64 * NW)
void slab_step_kernel(const float *__restrict__ uin, float *__restrict__ uout, const long s,
                      const Params P, const float g0, const float g1, const float g2,
                      const float b1, const float b2, const int tile_h, const int nfull,
                      const int cfl, DeviceDiag *diag, const SlabHalo H)
{
  constexpr bool ONE = step_one_barrier<M>();
  __shared__ StepLds<M, NW, ONE ? 2 : 1> L;
  const StepTileShape T = step_tile_shape<NW>(xcd_remap(blockIdx.x, gridDim.x), P.n1, tile_h,
                                              nfull);
  TileOut res{0, 0.0f, 0.0f};
  step_tile<M, NW, STEP_SLAB, ONE>(uin, uout, s, P, g0, g1, g2, b1, b2, T.j0, T.i0, T.i1,
                                   T.lane0, T.wl, T.nit, cfl, L, res, H);
  commit_tile(res, cfl, diag);
}

// The one-pass step of a bounded run (Neumann rows and / or reflecting
// columns): as step_kernel, with the reference's ghost fill applied to every
// stage's output (same per-cell arithmetic as the substage path).
template<class M, int NW>
__global__ __launch_bounds__(64 * NW)
void bounded_step_kernel(const float *__restrict__ uin, float *__restrict__ uout, const long s,
                         const Params P, const float g0, const float g1, const float g2,
                         const float b1, const float b2, const int tile_h, const int nfull,
                         const int cfl, DeviceDiag *diag)
{
  // one barrier per row for every module: the mirrored indices and signs
  // keep the bounded kernels at two wavefronts per SIMD either way (the
  // hydro-isot 3584^2 bounded step: 211.4 against 219.0 us, the same bits)
  constexpr bool ONE = true;
  __shared__ StepLds<M, NW, ONE ? 2 : 1> L;
  const StepTileShape T = step_tile_shape<NW>(xcd_remap(blockIdx.x, gridDim.x), P.n1, tile_h,
                                              nfull);
  TileOut res{0, 0.0f, 0.0f};
  step_tile<M, NW, STEP_BOUNDED, ONE>(uin, uout, s, P, g0, g1, g2, b1, b2, T.j0, T.i0, T.i1,
                                      T.lane0, T.wl, T.nit, cfl, L, res);
  commit_tile(res, cfl, diag);
}

// Batched one-pass step: a persistent grid of workgroups pulls step tiles of
// many periodic runs from one atomic counter (same step_tile as above, so a
// run gives the same bits alone and in a batch).  CFL: the variant that
// reduces the rates of the runs whose records ask for it (the reduction
// costs the hydro-isot kernel its third wavefront per SIMD, so batches
// without CFL runs take the variant without it).
template<class M, int NW, bool CFL>
__global__ __launch_bounds__(64 * NW)
void batch_step_kernel(const BatchRun *__restrict__ runs, const int nruns, const int total_tiles,
                       int *counter)
{
  __shared__ StepLds<M, NW> L;
  __shared__ int next;
  for(;;) {
    __syncthreads();                       // previous tile's LDS reads are done
    if(threadIdx.x == 0) next = atomicAdd(counter, 1);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(next);
    if(t >= total_tiles) break;
    int lo = 0, hi = nruns - 1;
    while(lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if(runs[mid].tile_begin <= t) lo = mid; else hi = mid - 1;
    }
    const BatchRun &R = runs[lo];
    // the run's tiles as a single-run launch of it cuts them (nstrips strips
    // of 64 NW lanes, then a packed narrow strip)
    const StepTileShape T = step_tile_shape<NW>(t - R.tile_begin, R.P.n1, R.tile_h, R.nstrips);
    TileOut res{0, 0.0f, 0.0f};
    const int cfl = CFL ? R.cfl : 0;
    step_tile<M, NW>(R.uin, R.uout, R.s, R.P, R.g[0], R.g[1], R.g[2], (float)(-5.0 / 9.0),
                     (float)(-153.0 / 128.0), T.j0, T.i0, T.i1, T.lane0, T.wl, T.nit, cfl, L,
                     res);
    commit_tile(res, cfl, R.diag);
  }
}

// Batched substage: a persistent grid of workgroups pulls tiles of many runs
// from one atomic counter, WPB consecutive tiles per workgroup (one per
// wavefront) behind a barrier.  Every tile runs the same process_tile as
// the single-run kernel, so a run gives the same bits alone and in a batch.
// The fetch is per workgroup, behind barriers, because a per-wavefront
// fetch (lane 0's atomic broadcast by a lane shuffle) compiled to a loop
// whose lanes 1..63 went round again without lane 0 after a tile that
// reduces CFL rates, reading a stale tile index for ever.
template<class M, int STAGE, int WPB>
__global__ __launch_bounds__(64 * WPB)
void batch_kernel(const BatchRun *__restrict__ runs, const int nruns, const int total_tiles,
                  int *counter)
{
  __shared__ int next;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  for(;;) {
    __syncthreads();                       // every wavefront is past the previous fetch
    if(threadIdx.x == 0) next = atomicAdd(counter, WPB);
    __syncthreads();
    const int t0 = __builtin_amdgcn_readfirstlane(next);
    if(t0 >= total_tiles) break;
    const int t = t0 + w;
    if(t >= total_tiles) continue;         // past the last tile: only the barriers
    // binary search for the run owning tile t (tiles are ordered by run)
    int lo = 0, hi = nruns - 1;
    while(lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if(runs[mid].tile_begin <= t) lo = mid; else hi = mid - 1;
    }
    const int r = lo;
    const BatchRun &R = runs[r];
    const int lt = t - R.tile_begin;
    const int strip = lt % R.nstrips, chunk = lt / R.nstrips;
    const int i0 = chunk * R.tile_h, i1 = min(R.P.n1, i0 + R.tile_h);
    TileOut res{0, 0.0f, 0.0f};
    const int cfl = STAGE == 2 ? R.cfl : 0;
    process_tile<M, STAGE, true>(R.uin, R.uout, R.v, R.s, R.P, R.beta, R.gdt, strip * STRIP, i0, i1,
                           cfl, res);
    commit_tile(res, cfl, R.diag);
  }
}

// ---------------------------------------------------------------------------
// Plain path, ghost zones in memory (the reference's operation split).
// ---------------------------------------------------------------------------

// v <- F(u) + beta v, one thread per cell, stencil read from the ghosted
// array (src/kick.cu:59-106 computes the same from a shared-memory tile).
template<class M>
__global__ __launch_bounds__(256)
void kick_kernel(float *__restrict__ v, const float *__restrict__ u, const long s,
                 const Params P, const float beta)
{
  constexpr int NV = M::NV, ND = M::ND;
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if(j >= P.n2) return;
  const float *c = u + (long)i * s + (long)j * NV;
  typename M::P_t x;
#pragma unroll
  for(int q = 0; q < NV; ++q) x.c[q] = c[q];
#pragma unroll
  for(int k = 1; k <= 3; ++k) {
#pragma unroll
    for(int q = 0; q < NV; ++q) {
      const float a = c[k * s + q], b = c[-k * s + q];
      x.dv[k - 1][q] = a - b;
      x.sv[k - 1][q] = a + b;
      const float e = c[k * NV + q], f = c[-k * NV + q];
      x.dh[k - 1][q] = e - f;
      x.sh[k - 1][q] = e + f;
    }
#pragma unroll
    for(int d = 0; d < ND; ++d) {
      const int q = M::D12V[d];
      x.mx[k - 1][d] = (c[k * s + k * NV + q] - c[-k * s + k * NV + q])
                     - (c[k * s - k * NV + q] - c[-k * s - k * NV + q]);
    }
  }
  float f[NV];
  M::rhs(P, x, i, j, f);
  float *o = v + (long)i * s + (long)j * NV;
#pragma unroll
  for(int q = 0; q < NV; ++q) o[q] = __builtin_fmaf(beta, o[q], f[q]);
}

// u <- u + dt v with the sticky NaN flag (src/drift.cu:23-32).
template<int NV>
__global__ __launch_bounds__(256)
void drift_kernel(float *__restrict__ u, const float *__restrict__ v, const long s,
                  const int n2, const float dt, DeviceDiag *diag)
{
  const int i = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= n2 * NV) return;
  const long o = (long)i * s + k;
  const float x = __builtin_fmaf(dt, v[o], u[o]);
  u[o] = x;
  if(x != x) atomicOr(&diag->flags, 1u);
}

// Ghost rows, dim 1, all columns including dim-2 ghosts (src/bcond.cu:31-36,45-57).
template<int NV>
__global__ void bcond1_kernel(float *u, const long s, const int n1, const int n2, const int p1)
{
  const int w = (n2 + 2 * HALF) * NV;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;   // float within the row
  const int g = blockIdx.y;                              // ghost row 0..5
  if(k >= w) return;
  const int lo = g < HALF;
  const int kk = lo ? g : g - HALF;
  const int dst = lo ? -(kk + 1) : n1 + kk;
  const int src = lo ? (p1 ? n1 - (kk + 1) : kk) : (p1 ? kk : n1 - 1 - kk);
  float *row0 = u - HALF * NV;
  row0[(long)dst * s + k] = row0[(long)src * s + k];
}

// Ghost columns, dim 2, all rows including dim-1 ghosts (src/bcond.cu:38-43,59-75).
template<int NV, int ODD>
__global__ void bcond2_kernel(float *u, const long s, const int n1, const int n2, const int p2)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;   // (ghost col, var)
  const int r = blockIdx.y - HALF;                       // row -3 .. n1+2
  if(t >= 2 * HALF * NV) return;
  const int q = t % NV, g = t / NV;
  const int lo = g < HALF;
  const int kk = lo ? g : g - HALF;
  const int dst = lo ? -(kk + 1) : n2 + kk;
  const int src = lo ? (p2 ? n2 - (kk + 1) : kk) : (p2 ? kk : n2 - 1 - kk);
  float *row = u + (long)r * s;
  const float x = row[(long)src * NV + q];
  row[(long)dst * NV + q] = (!p2 && q == ODD) ? -x : x;
}

template<class M>
__global__ __launch_bounds__(256)
void euler_kernel(float *u, const long s, const Params P, const float dt)
{
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if(j >= P.n2) return;
  float *c = u + (long)i * s + (long)j * M::NV;
  M::first_order(P, c, dt, i, j);
}

// CFL rates of u for a state no fused stage-2 epilogue has reduced (the
// first step, or after an upload): columns across threads, rows strided
// over gridDim.y, four rows' loads in flight per thread, one block
// reduction, and an atomic only from a block that raises the maximum.  (max is exact, so the order of the
// rows does not change the result.)
template<class M>
__global__ __launch_bounds__(256)
void cfl_kernel(const float *u, const long s, const Params P, DeviceDiag *diag)
{
  constexpr int NV = M::NV;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const float k1a = 1.5859783962656393f / P.Delta1, k1b = 1.5859783962656393f / P.Delta2;
  float la = 0.0f, ld = 0.0f;
  auto fold = [&](const float (&c)[NV], int i) {
    la = fmaxf(la, M::lam_adv(P, c, i, j, k1a, k1b));
    if constexpr(M::CELL_DIFF) ld = fmaxf(ld, M::lam_diff(P, c, i, j));
  };
  if(j < P.n2) {
    const int gy = gridDim.y;
    int i = blockIdx.y;
    for(; i + 3 * gy < P.n1; i += 4 * gy) {
      float c[4][NV];
#pragma unroll
      for(int k = 0; k < 4; ++k) gload_cell<NV>(gptr(u + (long)(i + k * gy) * s + (long)j * NV), c[k]);
#pragma unroll
      for(int k = 0; k < 4; ++k) fold(c[k], i + k * gy);
    }
    for(; i < P.n1; i += gy) {
      float c[NV];
      gload_cell<NV>(gptr(u + (long)i * s + (long)j * NV), c);
      fold(c, i);
    }
  }
  la = wave_max(la);
  ld = wave_max(ld);
  __shared__ float red[2][4];
  const int w = threadIdx.x >> 6;
  if((threadIdx.x & 63) == 0) { red[0][w] = la; red[1][w] = ld; }
  __syncthreads();
  if(threadIdx.x == 0) {
    for(int k = 1; k < 4; ++k) { la = fmaxf(la, red[0][k]); ld = fmaxf(ld, red[1][k]); }
    max_bits(&diag->lam_adv_bits, (double)la);
    max_bits(&diag->lam_diff_bits, (double)ld);
  }
}

// FP64 sums per row (pass 1) and over rows in fixed order (pass 2).
// out[0] = sum rho, out[1] = sum rho u1, out[2] = sum rho u2,
// out[3] = sum rho Omg R^2 (disk; R = sphr sin theta), all times the cell area
// in index space (the caller scales).
template<class M>
__global__ __launch_bounds__(256)
void sums_row_kernel(const float *u, const long s, const Params P, double *partial)
{
  const int i = blockIdx.x;
  double acc[4] = {0, 0, 0, 0};
  for(int j = threadIdx.x; j < P.n2; j += blockDim.x) {
    M::sums(P, u + (long)i * s + (long)j * M::NV, i, j, acc);
  }
  __shared__ double red[4][256];
#pragma unroll
  for(int k = 0; k < 4; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for(int o = 128; o > 0; o >>= 1) {
    if((int)threadIdx.x < o)
#pragma unroll
      for(int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
    __syncthreads();
  }
  if(threadIdx.x < 4) partial[(long)threadIdx.x * P.n1 + i] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256)
void sums_final_kernel(const double *partial, const int n1, double *out)
{
  __shared__ double red[256];
  for(int k = 0; k < 4; ++k) {
    double a = 0;
    for(int i = threadIdx.x; i < n1; i += blockDim.x) a += partial[(long)k * n1 + i];
    red[threadIdx.x] = a;
    __syncthreads();
    for(int o = 128; o > 0; o >>= 1) {
      if((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if(threadIdx.x == 0) out[k] = red[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

hipError_t launch_bcond(int module, float *u, long s, const Params &P, int p1, int p2, hipStream_t st)
{
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    constexpr int NV = M::NV;
    const int w = (P.n2 + 2 * HALF) * NV;
    bcond1_kernel<NV><<<dim3((w + 255) / 256, 2 * HALF), 256, 0, st>>>(u, s, P.n1, P.n2, p1);
    bcond2_kernel<NV, M::ODD><<<dim3(1, P.n1 + 2 * HALF), 64, 0, st>>>(u, s, P.n1, P.n2, p2);
    return hipGetLastError();
  });
}

hipError_t launch_kick(int module, float *v, const float *u, long s, const Params &P, float beta,
                       hipStream_t st)
{
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    kick_kernel<M><<<dim3((P.n2 + 255) / 256, P.n1), 256, 0, st>>>(v, u, s, P, beta);
    return hipGetLastError();
  });
}

hipError_t launch_drift(int module, float *u, const float *v, long s, const Params &P, float dt,
                        DeviceDiag *diag, hipStream_t st)
{
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    const int n = P.n2 * M::NV;
    drift_kernel<M::NV><<<dim3((n + 255) / 256, P.n1), 256, 0, st>>>(u, v, s, P.n2, dt, diag);
    return hipGetLastError();
  });
}

// Per-row table (Params::rowtab) of a module that reads one: the module's
// row_entry() for every row.
template<class M>
__global__ __launch_bounds__(256)
void rows_kernel(float *tab, const Params P)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if(i >= P.n1) return;
  reinterpret_cast<float4 *>(tab)[i] = M::row_entry(P, i);
}

hipError_t launch_rows(int module, float *tab, const Params &P, hipStream_t st)
{
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    if constexpr(!M::ROW_TABLE) {
      return hipSuccess;
    } else {
      rows_kernel<M><<<(P.n1 + 255) / 256, 256, 0, st>>>(tab, P);
      return hipGetLastError();
    }
  });
}

hipError_t launch_euler(int module, float *u, long s, const Params &P, float dt, hipStream_t st)
{
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    if constexpr(!M::HAS_FIRST_ORDER) return hipSuccess;
    euler_kernel<M><<<dim3((P.n2 + 255) / 256, P.n1), 256, 0, st>>>(u, s, P, dt);
    return hipGetLastError();
  });
}

template<class M, int STAGE>
static hipError_t launch_stage(const float *uin, float *uout, float *v, long s, const Params &P,
                               float beta, float gdt, const FusedShape &sh, int cfl,
                               DeviceDiag *diag, hipStream_t st)
{
  const int nstrips = (P.n2 + STRIP - 1) / STRIP;
  const int nchunks = (P.n1 + sh.tile_h - 1) / sh.tile_h;
  const int ntiles = nstrips * nchunks;
  switch(sh.waves) {
#define FG2_W(WPB)                                                              \
  case WPB:                                                                    \
    if(STAGE == 2 && cfl)                                                      \
      substage_kernel<M, STAGE, WPB, STAGE == 2><<<(ntiles + WPB - 1) / WPB, 64 * WPB, 0, st>>>( \
        uin, uout, v, s, P, beta, gdt, sh.tile_h, nstrips, ntiles, cfl, diag);  \
    else                                                                       \
      substage_kernel<M, STAGE, WPB, false><<<(ntiles + WPB - 1) / WPB, 64 * WPB, 0, st>>>( \
        uin, uout, v, s, P, beta, gdt, sh.tile_h, nstrips, ntiles, 0, diag);    \
    break;
  FG2_W(1) FG2_W(2) FG2_W(4)
#undef FG2_W
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_substage(int module, int stage, const float *uin, float *uout, float *v, long s,
                           const Params &P, float beta, float gdt, const FusedShape &sh, int cfl,
                           DeviceDiag *diag, hipStream_t st)
{
  if(sh.tile_h < 1) return hipErrorInvalidValue;
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    switch(stage) {
    case 0: return launch_stage<M, 0>(uin, uout, v, s, P, beta, gdt, sh, cfl, diag, st);
    case 1: return launch_stage<M, 1>(uin, uout, v, s, P, beta, gdt, sh, cfl, diag, st);
    case 2: return launch_stage<M, 2>(uin, uout, v, s, P, beta, gdt, sh, cfl, diag, st);
    }
    return hipErrorInvalidValue;
  });
}

hipError_t launch_cfl(int module, const float *u, long s, const Params &P, DeviceDiag *diag,
                      hipStream_t st)
{
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    // about 4096 blocks: 12-13 rows per thread at 3584^2
    const int gx = (P.n2 + 255) / 256;
    const int gy = std::max(1, std::min(P.n1, 4096 / gx));
    cfl_kernel<M><<<dim3(gx, gy), 256, 0, st>>>(u, s, P, diag);
    return hipGetLastError();
  });
}

hipError_t launch_step(int module, const float *uin, float *uout, long s, const Params &P,
                       double dt, const FusedShape &sh, int cfl, DeviceDiag *diag, hipStream_t st,
                       const SlabHalo *slab)
{
  const bool bounded = !P.p1 || !P.p2;
  if(slab && (bounded || P.n1 < SLAB_HALO || !slab->lo || !slab->hi)) return hipErrorInvalidValue;
  const float g0 = (float)(dt * (1.0 / 3.0)), g1 = (float)(dt * (15.0 / 16.0)),
              g2 = (float)(dt * (8.0 / 15.0));
  const float b1 = (float)(-5.0 / 9.0), b2 = (float)(-153.0 / 128.0);
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    if constexpr(!M::ONE_PASS) {
      return hipErrorInvalidValue;   // disk runs take the substage path
    } else {
    const int tile_h = std::max(1, sh.tile_h);
    int nfull = 0;
    const long tiles = step_launch_tiles(P.n1, P.n2, tile_h, sh.waves, &nfull);
    if(tiles < 1 || tiles > INT_MAX) return hipErrorInvalidValue;
    const int grid = (int)tiles;
    switch(sh.waves) {
#define FG2_S(NW)                                                                            \
    case NW: {                                                                               \
      if(slab)                                                                               \
        slab_step_kernel<M, NW><<<grid, 64 * NW, 0, st>>>(                                   \
            uin, uout, s, P, g0, g1, g2, b1, b2, tile_h, nfull, cfl, diag, *slab);             \
      else if(bounded)                                                                       \
        bounded_step_kernel<M, NW><<<grid, 64 * NW, 0, st>>>(                                \
            uin, uout, s, P, g0, g1, g2, b1, b2, tile_h, nfull, cfl, diag);                    \
      else                                                                                   \
        step_kernel<M, NW><<<grid, 64 * NW, 0, st>>>(uin, uout, s, P, g0, g1, g2,            \
                                                   b1, b2, tile_h, nfull, cfl, diag);          \
      break;                                                                                 \
    }
    FG2_S(1) FG2_S(2) FG2_S(4)
#undef FG2_S
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
    }
  });
}

// Resident workgroups per CU of a kernel family (hardware occupancy from its
// register and LDS use): kind 0 = the substage kernels (the least of the
// three stages), kind 1 = the one-pass step kernel, kind 2 = the batched
// substage kernels, kind 3 = the batched step kernel, kind 4 = the one-pass
// step kernel of bounded runs, kind 5 = the batched step kernel with the
// CFL reduction.
hipError_t kernel_occupancy(int module, int kind, int waves, int *blocks_per_cu)
{
  *blocks_per_cu = 0;
  if(waves != 1 && waves != 2 && waves != 4) return hipErrorInvalidValue;
  return by_module(module, [&](auto m) -> hipError_t {
    using M = decltype(m);
    int nb = 0;
    hipError_t e = hipSuccess;
    auto take = [&](const void *fn) {
      int b = 0;
      if(e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, 64 * waves, 0);
      nb = nb ? std::min(nb, b) : b;
    };
    if(kind == 0 || kind == 2) {
#define FG2_W(WPB)                                                                 \
      if(waves == WPB && kind == 0) {                                              \
        take((const void *)substage_kernel<M, 0, WPB, false>);                     \
        take((const void *)substage_kernel<M, 1, WPB, false>);                     \
        take((const void *)substage_kernel<M, 2, WPB, false>);                     \
      }                                                                            \
      if(waves == WPB && kind == 2) {                                              \
        take((const void *)batch_kernel<M, 0, WPB>);                               \
        take((const void *)batch_kernel<M, 1, WPB>);                               \
        take((const void *)batch_kernel<M, 2, WPB>);                               \
      }
      FG2_W(1) FG2_W(2) FG2_W(4)
#undef FG2_W
    } else {
      if constexpr(!M::ONE_PASS) {
        return hipErrorInvalidValue;
      } else {
        if(kind == 1) {
          if(waves == 1) take((const void *)step_kernel<M, 1>);
          if(waves == 2) take((const void *)step_kernel<M, 2>);
          if(waves == 4) take((const void *)step_kernel<M, 4>);
        } else if(kind == 4) {
          if(waves == 1) take((const void *)bounded_step_kernel<M, 1>);
          if(waves == 2) take((const void *)bounded_step_kernel<M, 2>);
          if(waves == 4) take((const void *)bounded_step_kernel<M, 4>);
        } else if(kind == 3) {
          if(waves == 1) take((const void *)batch_step_kernel<M, 1, false>);
          if(waves == 2) take((const void *)batch_step_kernel<M, 2, false>);
          if(waves == 4) take((const void *)batch_step_kernel<M, 4, false>);
        } else {
          if(waves == 1) take((const void *)batch_step_kernel<M, 1, true>);
          if(waves == 2) take((const void *)batch_step_kernel<M, 2, true>);
          if(waves == 4) take((const void *)batch_step_kernel<M, 4, true>);
        }
      }
    }
    *blocks_per_cu = nb;
    return e;
  });
}

hipError_t launch_batch_step(int module, const BatchRun *runs_dev, int nruns, int total_tiles,
                             int waves, int grid, int *counter, int cfl, hipStream_t st)
{
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    if constexpr(!M::ONE_PASS) {
      return hipErrorInvalidValue;
    } else {
      switch(waves) {
#define FG2_S(NW)                                                                           \
      case NW:                                                                              \
        if(cfl)                                                                             \
          batch_step_kernel<M, NW, true><<<grid, 64 * NW, 0, st>>>(runs_dev, nruns,         \
                                                                   total_tiles, counter);  \
        else                                                                                \
          batch_step_kernel<M, NW, false><<<grid, 64 * NW, 0, st>>>(runs_dev, nruns,        \
                                                                    total_tiles, counter); \
        break;
      FG2_S(1) FG2_S(2) FG2_S(4)
#undef FG2_S
      default: return hipErrorInvalidValue;
      }
      return hipGetLastError();
    }
  });
}

hipError_t launch_sums(int module, const float *u, long s, const Params &P, double *partial,
                       double *out, hipStream_t st)
{
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    sums_row_kernel<M><<<P.n1, 256, 0, st>>>(u, s, P, partial);
    sums_final_kernel<<<1, 256, 0, st>>>(partial, P.n1, out);
    return hipGetLastError();
  });
}

template<class M, int STAGE>
static hipError_t launch_batch_stage(const BatchRun *runs, int nruns, int total, int waves,
                                     int grid, int *counter, hipStream_t st)
{
  switch(waves) {
#define FG2_W(WPB)                                                                 \
  case WPB:                                                                       \
    batch_kernel<M, STAGE, WPB><<<grid, 64 * WPB, 0, st>>>(runs, nruns, total, counter); \
    break;
  FG2_W(1) FG2_W(2) FG2_W(4)
#undef FG2_W
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_substage(int module, int stage, const BatchRun *runs, int nruns, int total,
                                 int waves, int grid, int *counter, hipStream_t st)
{
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    switch(stage) {
    case 0: return launch_batch_stage<M, 0>(runs, nruns, total, waves, grid, counter, st);
    case 1: return launch_batch_stage<M, 1>(runs, nruns, total, waves, grid, counter, st);
    case 2: return launch_batch_stage<M, 2>(runs, nruns, total, waves, grid, counter, st);
    }
    return hipErrorInvalidValue;
  });
}

// ---------------------------------------------------------------------------
// Batched Euler pre-step and the device clock
// ---------------------------------------------------------------------------

// first_order of every cell of the records with an active pre-step: block b
// belongs to the last record whose fo_begin <= b (fo_begin is the running
// sum of the blocks of the records before, so of equal values the last is
// the one with blocks), and each thread takes one cell, as euler_kernel
// does, with the same per-cell arithmetic.
template<class M>
__global__ __launch_bounds__(256)
void batch_euler_kernel(const BatchRun *__restrict__ runs, const int nruns)
{
  const int b = blockIdx.x;
  int lo = 0, hi = nruns - 1;
  while(lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if(runs[mid].fo_begin <= b) lo = mid; else hi = mid - 1;
  }
  const BatchRun &R = runs[lo];
  if(!R.fo || !(R.edt > 0.0f)) return;     // no pre-step, or a run standing still
  const long cell = (long)(b - R.fo_begin) * 256 + threadIdx.x;
  const long n = (long)R.P.n1 * R.P.n2;
  if(cell >= n) return;
  const int i = (int)(cell / R.P.n2), j = (int)(cell % R.P.n2);
  float *c = const_cast<float *>(R.uin) + (long)i * R.s + (long)j * M::NV;
  M::first_order(R.P, c, R.edt, i, j);
}

hipError_t launch_batch_euler(int module, const BatchRun *runs_dev, int nruns, int blocks,
                              hipStream_t st)
{
  if(blocks < 1) return hipSuccess;
  return by_module(module, [&](auto m) {
    using M = decltype(m);
    if constexpr(!M::HAS_FIRST_ORDER) return hipSuccess;
    batch_euler_kernel<M><<<blocks, 256, 0, st>>>(runs_dev, nruns);
    return hipGetLastError();
  });
}

__device__ __forceinline__ unsigned long long splitmix64_dev(unsigned long long &state)
{
  unsigned long long z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// One thread per run.  The dt arithmetic is the host's dt rule (FP64,
// IEEE division, the same order of operations), so a run takes the same
// steps whether its dt is chosen here or by fg2_getdt_cfl.
__global__ __launch_bounds__(256)
void clock_kernel(BatchRun *set0, BatchRun *set1, BatchRun *set2, const int nsets,
                  const int step, const int nruns, const int draw_phases)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= nruns) return;
  DeviceDiag *d = set0[i].diag;
  double rule = d->fixed_dt;
  if(!(rule > 0.0)) {
    constexpr double RK3_IMAG = 1.7320508075688772, RK3_REAL = 2.5127453266183286;
    const double la = __longlong_as_double((long long)d->lam_adv_bits);
    const double ld = d->lam_diff_const >= 0.0
                        ? d->lam_diff_const
                        : __longlong_as_double((long long)d->lam_diff_bits) * d->k2s;
    double x = INFINITY;
    if(la > 0.0) { const double y = RK3_IMAG / la; x = y < x ? y : x; }
    if(ld > 0.0) { const double y = RK3_REAL / ld; x = y < x ? y : x; }
    rule = d->c * x;
  }
  const double left = d->t_final - d->t;
  double dt = rule < left ? rule : left;
  float r1 = set0[i].P.rand1, r2 = set0[i].P.rand2;
  if(dt > 0.0) {
    const double tn = d->t + dt;
    d->t = tn < d->t_final ? tn : d->t_final;
    ++d->steps;
    if(draw_phases) {                     // the host's randomize(), per step taken
      const double u1 = (double)(splitmix64_dev(d->rng) >> 11) * 0x1.0p-53;
      const double u2 = (double)(splitmix64_dev(d->rng) >> 11) * 0x1.0p-53;
      r1 = (float)(6.2831853071795865 * u1);
      r2 = (float)(6.2831853071795865 * u2);
    }
  } else {
    // a rule of 0 (or less) before t_final would hold the run still for
    // ever: flag it (DeviceDiag flags bit 1) and keep the rule for the error
    if(left > 0.0) {
      atomicOr(&d->flags, 2u);
      d->dt = rule;
    }
    dt = 0.0;
  }
  if(dt > 0.0 || !(left > 0.0)) d->dt = dt;
  d->lam_adv_bits = 0;                    // the step's epilogue reduces the new state's rates
  d->lam_diff_bits = 0;
  constexpr double G[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
  BatchRun *sets[3] = {set0, set1, set2};
  for(int k = 0; k < nsets; ++k) {
    BatchRun &R = sets[k][i];
    R.P.rand1 = r1;
    R.P.rand2 = r2;
    R.edt = (float)dt;
    if(step) { R.g[0] = (float)(dt * G[0]); R.g[1] = (float)(dt * G[1]); R.g[2] = (float)(dt * G[2]); }
    else R.gdt = (float)(dt * G[k]);
  }
}

hipError_t launch_clock(int module, BatchRun *set0, BatchRun *set1, BatchRun *set2, int nsets,
                        int step, int nruns, hipStream_t st)
{
  if(nruns < 1) return hipSuccess;
  clock_kernel<<<(nruns + 255) / 256, 256, 0, st>>>(set0, set1, set2, nsets, step, nruns,
                                                     module_info(module).random_phases);
  return hipGetLastError();
}

} // namespace fg2
