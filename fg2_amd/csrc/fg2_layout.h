// fg2_layout.h -- storage layout and per-launch constants shared by the host
// driver (capi.hip) and the kernels (kernels.hip).
//
// Storage: AoS, NVAR floats per cell, rows along dim 1 (slow axis, n1),
// cells along dim 2 (fast axis, n2), as in src/setup.cpp:60-72 and the .raw
// order of src/io.cpp:56-76.  Each array keeps HALF = 3 ghost rows above and
// below and 3 ghost cells left and right, so the reference's pointer
// convention (a pointer to interior cell (0,0), ghosts at negative offsets,
// row stride s in floats) holds.  Unlike the reference's cudaMallocPitch
// layout, interior cell (0,0) of every row sits at a 256-byte boundary: the
// row starts with ROW_PAD floats of which the last 3*NVAR are the left ghosts.
#pragma once
#include <cstdint>

namespace fg2 {

constexpr int HALF    = 3;   // ghost width for the 6th-order stencil (src/fg2.h:26-34)
constexpr int ORDER   = 6;
constexpr int ROW_PAD = 64;  // floats before interior column 0 in each row (256 B)
constexpr int STRIP   = 64 - 2 * HALF;  // output columns of one wavefront tile

// Row stride in floats: ROW_PAD + (n2 + HALF) * nvar, rounded up to 64
// floats so that every row starts on a 256-byte boundary.
inline constexpr long row_stride(int n2, int nv)
{
  return ((long)ROW_PAD + (long)(n2 + HALF) * nv + 63) / 64 * 64;
}
// Floats per array: n1 + 2 HALF rows.
inline constexpr long array_floats(int n1, int n2, int nv)
{
  return (long)(n1 + 2 * HALF) * row_stride(n2, nv);
}
// Offset of interior cell (0,0) from the start of the allocation.
inline constexpr long interior_offset(int n2, int nv)
{
  return (long)HALF * row_stride(n2, nv) + ROW_PAD;
}

// Everything a kernel needs to evaluate the right-hand side at one cell.
// All values are FP32, as the reference keeps them in R-typed __constant__
// symbols (src/deriv.h:22-23, src/*/eqns.cuh).  Stencil coefficients come
// with 1/Delta folded in, computed on the host in FP64 and rounded once.
struct Params {
  // geometry
  int   n1, n2, p1, p2;
  float Delta1, Delta2;
  // D1: a1 (f(+1)-f(-1)) + a2 (f(+2)-f(-2)) + a3 (f(+3)-f(-3)), per axis
  float a1[3], a2[3];
  // D11: b0 f0 + b1 (f(+1)+f(-1)) + ... , per axis
  float b1[4], b2[4];
  // D12: m[k] * mix_k
  float m12[3];
  // D11 + D22 centre coefficient, b1[0] + b2[0]
  float blap0;
  // module parameters, in the module's own order (see physics.h)
  float par[16];
  // hydro-isot first_order random phases (src/hydro-isot/Euler.cuh:24-25)
  float rand1, rand2;
  // per-row table of a module that reads one (ROW_TABLE), 4 floats per row
  // from the module's row_entry(); disk: {sphr = rin e^{(i+1/2) Delta1},
  // 1/sphr, sphr^2, gravity}
  const float *rowtab;
};

// A slab: rows [row0, row0 + n1) of a grid of n1g rows, periodic along dim
// 1, whose other rows live in other runs (on this GPU or another).  The
// one-pass step of a slab reads the SLAB_HALO = 9 rows on either side (three
// substages of 3 rows) from two halo buffers in the state's own row layout:
// lo points at interior cell (0,0) of global row row0 - 9, hi at that of
// row row0 + n1 (rows wrap mod n1g), each followed by 8 more rows at the
// run's row stride.
constexpr int SLAB_HALO = 3 * HALF;
struct SlabHalo {
  const float *lo, *hi;
  int row0, n1g;
};

// Diagnostics written by the kernels.  flags bit 0: a NaN was produced
// (the sticky contain_nan of src/drift.cu:21,30).  An Inf is not flagged: as
// in the reference, only a NaN fails a step (src/step.cpp:38-40).  Bit 1:
// the device clock's dt rule gave no positive step before t_final.
struct DeviceDiag {
  unsigned int flags;
  unsigned int pad;
  unsigned long long lam_adv_bits;   // max advective rate, FP64 bits (>= 0)
  unsigned long long lam_diff_bits;  // max diffusive rate, FP64 bits (>= 0)
  // Clock of a run whose steps are queued without the host (batches,
  // hipGraph replays, served runs).  Before each such step the clock update
  // takes dt = min(t_final - t, rule), the rule being fixed_dt when it is
  // > 0 and otherwise c min(sqrt3 / lam_adv, 2.5127 / lam_diff) from the
  // rates above (lam_diff = lam_diff_const, or when that is < 0 the per-cell
  // rate times k2s), advances t, draws the hydro-isot forcing phases from
  // rng, and clears the rates for the step's epilogue.  dt = 0 once t has
  // reached t_final: the run then stands still.
  double t, t_final, dt;
  double fixed_dt, c, lam_diff_const, k2s;
  unsigned long long rng;            // SplitMix64 state of the forcing phases
  long long steps;                   // steps of dt > 0 taken
};

} // namespace fg2
