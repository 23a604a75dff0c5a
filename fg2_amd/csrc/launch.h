// launch.h -- host-side launchers of the HIP kernels (kernels.hip), called
// by the C-ABI driver (capi.hip).  All launches are asynchronous on the given
// stream and return the hipError_t of the launch.
#pragma once
#include <hip/hip_runtime.h>
#include "fg2_layout.h"

namespace fg2 {

// Plain path: one kernel per reference operation, ghost zones in memory.
hipError_t launch_bcond(int module, float *u, long s, const Params &P, int p1, int p2,
                        hipStream_t st);                                   // src/bcond.cu:77-86
hipError_t launch_kick(int module, float *v, const float *u, long s, const Params &P,
                       float beta, hipStream_t st);                        // src/kick.cu:108-114
hipError_t launch_drift(int module, float *u, const float *v, long s, const Params &P,
                        float dt, DeviceDiag *diag, hipStream_t st);       // src/drift.cu:34-39
hipError_t launch_euler(int module, float *u, long s, const Params &P, float dt,
                        hipStream_t st);                                   // src/kick.cu:128-138

// Fused path: one launch per RK substage.  Reads u_in (boundaries by index
// mapping, ghost zones unused) and v, writes v (stages 0 and 1) and u_out.
// Stage 2 optionally reduces the CFL rates of u_out into diag.
struct FusedShape {
  int tile_h;      // rows marched by one wavefront
  int waves;       // wavefronts per workgroup (each owns a 64-column strip tile)
};
hipError_t launch_substage(int module, int stage, const float *uin, float *uout, float *v,
                           long s, const Params &P, float beta, float gdt,
                           const FusedShape &shape, int cfl, DeviceDiag *diag,
                           hipStream_t st);

// Whole RK3 step in one launch (hydro, hydro-isot): the three substages
// temporally blocked; reads uin, writes uout, v unused.  waves = wavefronts
// per workgroup (1, 2 or 4); the workgroup recomputes 9 halo columns on each
// side.  Non-periodic dimensions (P.p1, P.p2) get the reference's ghost fill
// at every stage.  With slab != NULL the run is a slab of a taller periodic
// grid and the rows beyond it come from slab->lo / slab->hi.
hipError_t launch_step(int module, const float *uin, float *uout, long s, const Params &P,
                       double dt, const FusedShape &shape, int cfl, DeviceDiag *diag,
                       hipStream_t st, const SlabHalo *slab = nullptr);

// Workgroups of a single-run step launch with tiles of tile_h rows and
// `waves` wavefronts: *nfull strips of 64 waves - 18 output columns (the
// last one partial when the columns do not divide), each cut into row
// chunks; and when the columns leave a last strip of at most 46 columns and
// waves > 1, that narrow strip's chunks packed `waves` to a workgroup, one
// per wavefront.
inline long step_launch_tiles(int n1, int n2, int tile_h, int waves, int *nfull)
{
  const int out = 64 * waves - 6 * HALF, narrow = 64 - 6 * HALF;
  const long nchunks = (n1 + tile_h - 1) / tile_h;
  int full = n2 / out;
  const int rem = n2 - full * out;
  const bool pack = waves > 1 && rem > 0 && rem <= narrow;
  if(rem > 0 && !pack) ++full;
  if(nfull) *nfull = full;
  return (long)full * nchunks + (pack ? (nchunks + waves - 1) / waves : 0);
}

// Fill the per-row table (Params::rowtab, 4 floats per row) of a run of a
// module that reads one (ROW_TABLE: disk, from Delta1 and M, rS, rin, eps).
hipError_t launch_rows(int module, float *tab, const Params &P, hipStream_t st);

// Resident workgroups per CU (hipOccupancyMaxActiveBlocksPerMultiprocessor)
// with `waves` wavefronts per workgroup of the substage kernels (kind 0, the
// least over the three stages), the one-pass step kernel (kind 1), the
// batched substage kernels (kind 2), the batched step kernel (kind 3), the
// one-pass step kernel of bounded runs (kind 4) or the batched step kernel
// with the CFL reduction (kind 5).
hipError_t kernel_occupancy(int module, int kind, int waves, int *blocks_per_cu);

// CFL rates (FP64 max over cells of the advective and diffusive rates)
// of a state, folded into diag->lam_*_bits.
hipError_t launch_cfl(int module, const float *u, long s, const Params &P, DeviceDiag *diag,
                      hipStream_t st);

// Diagnostics: FP64 sums of density, momentum and (disk) angular momentum
// over the interior, written to out[0..3] (deterministic two-pass reduction;
// partial must hold 4 * n1 doubles).
hipError_t launch_sums(int module, const float *u, long s, const Params &P, double *partial,
                       double *out, hipStream_t st);

// Initial conditions on the device (ics.hip); par = the run's 16 FP32
// module parameters.  ic: the module's own number of a named condition
// (its ic_of(); 0 = zeros).
hipError_t launch_ic(int module, int ic, float *u, long s, int n1, int n2, float l1, float l2,
                     const float *par, unsigned long long seed, hipStream_t st);

// Spectral diagnostics (spectra.hip).  Status codes of the spectral_*
// functions:
enum { SPECTRA_OK = 0, SPECTRA_HIP = 1, SPECTRA_FFT = 2, SPECTRA_NOMEM = 3, SPECTRA_SHAPE = 4 };
struct SpectralWork;
int spectral_bins(int n1, int n2);               // shells below the Galerkin cutoff
int spectral_create(SpectralWork **out, int n1, int n2);
void spectral_free(SpectralWork *W);
// Shell-summed energy E and vorticity Z spectra and the wavevectors per
// shell of the velocity (variables 1, 2) of an AoS state; host outputs of
// spectral_bins() entries each (any may be NULL); synchronises the stream.
int spectral_shells(SpectralWork *W, const float *u, long s, int nv, double *E, double *Z,
                    unsigned long long *counts, hipStream_t st);
// Spectral vorticity field, n1 x n2 floats row-major, copied to `host`
// (may be NULL); synchronises the stream.
int spectral_vorticity(SpectralWork *W, const float *u, long s, int nv, float *host,
                       hipStream_t st);

// Batched path: many independent runs in one persistent launch.
struct BatchRun {
  const float *uin; float *uout; float *v;
  long s;
  Params P;
  int tile_h, ntiles, nstrips, tile_begin;  // this run's tiles are [tile_begin, tile_begin+ntiles)
  float beta, gdt;       // substage batches: this stage's beta and gamma dt
  float g[3];            // step batches: gamma_k dt of the three stages
  int cfl;               // reduce the CFL rates of the step's new state into diag
  int fo, fo_begin;      // Euler pre-step active; the run's first block in the batch's
                         // Euler launch (the blocks of the records before it)
  float edt;             // the Euler pre-step's dt
  DeviceDiag *diag;
};
// Batched Euler pre-step (first_order of every record with fo set; such a
// run of n1 x n2 cells takes ceil(n1 n2 / 256) blocks of 256 threads from
// fo_begin on; records with edt = 0 stand still).
hipError_t launch_batch_euler(int module, const BatchRun *runs_dev, int nruns, int blocks,
                              hipStream_t st);
// Clock update of the runs of a batched step whose dt the device takes
// (DeviceDiag's clock): one thread per run; writes dt into the nsets record
// sets (stage k of set k for substage batches, set 0 holding g[] for step
// batches), with the forcing phases and the Euler dt.  module: the batch's.
hipError_t launch_clock(int module, BatchRun *set0, BatchRun *set1, BatchRun *set2, int nsets,
                        int step, int nruns, hipStream_t st);
// Batched one-pass step (every run periodic, module hydro or hydro-isot):
// records hold uin/uout, the tile split and g[3]; one launch per step.
hipError_t launch_batch_step(int module, const BatchRun *runs_dev, int nruns, int total_tiles,
                             int waves, int grid, int *counter, int cfl, hipStream_t st);
hipError_t launch_batch_substage(int module, int stage, const BatchRun *runs_dev, int nruns,
                                 int total_tiles, int waves, int grid, int *counter,
                                 hipStream_t st);

} // namespace fg2
