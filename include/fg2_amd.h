/* fg2_amd.h -- C ABI of the MI355X (gfx950) implementation of fg2's
 * time-step path: kick, drift, boundary fill, first_order, RK3 step, time
 * step, the parameter registry, and batched execution of many runs.
 *
 * Every entry point returns an int status (FG2_OK = 0) unless stated.  All
 * device work of a run is ordered on the run's own HIP stream; fg2_step and
 * fg2_solve synchronise that stream once per step, as the reference's step()
 * does (src/step.cpp:36).  Pointers named u, v are device pointers to
 * interior cell (0,0) of an array in the run's layout (fg2_layout): NVAR
 * floats per cell (AoS), row stride `stride` floats, 3 ghost cells/rows on
 * every side at negative / past-the-end offsets -- the convention of the
 * reference's global::u / global::v (src/setup.cpp:60-72).
 *
 * Not reentrant per run: calls on one run must not overlap.  Different runs
 * may be driven from different host threads.
 */
#ifndef FG2_AMD_H
#define FG2_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fg2_run fg2_run;
typedef struct fg2_batch fg2_batch;

enum {
  FG2_OK        = 0,
  FG2_ERR_NAN   = 1,  /* the state holds a NaN after a step; the reference returns
                         cudaErrorInvalidValue here (src/step.cpp:38-40)            */
  FG2_ERR_HIP   = 2,  /* a HIP runtime error (src/step.cpp:42 returns the CUDA error) */
  FG2_ERR_SHAPE = 3,  /* grid or tile shape not supported                            */
  FG2_ERR_ARG   = 4,  /* bad argument: unknown module or parameter, NULL pointer     */
  FG2_ERR_NOMEM = 5   /* device allocation failed (setup() returns 0, src/setup.cpp:62-72) */
};

/* Step implementation used by fg2_step / fg2_solve. */
enum {
  FG2_PATH_FUSED    = 0, /* default: one launch per RK3 step (all three substages
                            temporally blocked, any boundaries) for hydro and
                            hydro-isot; one fused kick+drift launch per substage
                            for disk                                                  */
  FG2_PATH_PLAIN    = 1, /* bcond, kick, drift launches per substage (src/step.cpp:30-35) */
  FG2_PATH_SUBSTAGE = 2  /* one fused kick+drift launch per substage, always         */
};

/* Kernel of the fused path for hydro and hydro-isot (fg2_set_kernel). */
enum {
  FG2_KERNEL_AUTO     = 0, /* default: the one-pass step kernel, except that a grid
                              too small to fill most of the GPU with one-pass
                              tiles (hydro up to about 900^2, hydro-isot up to
                              about 1100^2) with an automatic tile takes the
                              substage kernels                                     */
  FG2_KERNEL_ONE_PASS = 1, /* the one-pass step kernel, whatever the grid size     */
  FG2_KERNEL_SUBSTAGE = 2  /* one fused kick+drift launch per substage             */
};

/* Time-step rule used by fg2_solve. */
enum {
  FG2_DT_CLOSED = 0,  /* c / (n1/l1 + n2/l2), src/getdt.cpp:26 (default)               */
  FG2_DT_CFL    = 1   /* FP64 min over cells of the advective and diffusive limits    */
};

/* Text of the last error on this host thread (never NULL). */
const char *fg2_last_error(void);

/* The physics modules the library is built with (the reference builds one,
 * chosen by its Makefile's -Isrc/<module>, Makefile:19-22): the count, and
 * for module k its name (the argument of fg2_setup), NVAR and its state
 * variables, comma-separated in storage order (src/<module>/state.h).
 * fg2_module returns FG2_ERR_ARG for k out of range; outputs may be NULL. */
int fg2_module_count(void);
int fg2_module(int k, const char **name, int *nvar, const char **fields);

/* ---- set-up: src/setup.cpp:32-87, src/config.cpp:24-41, src/para.cpp:25-40 ---- */

/* Allocate a run of `module` ("hydro", "hydro-isot" or "disk") on n1 x n2
 * cells on the current HIP device; replaces setup(n1, n2) (src/fg2.h:74,
 * src/setup.cpp:32).  *bytes (may be NULL) receives the device bytes
 * allocated.  Defaults as src/global.cpp:28-29: l1 = l2 = 1, c = 0.5,
 * periodic in both dimensions; module parameters at their reference values.
 * n1, n2 >= 3. */
int fg2_setup(fg2_run **run, const char *module, int n1, int n2, size_t *bytes);

/* Free a run (the reference frees at exit, src/setup.cpp:24-30). */
void fg2_free(fg2_run *run);

/* Geometry and boundary flags (global::l1, l2, c, p1, p2; src/fg2.h:54-56). */
int fg2_set_geometry(fg2_run *run, double l1, double l2, int p1, int p2, double c);
/* out[7] = l1, l2, p1, p2, c, Delta1, Delta2 */
int fg2_get_geometry(const fg2_run *run, double *out);

/* Apply the module's settings() and compute Delta = l / n; replaces config()
 * (src/fg2.h:75, src/config.cpp:24-41). */
int fg2_config(fg2_run *run);

/* "name=val" (leading dashes ignored) -> sets module parameter para_<name>;
 * returns "name = val" in a buffer owned by the run, or NULL if the module
 * has no such parameter.  Replaces para() (src/fg2.h:73, src/para.cpp:25-40). */
const char *fg2_para(fg2_run *run, const char *assignment);
int fg2_get_param(const fg2_run *run, const char *name, double *value);
/* Seed of the hydro-isot first_order random phases (libc rand() in
 * src/hydro-isot/Euler.cuh:39-44). */
int fg2_seed(fg2_run *run, unsigned long long seed);

/* Layout: NVAR, row stride in floats, and the interior pointers of the
 * current state u and the swap array v (global::u, global::v). */
int fg2_layout(const fg2_run *run, int *nvar, long *stride);
float *fg2_u(fg2_run *run);
float *fg2_v(fg2_run *run);
/* The run's HIP stream (hipStream_t). */
void *fg2_stream(fg2_run *run);

/* Host <-> device copies of the interior, n1*n2*NVAR floats in .raw order
 * (no ghosts); replace the cudaMemcpy2D of src/init.cpp:39-41 and
 * src/io.cpp:44-45,70-72. */
int fg2_upload(fg2_run *run, const float *host);
int fg2_download(fg2_run *run, float *host);
int fg2_upload_v(fg2_run *run, const float *host);
int fg2_download_v(fg2_run *run, float *host);

/* ---- spectral diagnostics (tools/spec.pro, tools/vis.pro) ---- */

/* Shell-integrated spectra of the state's velocity (variables 1 and 2) on
 * integer wavenumbers (tools/getk.pro:18-33): E[b] = sum over the shell of
 * (|u1^|^2 + |u2^|^2) / 2 (tools/spec.pro:25-31), Z[b] = sum of
 * |omega^|^2 / 2 with omega^ = i k1 u2^ - i k2 u1^ and the Nyquist
 * wavenumbers zeroed (tools/spec.pro:58-63), counts[b] = wavevectors in the
 * shell; u^ is the forward DFT divided by n1 n2, as IDL's fft.  Shell b
 * holds |k| in [b + 1/2, b + 3/2), shell 0 holds (0, 3/2), the last shell is
 * closed, b < n = ceil(min(n1, n2) / 2 + 0.99) (tools/oned.pro:18-44).
 * On input *nbins is the length of the arrays (each may be NULL); on
 * output it is n.  Too short arrays are FG2_ERR_ARG.  FP32 FFT (hipFFT),
 * FP64 shell sums.  Replaces the IDL post-processing of .raw frames. */
int fg2_spectrum(fg2_run *run, double *E, double *Z, unsigned long long *counts, int *nbins);

/* Spectral vorticity field of the state, real part of the inverse DFT of
 * omega^ (tools/vis.pro:57-63), n1*n2 floats row-major into `host`. */
int fg2_vorticity(fg2_run *run, float *host);

/* ---- initial conditions and .raw snapshots ---- */

/* Fill the state with the module's initial condition `name`, evaluated on
 * the device at the cell centres; replaces init(name) (src/fg2.h:70,
 * src/init.cpp:26-42, src/{hydro,hydro-isot,disk}/init.h).  Requires
 * fg2_config.  hydro: "blast", "bow", "implos", "KH", "Sod", anything else
 * zeros; hydro-isot: "Kf", anything else zeros; disk: the Hawley torus
 * whatever the name.  The rand() draws of KH and Kf come from a SplitMix64
 * stream seeded with `seed` (draw q = output q+1 from state seed, in the
 * reference's row-major draw order), mapped to ((z >> 11) + 1/2) 2^-53. */
int fg2_init(fg2_run *run, const char *name, unsigned long long seed);

/* Write the state as a .raw snapshot: int32[6] {n1, n2, NVAR, 4, p1, p2},
 * double time, float[3] {l1, l2, c}, then n1*n2*NVAR floats in row-major
 * AoS order; replaces dump(name, time) (src/fg2.h:72, src/io.cpp:56-76). */
int fg2_dump(fg2_run *run, const char *path, double time);

/* Read a .raw snapshot into the state and take p1, p2, l1, l2, c from it;
 * *time (may be NULL) receives its time.  A header whose n1, n2, NVAR or
 * word size differs from the run's is FG2_ERR_SHAPE ("input data is not
 * compatible with the setup"); replaces load(name) (src/fg2.h:71,
 * src/io.cpp:24-54).  As in the reference, call fg2_config afterwards. */
int fg2_load(fg2_run *run, const char *path, double *time);

/* ---- device operations, 1:1 with src/fg2.h:81-84 (asynchronous) ---- */

/* Point-wise first_order pre-step; replaces Euler(u, dt) (src/kick.cu:128-138). */
int fg2_euler(fg2_run *run, float *u, float dt);
/* Ghost fill, dim 1 then dim 2; replaces bcond(u, p1, p2) (src/bcond.cu:77-86). */
int fg2_bcond(fg2_run *run, float *u, int p1, int p2);
/* v <- F(u) + beta v on the interior (u's ghosts must be filled); replaces
 * kick(v, u, t, beta) (src/kick.cu:108-114).  t is unused, as in the reference. */
int fg2_kick(fg2_run *run, float *v, const float *u, float t, float beta);
/* u <- u + dt v on the interior, sticky NaN flag; replaces drift(u, v, dt)
 * (src/drift.cu:34-39). */
int fg2_drift(fg2_run *run, float *u, const float *v, float dt);

/* ---- integration: src/fg2.h:77-79 ---- */

/* Closed-form time step c / (n1/l1 + n2/l2); replaces getdt()
 * (src/getdt.cpp:22-29).  Returns the step (no status). */
double fg2_getdt(const fg2_run *run);
/* CFL time step of the current state: c * min(sqrt(3)/lambda_adv,
 * 2.5127/lambda_diff), the rates reduced over all cells in FP64 on the
 * device.  New: the reference leaves this as a TODO (src/getdt.cpp:26-28). */
int fg2_getdt_cfl(fg2_run *run, double *dt);

/* Path and time-step rule (FG2_PATH_*, FG2_DT_*). */
int fg2_set_path(fg2_run *run, int path);
int fg2_set_dt_mode(fg2_run *run, int mode);
/* Fused-kernel tile: rows per tile and wavefronts per workgroup (1, 2 or
 * 4; the one-pass step kernel gives each workgroup 64 waves - 18 output
 * columns, and a last strip of at most 46 columns one wavefront per row
 * chunk; the substage kernel gives each wavefront 58).  With the automatic
 * kernel choice, a tile set here steps a module that has the one-pass
 * kernel with that kernel; tile_h = 0 restores the automatic shape, which
 * steps a grid too small to fill most of the GPU with one-pass tiles
 * (hydro up to about 900^2) with the substage kernels and short tiles. */
int fg2_set_tile(fg2_run *run, int tile_h, int waves);
int fg2_get_tile(const fg2_run *run, int *tile_h, int *waves);
/* Fix the fused path's kernel of a hydro or hydro-isot run (FG2_KERNEL_*;
 * FG2_KERNEL_ONE_PASS on disk is FG2_ERR_ARG).  New: the reference's step
 * has one form (src/step.cpp:22-43).  A single small run steps faster on
 * the substage kernels; many small runs launched side by side on their own
 * streams go faster on the one-pass kernel, one launch per step. */
int fg2_set_kernel(fg2_run *run, int kernel);
/* The kernel a step of the run takes now: FG2_KERNEL_ONE_PASS or
 * FG2_KERNEL_SUBSTAGE (disk, FG2_PATH_SUBSTAGE, or the automatic choice on
 * a small grid); 0 on the plain path.  New, as fg2_set_kernel. */
int fg2_get_kernel(const fg2_run *run, int *kernel);
/* Time the run's fused kernel on its own state for candidate tile shapes
 * and keep the fastest (cached per module and grid size for the process);
 * replaces the block-size heuristic of setup() (src/setup.cpp:39-58).  With
 * the automatic kernel choice, a grid that it steps with the substage
 * kernels has the one-pass kernel's shape timed too, and the faster kernel
 * becomes the run's kernel (fg2_set_kernel).  The state is unchanged;
 * *tile_h, *waves (may be NULL) receive the choice. */
int fg2_autotune(fg2_run *run, int *tile_h, int *waves);

/* One low-storage RK3 step of size dt at time t: first_order, then three
 * substages; synchronises the run's stream and checks the NaN flag.
 * Replaces step(t, dt) (src/step.cpp:22-43): returns FG2_OK, FG2_ERR_NAN
 * or FG2_ERR_HIP. */
int fg2_step(fg2_run *run, double t, double dt);
/* Same as fg2_step without the synchronisation and the NaN check, so that
 * many steps can be queued or captured in a HIP graph; check the flag
 * later with fg2_check. */
int fg2_step_async(fg2_run *run, double t, double dt);
/* One fused substage (0, 1 or 2) of an RK3 step of size dt: reads the
 * state and v, writes v (stages 0, 1) and the new state
 * (src/step.cpp:30-35 does bcond, kick, drift per substage). */
int fg2_substage(fg2_run *run, int stage, double dt);
/* Synchronise the run's stream; FG2_ERR_NAN if any step since the last
 * check produced a NaN (the flag is then cleared). */
int fg2_check(fg2_run *run);

/* Advance from *t to t_final with dt = min(t_final - t, getdt) per step, as
 * the inner loop of solve() (src/solve.cpp:37-49); *t and *steps (may be
 * NULL) receive the final time and the number of steps.  On a NaN the state
 * of the failing step is kept and FG2_ERR_NAN is returned (the reference
 * dumps it, src/solve.cpp:41-46). */
int fg2_solve(fg2_run *run, double *t, double t_final, int *steps);

/* ---- slabs: one grid cut along dim 1 over several runs (GPUs) ----
 * No counterpart in the reference, which runs one grid on one GPU
 * (src/setup.cpp:60-72).  For callers that shard a periodic hydro or
 * hydro-isot grid by rows, one slab per GPU.  A slab's one-pass step reads
 * FG2_SLAB_HALO rows on either side from two halo buffers the run owns;
 * each step the caller fills them from the neighbouring slabs (RCCL
 * send/recv, peer copies, or fg2_slab_exchange when all slabs live in one
 * process).  A slab steps to the same bits as the whole grid. */
#define FG2_SLAB_HALO 9
/* Make the run rows [row0, row0 + n1) of a periodic grid of n1_global rows
 * (n1_global = 0: a whole grid again).  Needs n1 >= FG2_SLAB_HALO.  After
 * this, l1 (fg2_set_geometry) is the length of the whole grid, and getdt,
 * Delta1 and the hydro-isot forcing use the whole grid's rows.  fg2_step,
 * fg2_solve, fg2_substage and batches refuse a slab. */
int fg2_slab(fg2_run *run, int row0, int n1_global);
/* Device byte ranges of one halo exchange, each `bytes` long and contiguous
 * (FG2_SLAB_HALO rows in the run's layout, row padding included):
 * recv_lo <- the last rows of the slab above, recv_hi <- the first rows of
 * the slab below; send_first / send_last are this slab's first / last rows
 * in the current state (they move with every step).  Any may be NULL. */
int fg2_slab_halo(fg2_run *run, float **recv_lo, float **recv_hi, const float **send_first,
                  const float **send_last, size_t *bytes);
/* Fill the halos of runs[0..nslabs) (consecutive slabs of one grid, in
 * order, covering it; one device) from their neighbours with device copies
 * on the receivers' streams, ordered after every slab's queued work. */
int fg2_slab_exchange(fg2_run *const *runs, int nslabs);
/* One RK3 step of size dt of a slab from its current halos, queued on the
 * run's stream without a synchronisation (check NaN with fg2_check).  The
 * CFL rates (FG2_DT_CFL) cover the slab's own cells: the grid's dt is the
 * least over the slabs.  Hydro-isot stochastic forcing is refused. */
int fg2_slab_step(fg2_run *run, double dt);

/* FP64 diagnostics of the current state: out[0] = sum rho, out[1..2] =
 * sum rho u1, sum rho u2 (disk: rho ur, rho uz), out[3] = sum rho Omg R^2
 * (disk), each times the cell volume element (index-space for the
 * Cartesian modules, r^3 sin(theta) for disk). */
int fg2_sums(fg2_run *run, double *out);

/* Perf model of the reference (src/config.cpp:30-36): flop and bits per
 * step; and the algorithmic HBM bytes per cell per step of the fused path,
 * 10 NVAR sizeof(float). */
int fg2_perf_model(const fg2_run *run, double *flops, double *bits, double *alg_bytes_per_cell);

/* ---- batched execution: many independent runs advanced together ---- */

/* A batch of runs of one module.  Each fg2_batch_step advances every run of
 * the batch by one RK3 step of its own dt (dts[k]) with persistent launches
 * for the whole batch: one (the one-pass step, when every run is periodic
 * hydro or hydro-isot) or three (one per substage: disk, or a batch with a
 * Neumann or reflecting run), after one batched launch of the Euler
 * pre-step of the runs that have one (hydro-isot stochastic forcing, f k >
 * 0; disk radiation equilibrium, ff != 0; src/kick.cu:128-138).  The runs
 * keep their own state; a run gives the same bits alone (fg2_step, fused
 * path) and in a batch. */
int fg2_batch_create(fg2_batch **batch, fg2_run *const *runs, int nruns, int waves, int tile_h);
void fg2_batch_free(fg2_batch *batch);
int fg2_batch_step_async(fg2_batch *batch, const double *dts);
int fg2_batch_step(fg2_batch *batch, const double *dts);
/* Capture `steps` batch steps into a HIP graph and replay it `replays`
 * times (an odd `steps` only once).  The graph is kept: a later call with
 * the same `steps`, while the batch's workgroup shape and tiles are
 * unchanged, replays it without a new capture.  The per-run launch records
 * (state buffers, dts) are refilled on every call.  Forced hydro-isot runs
 * (f k > 0) draw their phases on the device. */
int fg2_batch_graph_run(fg2_batch *batch, const double *dts, int steps, int replays);
/* Advance run k of the batch from t[k] to t_final[k], as fg2_solve advances
 * one run (src/solve.cpp:37-49: dt = min(t_final - t, getdt()), t = min(t +
 * dt, t_final)), with every dt chosen on the device: each run's dt mode
 * (fg2_set_dt_mode) -- the closed form, or the FP64 CFL rule from the rates
 * its previous step's epilogue reduced -- so no step waits for the host.
 * Steps go in hipGraphs of 16 (kept across calls) and the host reads the
 * runs' times once per graph; a run that reaches its end first takes steps
 * of dt = 0, which leave its state as it is, until all have, or until
 * max_steps steps.  t[k] receives the time reached, steps[k] (may be NULL)
 * the steps taken.  A run gives the same bits, times and step count here
 * as through fg2_solve.  Returns FG2_OK or FG2_ERR_NAN; FG2_ERR_ARG for a
 * run whose dt rule gives no positive step before its end time. */
int fg2_batch_solve(fg2_batch *batch, double *t, const double *t_final, int max_steps, int *steps);
void *fg2_batch_stream(fg2_batch *batch);
/* Choose the batch's tile height and workgroup shape by timing batched
 * steps of size dts on the runs' states (states unchanged).  Each height h
 * is timed with every run cut into h-row tiles and with every run cut into
 * the same number of equal tiles; *tile_h receives h. */
int fg2_batch_autotune(fg2_batch *batch, const double *dts, int *tile_h, int *waves);

/* One synchronous batched step of runs[0..nruns), as fg2_batch_step of a
 * batch made by fg2_batch_create(runs, nruns, waves, tile_h) -- same
 * kernels, same tiles, same bits -- that allocates nothing: no hipMalloc,
 * hipFree or other allocating call, on any call.  The launch records and
 * tile counters live in a fixed device array of the library (about 400 KiB
 * per GPU, held until the process ends), so nruns is at most
 * FG2_STATIC_MAX_RUNS (more is FG2_ERR_SHAPE).  The runs' states stay in
 * their own buffers.  Runs in the current device; the launch goes on
 * runs[0]'s stream after every run's queued work, and the call returns
 * after the step with FG2_OK or FG2_ERR_NAN as fg2_batch_step.  Calls that
 * overlap on one GPU share the array and race (one process serialises its
 * own calls).  No counterpart in the reference, which has no batching: for
 * callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). */
#define FG2_STATIC_MAX_RUNS 512
int fg2_batch_step_static(fg2_run *const *runs, int nruns, const double *dts, int waves,
                          int tile_h);

/* ---- serving: continuous batching of runs of different sizes ---- */

typedef struct fg2_server fg2_server;

/* A server that keeps admitted runs' state on the current device in an
 * arena of hbm_budget bytes (allocated once; a run holds two state arrays)
 * and advances all active runs together each tick (one batched one-pass
 * launch per module, after the batched Euler pre-step of the runs that have
 * one; served runs are periodic hydro or hydro-isot).  waves, tile_h: the
 * batched kernel's wavefronts per workgroup and tile height; waves = 0
 * takes two-wavefront workgroups and tile_h = 0 chooses the height per
 * launch from the active runs (rounds of resident workgroups times the
 * tile's cost). */
int fg2_server_create(fg2_server **server, size_t hbm_budget, int waves, int tile_h);
void fg2_server_free(fg2_server *server);
/* Queue a request of `steps` RK3 steps (closed-form dt per step) from the
 * initial state init (n1*n2*NVAR floats, .raw order, copied); `now` is the
 * caller's clock, used for the latency record. */
int fg2_server_submit(fg2_server *server, const char *module, int n1, int n2, int steps,
                      const float *init, double now, long long *id);
/* Queue a request shaped like the configured run `like`: its module and
 * grid, geometry (periodic only), parameters (fg2_para), dt mode
 * (fg2_set_dt_mode: the closed form, or the FP64 CFL rule chosen on the
 * device each step) and forcing stream (fg2_seed).  Only like's host-side
 * settings are read.  The result is the same bits as `steps` fg2_step calls
 * on a run set up like `like`, with dt = fg2_getdt or fg2_getdt_cfl before
 * each.  Per-run mode serves only fixed-dt runs without a pre-step. */
int fg2_server_submit_run(fg2_server *server, const fg2_run *like, int steps, const float *init,
                          double now, long long *id);
/* Retire finished runs, admit queued requests that fit the budget, and
 * launch one step of every active run.  Host copies of states into and out
 * of the server's staging buffers are spread over ticks (at most 8 MiB per
 * tick), so a large request takes several ticks to upload or to deliver.
 * Counts of active, queued (waiting or uploading) and finished requests
 * are returned (pointers may be NULL). */
int fg2_server_tick(fg2_server *server, double now, int *active, int *queued, int *finished);
/* times[3] = arrival, admission, completion (caller's clock; -1 if not yet);
 * *status = -1 while running, else FG2_OK / FG2_ERR_NAN; out (may be NULL)
 * receives the final state once finished. */
int fg2_server_result(fg2_server *server, long long id, float *out, double *times, int *status);
/* Total cell-updates launched so far and the device bytes held by runs. */
int fg2_server_stats(fg2_server *server, long long *cell_steps, size_t *bytes_in_use);

/* How a tick advances the active runs: FG2_SERVE_BATCHED (default), one
 * batched one-pass launch per module; FG2_SERVE_PER_RUN, each run's own
 * one-pass step launch (the one-pass kernel's automatic shape for a run of
 * its size, as with FG2_KERNEL_ONE_PASS) on one of 8 streams;
 * FG2_SERVE_PER_RUN_PLAIN, each run's plain kernels (ghost
 * fill, kick, drift per substage, the reference's operation split, in place
 * on the run's state with its second array as v) on one of 8 streams.  The
 * per-run modes are the baselines the batched mode is measured against,
 * with the same admission, arena and copies, and give the same bits.
 * Changes only while idle. */
enum { FG2_SERVE_BATCHED = 0, FG2_SERVE_PER_RUN = 1, FG2_SERVE_PER_RUN_PLAIN = 2 };
int fg2_server_set_mode(fg2_server *server, int mode);

#ifdef __cplusplus
}
#endif

#endif /* FG2_AMD_H */
