"""Benchmark of the fused RK3 time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--no-batch] [--no-serve] [--no-disk] [--no-spectrum] [--no-slab]
                    [--no-bounded] [--no-config1] [--no-sweep] [--disk-only]
                    [--slab-only] [--bounded-only] [--headline-only] [--batch-only]
                    [--serve-only]

Workload (the `value`): hydro module, Kelvin-Helmholtz shear layer on a
3584 x 3584 periodic grid, FP32, one launch of the one-pass step kernel per
RK3 step (the three substages temporally blocked), fixed dt from the CFL
limit of the initial state.  Inputs are resident in HBM before the timed
region.  Prints one JSON line with the whole-job cell-updates/s, the
roofline of the step kernel, and beside it: a served request stream (and
the server's capacity, batched against per-run launches), 64 batched
hydro-isot runs, the disk module on 2048 x 2048 (Neumann / reflecting
boundaries, three substage launches per step), the energy and vorticity
spectra of the 3584 x 3584 state, the KH grid with Neumann rows and
reflecting columns, hydro 256^2, a hydro-isot resolution sweep at the
reference's benchmark sizes, the KH grid cut into 8 slabs of 448 rows (one
GPU's share, with its halo exchange), the substage path on the headline
workload, and the host CPU baseline (the oracle's FP32 instantiation on all
host cores, on a bounded sample of the same workload).

The other GPU legs run before the headline, and the headline's own
substage-path leg (100 steps of the same workload) is queued back to back
with its warm-up and timed steps: after 5 ms or more without work the step
takes 250-290 us per launch (measured with one barrier per stage), and it
returns to its steady launch time only after some 100 launches (25 ms) of
sustained load (the GPU's clocks ramp up), so a headline of 20 steps after
an idle gap would time the ramp.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
N = 3584
METRIC = "cell-updates/s per RK3 step + % HBM peak; p50/p99 run latency (batched)"


class HipEvents:
    """hipEvent timing on an explicit stream (the run's own stream)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float),
                                                 ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value

    def sync(self, stream):
        assert self.hip.hipStreamSynchronize(ctypes.c_void_p(stream)) == 0


VALU_PEAK = 256 * 4 * 2.4e9 / 4   # wave64 FP32 VALU instructions/s: 4 cycles on a 16-lane SIMD


def load_counters():
    """HBM bytes and VALU wave-instructions per launch of the time-step
    kernel, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by
    the copy-kernel calibration) and SQ_INSTS_VALU counts of this workload
    (None where the summary is absent)."""
    path = os.path.join(ROOT, "profiles", "step_kernel_hydro_3584_pmc.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch"), d.get("valu_insts_per_launch")


def cpu_baseline(budget_s=12.0):
    """The oracle's FP32 instantiation (OpenMP over rows, all host cores) on
    the same 3584^2 KH workload, for as many steps as fit in budget_s (about
    12 s of host work)."""
    from oracle.oracle import Oracle
    from fg2_amd import ics
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    x = ics.hydro_kh(N, N, seed=2)
    o = Oracle("hydro", N, N, fp64=False, geometry=dict(l1=1.0, l2=1.0, p1=1, p2=1, c=0.5))
    o.upload(x)
    dt = o.getdt_cfl()
    o.step(0.0, dt)             # warm-up: thread pool, first touch of the arrays
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(steps * dt, dt)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 1000:
            break
    rate = N * N * steps / el
    return {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "sample": f"hydro KH {N}x{N} FP32 (the oracle's FP32 instantiation, OpenMP, "
                      f"{steps} timed RK3 steps after one warm-up step) on the host "
                      f"({el:.1f} s, OMP_NUM_THREADS={threads})"}


def bench_batched(ev, steps=500, nruns=64, seed=7):
    """64 independent hydro-isot runs (c_s = 1, box 2 pi) with sides
    log-uniform in [128, 1024], multiples of 32, seeded Gaussian density
    perturbations (delta rho / rho = 0.05, |k| <= 8), `steps` RK3 steps each
    (fg2_amd.ics.batched_workload, the workload the GPU tests check).
    Batched: one persistent launch per step for all runs (tile shape from the
    autotuner), 50 steps captured in a HIP graph once and replayed.  Baseline: each run launching its own
    single-run kernel, issued round-robin on the runs' own streams."""
    import fg2_amd
    from fg2_amd import ics
    n1s, n2s, xs = ics.batched_workload(nruns, seed)
    cells = int(np.sum(n1s.astype(np.int64) * n2s))

    runs = [fg2_amd.Run("hydro-isot", int(a), int(b)) for a, b in zip(n1s, n2s)]
    for r, x in zip(runs, xs):
        # the one-pass kernel on every run: one launch per step.  The small
        # runs' automatic substage kernels (three launches) are faster alone
        # but slower side by side.
        r.set_kernel(fg2_amd.KERNEL_ONE_PASS)
        r.upload(x)
    dts = [r.getdt() for r in runs]
    # per-run launches
    for r, d in zip(runs, dts):
        r.step_async(0.0, d)
    for r in runs:
        ev.sync(r.stream())
    t0 = time.perf_counter()
    for k in range(steps):
        for r, d in zip(runs, dts):
            r.step_async(k * d, d)
    for r in runs:
        ev.sync(r.stream())
    t_plain = time.perf_counter() - t0
    for r in runs:
        assert r.check() == 0
        r.close()

    # per-run launches of the plain path (bcond, kick, drift per substage, the
    # reference's operation split), on a shorter sample of steps
    runs = [fg2_amd.Run("hydro-isot", int(a), int(b), path=fg2_amd.PATH_PLAIN)
            for a, b in zip(n1s, n2s)]
    for r, x in zip(runs, xs):
        r.upload(x)
    plain_steps = 50
    for r, d in zip(runs, dts):
        r.step_async(0.0, d)
    for r in runs:
        ev.sync(r.stream())
    t0 = time.perf_counter()
    for k in range(plain_steps):
        for r, d in zip(runs, dts):
            r.step_async(k * d, d)
    for r in runs:
        ev.sync(r.stream())
    t_opsplit = (time.perf_counter() - t0) * steps / plain_steps
    for r in runs:
        assert r.check() == 0
        r.close()

    runs = [fg2_amd.Run("hydro-isot", int(a), int(b)) for a, b in zip(n1s, n2s)]
    for r, x in zip(runs, xs):
        r.upload(x)
    batch = fg2_amd.Batch(runs, waves=4, tile_h=0)
    tile = batch.autotune(dts)
    gsteps = 50                                       # steps per graph
    batch.graph_run(dts, steps=gsteps, replays=1)     # warm-up: captures the graph
    t0 = time.perf_counter()
    batch.graph_run(dts, steps=gsteps, replays=steps // gsteps)   # replays of the kept graph
    t_batch = time.perf_counter() - t0
    bad = sum(r.check() != 0 for r in runs)
    batch.close()
    for r in runs:
        r.close()
    return {"workload": f"{nruns} hydro-isot runs, sides log-uniform 128..1024 (x32), "
                        f"{steps} RK3 steps each, FP32",
            "cells": cells, "steps": steps,
            "batched_cell_updates_per_s": cells * steps / t_batch,
            "per_run_cell_updates_per_s": cells * steps / t_plain,
            "speedup": t_plain / t_batch, "batched_s": t_batch, "per_run_s": t_plain,
            "per_run_plain_path_cell_updates_per_s": cells * steps / t_opsplit,
            "speedup_vs_plain_path": t_opsplit / t_batch,
            "baselines": "per_run: each run's own one-pass step kernel, round-robin on the "
                         "runs' streams; per_run_plain_path: each run's bcond/kick/drift "
                         "launches (9 per step), timed over 50 steps and scaled",
            "tile": {"tile_h": tile[0], "waves": tile[1]}, "nan_runs": bad,
            "graph": f"a hipGraph of {gsteps} batched steps, captured once before the timed "
                     f"region and replayed {steps // gsteps} times"}


def _serve(arrivals, mods, n1s, n2s, steps, states, budget_gib, waves=0, tile_h=0,
           mode=0):
    """Drive a server through a request stream: requests are submitted when
    the wall clock passes their arrival time (all at once for arrivals of
    0), the server ticks as fast as the host allows.  Returns the latency
    array, the status count of failed requests, cell-steps, wall time, ticks
    and host time inside tick()."""
    import fg2_amd
    nreq = len(arrivals)
    srv = fg2_amd.Server(int(budget_gib * 2 ** 30), waves=waves, tile_h=tile_h)
    srv.set_mode(mode)
    ids = []
    k = 0
    ticks, tick_s = 0, 0.0
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        while k < nreq and arrivals[k] <= now:
            ids.append(srv.submit(mods[k], states[k], int(steps[k]), arrivals[k]))
            k += 1
        th0 = time.perf_counter()
        act, que, fin = srv.tick(th0 - t0)
        tick_s += time.perf_counter() - th0
        ticks += 1
        if k == nreq and fin == nreq:
            break
        if act == 0 and que == 0 and k < nreq:
            time.sleep(max(0.0, arrivals[k] - (time.perf_counter() - t0)))
    wall = time.perf_counter() - t0
    lat, bad = [], 0
    for i in ids:
        st, (ta, tadm, td), _ = srv.result(i, fetch=False)
        lat.append(td - ta)
        bad += st != 0
    cell_steps, _ = srv.stats()
    srv.close()
    return np.array(lat), int(bad), cell_steps, wall, ticks, tick_s


def bench_serving(nreq=1000, rate=100.0, seed=11, budget_gib=8.0, waves=0, tile_h=0,
                  ncap=300, nplain=100):
    """Poisson stream of run requests (rate per second) mixing hydro and
    hydro-isot, sides log-uniform in 64..2048, steps uniform in 50..2000,
    seeded random-field initial states (fg2_amd.ics.serving_workload);
    continuous batching under an HBM budget.  Reports p50 / p99 completion
    latency and aggregate cell-updates/s of the Poisson stream (whose rate is
    the offered load), then the server's capacity: the first `ncap` requests
    all arriving at once (the queue never drains until the end), served
    batched and served per run (each run's own one-pass launch per step on
    8 streams, same admission and arena), and their ratio; and on the first
    `nplain` requests the same against each run's plain kernels."""
    from fg2_amd import ics
    arrivals, mods, n1s, n2s, steps, make = ics.serving_workload(nreq, rate, seed)
    # initial states are generated before the timed loop (host set-up)
    states = [make(k) for k in range(nreq)]
    lat, bad, cell_steps, wall, ticks, tick_s = _serve(arrivals, mods, n1s, n2s, steps, states,
                                                      budget_gib, waves, tile_h)
    out = {"workload": f"{nreq} requests, Poisson {rate:g}/s, hydro + hydro-isot, "
                       "sides log-uniform 64..2048, steps U[50,2000], FP32",
           "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
           "mean_s": float(lat.mean()), "cell_updates_per_s": cell_steps / wall,
           "note": "cell_updates_per_s of the Poisson stream is the offered load (the "
                   "arrivals span about nreq / rate seconds); capacity is below",
           "wall_s": wall, "ticks": ticks, "host_s_in_tick": tick_s,
           "host_frac_of_wall": tick_s / wall,
           "hbm_budget_gib": budget_gib,
           "waves": waves if waves > 0 else 2,
           "tile_h": tile_h if tile_h > 0 else "chosen per launch (rounds of resident workgroups)",
           "nan_requests": bad}
    zero = np.zeros(ncap)
    cap = {}
    for name, mode in (("batched", 0), ("per_run", 1)):
        lat, bad, cell_steps, wall, ticks, tick_s = _serve(zero, mods[:ncap], n1s[:ncap],
                                                          n2s[:ncap], steps[:ncap],
                                                          states[:ncap], budget_gib, waves,
                                                          tile_h, mode)
        cap[name] = {"cell_updates_per_s": cell_steps / wall, "wall_s": wall, "ticks": ticks,
                     "host_s_in_tick": tick_s, "p50_s": float(np.percentile(lat, 50)),
                     "p99_s": float(np.percentile(lat, 99)), "nan_requests": bad}
    # the plain-kernel baseline is 8-10x slower: a shorter saturated queue
    # (the first nplain requests), measured batched and per run alike
    plain = {}
    for name, mode in (("batched", 0), ("per_run_plain", 2)):
        z = np.zeros(nplain)
        lat, bad, cell_steps, wall, ticks, tick_s = _serve(z, mods[:nplain], n1s[:nplain],
                                                          n2s[:nplain], steps[:nplain],
                                                          states[:nplain], budget_gib, waves,
                                                          tile_h, mode)
        plain[name] = {"cell_updates_per_s": cell_steps / wall, "wall_s": wall, "ticks": ticks,
                       "nan_requests": bad}
    out["capacity_cell_updates_per_s"] = cap["batched"]["cell_updates_per_s"]
    out["capacity"] = {"workload": f"the first {ncap} requests of the stream, all arriving "
                                   "at t = 0 (a saturated queue)",
                       "batched": cap["batched"], "per_run": cap["per_run"],
                       "speedup_vs_per_run": cap["batched"]["cell_updates_per_s"] /
                                             cap["per_run"]["cell_updates_per_s"],
                       "per_run_note": "each admitted run launches its own one-pass step "
                                       "kernel (its automatic one-pass shape, as "
                                       "fg2_set_kernel(FG2_KERNEL_ONE_PASS) gives a run of its "
                                       "size) per tick, on 8 streams; same admission, "
                                       "arena and copies",
                       "plain": {"workload": f"the first {nplain} requests of the stream, all "
                                             "arriving at t = 0",
                                 "batched": plain["batched"],
                                 "per_run_plain": plain["per_run_plain"],
                                 "speedup_vs_per_run_plain": plain["batched"]["cell_updates_per_s"] /
                                                             plain["per_run_plain"]["cell_updates_per_s"],
                                 "note": "per_run_plain: each admitted run launches the plain "
                                         "kernels (bcond, kick, drift per substage, 9 per "
                                         "step) per tick on 8 streams; same admission, arena "
                                         "and copies"}}
    return out


def bench_disk(ev, n=2048, steps=200, warmup=10):
    """The disk module on its (log r, theta) grid, n x n, Neumann in r and
    reflecting in theta, so three substage launches per RK3 step: the
    Keplerian disc of the precision study (point mass with softening length
    0.05, es = 1e6),
    fixed dt = 0.9 x the FP64 CFL step of the initial state."""
    import fg2_amd
    from fg2_amd import ics
    x = ics.disk_kepler(n, n, seed=4, h_over_r=0.05, eps=0.05)
    run = fg2_amd.Run("disk", n, n, params={"es": 1.0e6, "rS": 0.0, "eps": 0.05})
    run.upload(x)
    dt = 0.9 * run.getdt_cfl()
    stream = run.stream()

    def timed():
        for k in range(warmup):
            run.step_async(k * dt, dt)
        assert run.check() == 0
        e0, e1 = ev.event(), ev.event()
        ev.sync(stream)
        t0 = time.perf_counter()
        ev.record(e0, stream)
        for k in range(steps):
            run.step_async(k * dt, dt)
        ev.record(e1, stream)
        ev.sync(stream)
        return time.perf_counter() - t0, ev.elapsed_ms(e0, e1), run.check()

    wall, gpu_ms, status = timed()
    tile_h, waves = run.tile()
    tuned = run.autotune()                             # the autotuner's shape, timed the same way
    wall_t, _, status_t = timed()
    run.close()
    cells = n * n
    launch_s = gpu_ms * 1e-3 / (3 * steps)
    return {"workload": f"disk Keplerian disc {n}x{n} (softening 0.05), Neumann r / reflecting theta, FP32, "
                        "three substage launches per RK3 step, fixed 0.9 x CFL dt",
            "steps": steps, "dt": dt, "status": max(status, status_t),
            "ms_per_step": 1e3 * wall / steps,
            "cell_updates_per_s": cells * steps / wall,
            "hbm_frac_alg": cells * steps / wall * 240 / HBM_PEAK,
            "kernel": "substage_kernel<Disk,stage,waves>",
            "avg_launch_us": launch_s * 1e6,
            "tile": {"tile_h": tile_h, "waves": waves},
            "autotuned": {"tile_h": tuned[0], "waves": tuned[1],
                          "ms_per_step": 1e3 * wall_t / steps}}


def bench_slab(ev, dt, steps=200, warmup=10, nslabs=8, n=N):
    """The hydro KH grid cut along dim 1 into nslabs slabs (one per GPU when
    sharded), each stepped by the slab kernel from 9 halo rows per side.
    share: one GPU's slab, (n / nslabs) x n, as a one-slab ring (its halos
    are its own opposite edge, l1 scaled so Delta1 stays 1/n), per step one
    halo exchange (2 x 9 rows of device copies) and one slab step, both on
    the run's stream; whole: all nslabs slabs on this GPU, one exchange and
    nslabs slab steps per grid step.  dt: the whole grid's CFL step."""
    import fg2_amd
    from fg2_amd import ics
    h = n // nslabs
    x = ics.hydro_kh(n, n, seed=2)

    def timed(step, stream):
        for k in range(warmup):
            step()
        ev.sync(stream)
        e0, e1 = ev.event(), ev.event()
        t0 = time.perf_counter()
        ev.record(e0, stream)
        for k in range(steps):
            step()
        ev.record(e1, stream)
        ev.sync(stream)
        return time.perf_counter() - t0, ev.elapsed_ms(e0, e1) / steps

    share = fg2_amd.Run("hydro", h, n, geometry=dict(l1=h / n, l2=1.0, p1=1, p2=1, c=0.5))
    share.slab(0, h)
    share.upload(x[:h])
    halo_bytes = share.slab_halo()["bytes"]

    def share_step():
        fg2_amd.slab_exchange([share])
        share.slab_step(dt)

    wall_s, ms_s = timed(share_step, share.stream())
    _, ms_x = timed(lambda: fg2_amd.slab_exchange([share]), share.stream())
    status = share.check()
    tile_h, waves = share.tile()
    share.close()

    runs = []
    for k in range(nslabs):
        r = fg2_amd.Run("hydro", h, n, geometry=dict(l1=1.0, l2=1.0, p1=1, p2=1, c=0.5))
        r.slab(k * h, n)
        r.upload(x[k * h:(k + 1) * h])
        runs.append(r)
    def whole_step():
        fg2_amd.slab_exchange(runs)
        for r in runs:
            r.slab_step(dt)

    def sync_all():
        for r in runs:
            ev.sync(r.stream())

    for k in range(warmup):
        whole_step()
    sync_all()
    t0 = time.perf_counter()
    for k in range(steps):
        whole_step()
    sync_all()
    wall_w = time.perf_counter() - t0
    status = max(status, max(r.check() for r in runs))
    for r in runs:
        r.close()
    cells = h * n
    kernel_ms = ms_s - ms_x
    return {"workload": f"hydro KH {n}x{n} cut into {nslabs} slabs of {h}x{n}; share = one "
                        "slab with its 2 x 9 halo rows exchanged every step",
            "steps": steps, "status": status,
            "share": {"ms_per_step": ms_s, "kernel_ms_per_step": kernel_ms,
                      "exchange_ms_per_step": ms_x, "exchange_bytes_per_step": 2 * halo_bytes,
                      "cell_updates_per_s": cells / (ms_s * 1e-3),
                      "hbm_frac_alg": cells * 10 * 4 * 4 / (kernel_ms * 1e-3) / HBM_PEAK,
                      "tile": {"tile_h": tile_h, "waves": waves}},
            "whole_on_one_gpu": {"ms_per_step": 1e3 * wall_w / steps,
                                 "cell_updates_per_s": n * n * steps / wall_w}}


def bench_bounded(ev, steps=200, warmup=10, n=N):
    """The KH workload with Neumann rows and reflecting columns (p1 = p2 = 0):
    the one-pass step kernel with the ghost fill applied to every stage
    (bounded_step_kernel), against the substage path on the same state.
    Fixed dt = the FP64 CFL step of the initial state."""
    import fg2_amd
    from fg2_amd import ics
    x = ics.hydro_kh(n, n, seed=2)
    out = {"workload": f"hydro KH {n}x{n}, Neumann rows and reflecting columns, FP32, "
                       "fixed CFL dt", "steps": steps}
    for name, path in (("one_pass", fg2_amd.PATH_FUSED), ("substage", fg2_amd.PATH_SUBSTAGE)):
        run = fg2_amd.Run("hydro", n, n, geometry=dict(l1=1.0, l2=1.0, p1=0, p2=0, c=0.5),
                          path=path)
        run.upload(x)
        dt = run.getdt_cfl()
        stream = run.stream()
        for k in range(warmup):
            run.step_async(k * dt, dt)
        ev.sync(stream)
        e0, e1 = ev.event(), ev.event()
        ev.record(e0, stream)
        for k in range(steps):
            run.step_async(k * dt, dt)
        ev.record(e1, stream)
        ms = ev.elapsed_ms(e0, e1) / steps
        tile_h, waves = run.tile()
        out[name] = {"ms_per_step": ms, "cell_updates_per_s": n * n / (ms * 1e-3),
                     "status": run.check(), "tile": {"tile_h": tile_h, "waves": waves}}
        run.close()
    out["speedup"] = out["substage"]["ms_per_step"] / out["one_pass"]["ms_per_step"]
    return out


def bench_spectrum(n=N, calls=10):
    """Energy and vorticity shell spectra (fg2_spectrum) of the hydro KH
    state on n x n: pack u1 + i u2, one FP32 complex 2-D FFT, FP64 shell sums;
    results copied to the host.  The first call (FFT plan) is untimed."""
    import fg2_amd
    run = fg2_amd.Run("hydro", n, n)
    run.init("KH", 2)
    s = run.spectrum()
    t0 = time.perf_counter()
    for _ in range(calls):
        s = run.spectrum()
    wall = (time.perf_counter() - t0) / calls
    run.close()
    return {"workload": f"energy + vorticity shell spectra of hydro KH {n}x{n} (device IC), "
                        "FP32 FFT, FP64 shell sums, host copy of the spectra",
            "ms_per_call": 1e3 * wall, "cells_per_s": n * n / wall, "shells": len(s["E"])}


def bench_config1(ev, steps=100, reps=20):
    """Config 1 on the GPU: hydro 256^2 periodic, Gaussian random velocity
    (|k| <= 4, rms 0.01 c_s), closed-form dt, `steps` RK3 steps from the same
    state, repeated `reps` times (the least is kept).  A grid this small is
    latency bound: 256^2 cells fill a fraction of one round of workgroups."""
    import fg2_amd
    from fg2_amd import ics
    x = ics.hydro_random(256, 256, seed=1)
    run = fg2_amd.Run("hydro", 256, 256)
    dt = run.getdt()
    stream = run.stream()
    best = None
    for rep in range(reps):
        run.upload(x)
        ev.sync(stream)
        e0, e1 = ev.event(), ev.event()
        ev.record(e0, stream)
        for k in range(steps):
            run.step_async(k * dt, dt)
        ev.record(e1, stream)
        ms = ev.elapsed_ms(e0, e1)
        best = ms if best is None else min(best, ms)
    status = run.check()
    tile_h, waves = run.tile()
    kernel = "one_pass" if run.kernel() == fg2_amd.KERNEL_ONE_PASS else "substage"
    run.close()
    return {"workload": f"hydro 256x256 random velocity, {steps} RK3 steps on the GPU "
                        f"(least of {reps} repetitions from the same state)",
            "ms_total": best, "us_per_step": 1e3 * best / steps,
            "cell_updates_per_s": 256 * 256 * steps / (best * 1e-3), "status": status,
            "tile": {"tile_h": tile_h, "waves": waves}, "kernel": kernel}


def cpu_config1(steps=100):
    """Config 1 on the host: hydro 256^2 periodic, Gaussian random velocity
    (|k| <= 4, rms 0.01 c_s), closed-form dt, `steps` RK3 steps, with the
    oracle's FP32 and FP64 instantiations (OpenMP over rows)."""
    from oracle.oracle import Oracle
    from fg2_amd import ics
    x = ics.hydro_random(256, 256, seed=1)
    out = {}
    for name, fp64 in (("fp32", False), ("fp64", True)):
        o = Oracle("hydro", 256, 256, fp64=fp64)
        o.upload(x)
        dt = o.getdt()
        o.run(0.0, dt, 2)       # warm-up
        o.upload(x)
        t0 = time.perf_counter()
        assert o.run(0.0, dt, steps) == 0
        el = time.perf_counter() - t0
        out[name] = 256 * 256 * steps / el
    return {"workload": f"hydro 256x256 random velocity, {steps} RK3 steps "
                        "(after 2 warm-up steps)", "unit": "cell-updates/s",
            **out}


# The reference's published hydro-isot sweep on a Tesla M2090 (best block
# shape 4 x 128, bench/bench.txt:57-73): side -> ms per RK3 step.
REF_ISOT_M2090_MS = {1024: 5.468, 2048: 21.188, 3584: 71.384, 4096: 84.107, 8192: 337.530,
                     14336: 1036.88}


def bench_sweep(ev, sides=(128, 256, 512, 1024, 2048, 3584, 4096, 8192, 14336), steps=40,
                warmup=20):
    """hydro-isot on n x n periodic grids from the device initial condition
    Kf (the module's own IC, SplitMix64 noise), the automatic launch shape,
    closed-form dt: ms per RK3 step at each side, the reference's M2090
    figure where it publishes one, and the kernel the automatic choice takes
    (the substage kernels on grids too small for the one-pass kernel) with
    its tile."""
    import fg2_amd
    out = []
    for n in sides:
        run = fg2_amd.Run("hydro-isot", n, n)
        run.init("Kf", 5)
        dt = run.getdt()
        st = run.stream()
        for k in range(warmup):
            run.step_async(k * dt, dt)
        ev.sync(st)
        e0, e1 = ev.event(), ev.event()
        ev.record(e0, st)
        for k in range(steps):
            run.step_async(k * dt, dt)
        ev.record(e1, st)
        ms = ev.elapsed_ms(e0, e1) / steps
        tile_h, waves = run.tile()
        row = {"n": n, "ms_per_step": ms, "cell_updates_per_s": n * n / (ms * 1e-3),
               "status": run.check(), "tile": {"tile_h": tile_h, "waves": waves},
               "kernel": "one_pass" if run.kernel() == fg2_amd.KERNEL_ONE_PASS else "substage"}
        if n in REF_ISOT_M2090_MS:
            row["reference_m2090_ms_per_step"] = REF_ISOT_M2090_MS[n]
            row["vs_reference_m2090"] = REF_ISOT_M2090_MS[n] / ms
        out.append(row)
        run.close()
    return {"workload": "hydro-isot n x n periodic, device IC Kf, closed-form dt, "
                        f"{steps} timed RK3 steps after {warmup}",
            "rows": out}


def headline(ev, run, dt, steps, warmup, tile=0, waves=4, preload=100):
    """The timed region of the headline, on the run's stream.

    First `preload` steps of the same workload on the substage path (three
    launches per step; timed by events, reported as the substage figure),
    queued back to back with the W untimed one-pass warm-up steps and the
    timed region, so the GPU stays loaded up to the first timed launch: its
    clocks drop after about 5 ms without work and take some 25 ms of
    sustained load to come back, which the first launches of a fresh or idle
    GPU would otherwise measure.  Then exactly K one-pass steps between two
    events (wall clock around them too), and 25 more launches with an event
    after each, for the median launch time."""
    import fg2_amd
    stream = run.stream()
    run.set_path(fg2_amd.PATH_SUBSTAGE)
    run.set_tile(32, 4)
    s0, s1 = ev.event(), ev.event()
    ev.record(s0, stream)
    for k in range(preload):
        run.step_async(k * dt, dt)
    ev.record(s1, stream)
    run.set_path(fg2_amd.PATH_FUSED)
    run.set_tile(tile, waves)
    for k in range(warmup):
        run.step_async(k * dt, dt)
    sums0 = run.sums()                                  # FP64 mass and momentum
    assert run.check() == 0
    substage_ms = ev.elapsed_ms(s0, s1) / preload
    e0, e1 = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(e0, stream)
    for k in range(steps):
        run.step_async(k * dt, dt)
    ev.record(e1, stream)
    ev.sync(stream)
    wall = time.perf_counter() - t0
    gpu_ms = ev.elapsed_ms(e0, e1)
    status = run.check()
    sums1 = run.sums()
    evs = [ev.event() for _ in range(26)]
    ev.record(evs[0], stream)
    for k in range(25):
        run.step_async(k * dt, dt)
        ev.record(evs[k + 1], stream)
    series = [1e3 * ev.elapsed_ms(evs[k], evs[k + 1]) for k in range(25)]
    status = max(status, run.check())
    return wall, gpu_ms, status, sums0, sums1, series, substage_ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--waves", type=int, default=4)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-disk", action="store_true")
    ap.add_argument("--no-spectrum", action="store_true")
    ap.add_argument("--no-slab", action="store_true")
    ap.add_argument("--no-bounded", action="store_true")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--disk-only", action="store_true",
                    help="run only the disk leg and print its JSON (profiling)")
    ap.add_argument("--slab-only", action="store_true",
                    help="run only the slab leg and print its JSON (profiling)")
    ap.add_argument("--bounded-only", action="store_true",
                    help="run only the bounded-grid leg and print its JSON (profiling)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline step and its JSON line (profiling)")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched leg and print its JSON (profiling)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving leg and print its JSON (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the time-step path runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import fg2_amd
    from fg2_amd import ics

    ev = HipEvents()
    if args.disk_only:
        print(json.dumps({"disk_2048": bench_disk(ev, steps=args.steps, warmup=args.warmup)}),
              flush=True)
        return
    if args.bounded_only:
        print(json.dumps({"bounded": bench_bounded(ev, steps=args.steps, warmup=args.warmup)}),
              flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batched": bench_batched(ev)}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"serving": bench_serving()}), flush=True)
        return
    x = ics.hydro_kh(N, N, seed=2)
    run = fg2_amd.Run("hydro", N, N, geometry=dict(l1=1.0, l2=1.0, p1=1, p2=1, c=0.5))
    run.set_tile(args.tile, args.waves)
    run.upload(x)
    dt = run.getdt_cfl()
    if args.slab_only:
        run.close()
        print(json.dumps({"slabs": bench_slab(ev, dt, steps=args.steps, warmup=args.warmup)}),
              flush=True)
        return
    if args.headline_only:
        for flag in ("no_batch", "no_serve", "no_disk", "no_spectrum", "no_slab", "no_bounded",
                     "no_config1", "no_sweep", "no_cpu"):
            setattr(args, flag, True)
    # the other GPU legs first (see the module docstring), the GPU-heavy
    # slab leg last before the headline
    legs = {}
    if not args.no_serve:
        legs["serving"] = bench_serving()
    if not args.no_batch:
        legs["batched"] = bench_batched(ev)
    if not args.no_disk:
        legs["disk_2048"] = bench_disk(ev)
    if not args.no_spectrum:
        legs["spectrum"] = bench_spectrum()
    if not args.no_config1:
        legs["config1_hydro_256"] = bench_config1(ev)
    if not args.no_sweep:
        legs["sweep_hydro_isot"] = bench_sweep(ev)
    if not args.no_bounded:
        legs["bounded"] = bench_bounded(ev)
    if not args.no_slab:
        legs["slabs"] = bench_slab(ev, dt)

    wall, gpu_ms, status, sums0, sums1, series, substage_ms = headline(
        ev, run, dt, args.steps, args.warmup, tile=args.tile, waves=args.waves)
    if status != 0:
        print("NaN in the benchmark state", file=sys.stderr)
        sys.exit(1)
    conservation = {"steps": args.steps,
                    "mass_rel_drift": float(sums1[0] / sums0[0] - 1.0),
                    "momentum1_drift_per_mass": float((sums1[1] - sums0[1]) / sums0[0]),
                    "momentum2_drift_per_mass": float((sums1[2] - sums0[2]) / sums0[0]),
                    "note": "FP64 sums over the grid of rho, rho u1, rho u2 before and after "
                            "the timed steps; the scheme evolves ln rho, so mass is not "
                            "conserved to round-off"}

    cells = N * N
    value = cells * args.steps / wall
    ms_step = 1e3 * wall / args.steps
    nv = 4
    alg_step = 10 * nv * 4 * cells                      # bytes per RK3 step
    # periodic hydro runs take one launch of the one-pass step kernel per step
    launch_s = gpu_ms * 1e-3 / args.steps
    traffic, valu_insts = load_counters()
    tile_h, waves = run.tile()
    pm = run.perf_model()
    # the reference's own perf model (src/config.cpp:30-36, printed per frame
    # by src/solve.cpp:55-57) at the measured step time
    ref_model = {"flop_per_step": pm["flops"], "bytes_per_step": pm["bits"] / 8,
                 "gflops": pm["flops"] / (launch_s * 1e9),
                 "gbytes_per_s": pm["bits"] / 8 / (launch_s * 1e9),
                 "note": "eqns_flop + 2 NVAR flop per cell and substage, and 18 NVAR "
                         "R-words per cell per step plus ghosts, as the reference counts"}
    roofline = {
        "bound": "valu",
        "achieved": None if valu_insts is None else valu_insts / launch_s / 1e9,
        "peak": VALU_PEAK / 1e9, "unit": "G wave-instructions/s",
        "frac": None if valu_insts is None else valu_insts / launch_s / VALU_PEAK,
        "traffic": traffic,
        "kernel": "step_kernel<Hydro,4>: one launch per RK3 step, the three substages "
                  "temporally blocked",
        "insts_per_launch": valu_insts,
        "avg_launch_us": launch_s * 1e6,
        "median_launch_us": float(np.median(series)),
        "launch_series_us": series,
        "note": "the bound that holds is VALU issue: SQ_INSTS_VALU per launch (rocprofv3 "
                "counters, profiles/) / the live launch time, against 256 CUs x 4 SIMDs x "
                "2.4 GHz / 4 cycles per FP32 wave64 instruction (packed FP32 off)",
        "hbm": None if traffic is None else {
            "achieved": traffic / launch_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": traffic / launch_s / HBM_PEAK,
            "note": "measured HBM traffic per launch (FETCH_SIZE + WRITE_SIZE, calibrated "
                    "on a copy kernel) / the live launch time"},
        "alg_substage_equivalent": {
            "bytes_per_launch": alg_step, "achieved": alg_step / launch_s / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": alg_step / launch_s / HBM_PEAK,
            "note": "B_alg = 10 NVAR 4 B per cell: the bytes three one-pass substage "
                    "launches would move; this temporally blocked kernel reads and writes "
                    "the state once, so this is not an HBM fraction"},
        "alg_flops": {
            "flops_per_launch": pm["flops"], "achieved": pm["flops"] / launch_s / 1e12,
            "peak": 78.6, "unit": "TFLOP/s",
            "frac": pm["flops"] / launch_s / 78.6e12,
            "note": "the reference's flop model (eqns_flop + 2 NVAR per cell and substage) "
                    "against MI355X FP32 vector FMA issue without packed FP32 "
                    "(157.3 TFLOP/s with it)"},
    }

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"hydro Kelvin-Helmholtz {N}x{N} periodic, FP32, "
                               "one-pass fused RK3 step, fixed CFL dt",
                   "n1": N, "n2": N, "module": "hydro", "dt": dt,
                   "tile_h": tile_h, "waves": waves},
        "ns_per_cell_step": 1e9 * wall / (cells * args.steps),
        "hbm_frac_alg": value * 10 * nv * 4 / HBM_PEAK,
        "hbm_frac_alg_note": "cell-updates/s x B_alg (160 B per hydro cell, substage-"
                             "equivalent) / 8 TB/s; the measured HBM fraction is roofline.hbm",
        "roofline": roofline,
        "reference_perf_model": ref_model,
        "substage_path_ms_per_step": substage_ms,
        "conservation": conservation,
        "batched": None,
        "serving": None,
        "disk_2048": None,
        "spectrum": None,
        "slabs": None,
        "bounded": None,
        "config1_hydro_256": None,
        "sweep_hydro_isot": None,
        "cpu_baseline": None,
    }
    out.update(legs)
    if out["slabs"] is not None:
        out["slabs"]["share"]["vs_whole_grid_share"] = \
            ms_step / 8 / out["slabs"]["share"]["ms_per_step"]
    run.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_baseline"]["config1_hydro_256"] = cpu_config1()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
