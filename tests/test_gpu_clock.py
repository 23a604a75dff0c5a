"""GPU tests of steps whose dt is chosen on the device, and of batched and
served runs with an Euler pre-step (first_order).

  - fg2_batch_solve against fg2_solve run by run, for one-pass batches
    (periodic hydro-isot, hydro), substage batches (disk, with per-cell
    diffusive rates; hydro runs of mixed periodicity) and both dt modes:
    the same state bits, the same end times and the same step counts,
    each run with its own end time (runs that finish first stand still);
  - a run of a CFL batch against the FP64 oracle's CFL loop
    (src/solve.cpp:37-49 with the CFL rule), within 1e-5 relative L2;
  - batched steps of runs with an active first_order -- hydro-isot
    stochastic forcing (f k > 0, src/hydro-isot/Euler.cuh:21-44) and disk
    radiation equilibrium (ff != 0, src/disk/Euler.cuh:19-69) -- eager,
    graph-replayed and through the no-allocation entry point, against the
    same runs stepped alone, bit for bit;
  - served requests shaped by a configured run: CFL steps, the forcing and
    non-default parameters, each result the bits of the run stepped alone.
"""
import sys

import numpy as np
import pytest

from helpers import DISK_PARAMS, rel_l2, state

pytestmark = pytest.mark.gpu

FORCED = {"f": 0.05, "k": 2.0}          # f k > 0: the stochastic forcing is on
# free-free opacity on: the radiation-equilibrium pre-step runs
RADIATIVE = dict(DISK_PARAMS, ff=1e-6, ar=1.0, dd=5e-4, nu=5e-4, ed=1e-3, rd=1e-3)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _runs(fg2, module, sizes, seed, params=None, geometries=None, cfl=False, seeds=None):
    out = []
    for k, (a, b) in enumerate(sizes):
        geo = geometries[k] if geometries else None
        r = fg2.Run(module, a, b, params=params, geometry=geo,
                    seed=(seeds[k] if seeds else 0))
        if cfl:
            r.set_dt_mode(fg2.DT_CFL)
        r.upload(state(module, a, b, seed=seed + k))
        out.append(r)
    return out


def _solve_vs_alone(fg2, module, sizes, seed, params=None, geometries=None, cfl=False,
                    spans=None):
    a = _runs(fg2, module, sizes, seed, params, geometries, cfl)
    b = _runs(fg2, module, sizes, seed, params, geometries, cfl)
    dt0 = [r.getdt_cfl() if cfl else r.getdt() for r in a]
    spans = spans or [6.5 + 2 * k for k in range(len(sizes))]
    t_final = [s * d for s, d in zip(spans, dt0)]
    batch = fg2.Batch(b, waves=4 if module != "disk" else 1, tile_h=0)
    # a bound on the steps keeps the solve finite whatever the dt rule gives
    code, t, steps = batch.solve(np.zeros(len(sizes)), t_final, max_steps=4096)
    assert code == 0
    for k, r in enumerate(a):
        code, tk, mk = r.solve(0.0, t_final[k])
        assert code == 0
        assert t[k] == tk == t_final[k], (k, t[k], tk, t_final[k])
        assert steps[k] == mk, (k, steps[k], mk)
        np.testing.assert_array_equal(b[k].download(), r.download(), err_msg=f"run {k} {sizes[k]}")
    batch.close()
    for r in a + b:
        r.close()
    return steps


@pytest.mark.parametrize("cfl", [True, False])
def test_batch_solve_isot_equals_solve(fg2, cfl):
    """A one-pass batch of periodic hydro-isot runs of four sizes, each to
    its own end time (the shortest stands still for the last steps)."""
    sizes = [(130, 200), (64, 300), (257, 96), (90, 90)]
    _say(f"case batch solve hydro-isot cfl={cfl}")
    steps = _solve_vs_alone(fg2, "hydro-isot", sizes, 40, cfl=cfl, spans=[5.5, 23.25, 12.0, 40.0])
    assert len(set(steps.tolist())) > 1


def test_batch_solve_hydro_cfl_equals_solve(fg2):
    sizes = [(96, 128), (160, 64), (72, 72)]
    _say("case batch solve hydro cfl")
    _solve_vs_alone(fg2, "hydro", sizes, 50, cfl=True)


def test_batch_solve_disk_cfl_equals_solve(fg2):
    """A substage batch: disk runs (per-cell diffusive rates), CFL mode."""
    sizes = [(96, 80), (64, 128)]
    _say("case batch solve disk cfl")
    _solve_vs_alone(fg2, "disk", sizes, 60, params=DISK_PARAMS, cfl=True, spans=[4.5, 9.0])


def test_batch_solve_mixed_periodicity_cfl(fg2):
    """A substage batch of hydro runs: periodic, Neumann rows, reflecting
    columns."""
    sizes = [(80, 96), (96, 64), (64, 64)]
    geos = [dict(p1=1, p2=1), dict(p1=0, p2=1), dict(p1=1, p2=0)]
    _say("case batch solve hydro mixed periodicity cfl")
    _solve_vs_alone(fg2, "hydro", sizes, 70, geometries=geos, cfl=True)


def test_batch_solve_long_and_capped(fg2):
    """More steps than one graph holds (graph replays, then single launches
    for the remainder), and a max_steps cap below what the end time needs."""
    sizes = [(100, 128), (64, 96)]
    a = _runs(fg2, "hydro-isot", sizes, 80, cfl=True)
    b = _runs(fg2, "hydro-isot", sizes, 80, cfl=True)
    t_final = [1e9, 1e9]
    batch = fg2.Batch(b, waves=4)
    code, t, steps = batch.solve([0.0, 0.0], t_final, max_steps=37)
    assert code == 0 and steps.tolist() == [37, 37]
    for k, r in enumerate(a):
        tt = 0.0
        for _ in range(37):
            dt = r.getdt_cfl()
            assert r.step(tt, dt) == 0
            tt += dt
        assert t[k] == tt
        np.testing.assert_array_equal(b[k].download(), r.download())
    # the rates of the final state are kept: a second solve continues alike
    tf2 = [t[0] + 3.5 * a[0].getdt_cfl(), t[1] + 1e-12]
    code, t2, steps2 = batch.solve(t, tf2, max_steps=4096)
    assert code == 0 and t2.tolist() == tf2
    for k, r in enumerate(a):
        code, tk, mk = r.solve(t[k], tf2[k])
        assert code == 0 and tk == t2[k] and mk == steps2[k]
        np.testing.assert_array_equal(b[k].download(), r.download())
    batch.close()
    for r in a + b:
        r.close()


def test_batch_solve_cfl_vs_oracle(fg2):
    """One run of a CFL batch follows the FP64 oracle's CFL loop: the same
    step count to the same end time, within 1e-5 relative L2 per field."""
    from oracle.oracle import Oracle
    sizes = [(64, 80), (96, 96)]
    b = _runs(fg2, "hydro", sizes, 90, cfl=True)
    x = state("hydro", 64, 80, seed=90)
    o = Oracle("hydro", 64, 80)
    o.upload(x.astype(np.float32).astype(np.float64))
    T = 10.5 * o.getdt_cfl()
    batch = fg2.Batch(b, waves=4)
    code, t, steps = batch.solve([0.0, 0.0], [T, 4.0 * b[1].getdt_cfl()], max_steps=4096)
    assert code == 0 and t[0] == T
    tt, k = 0.0, 0
    while tt < T:
        dt = min(T - tt, o.getdt_cfl())
        assert o.step(tt, dt) == 0
        tt = min(tt + dt, T)
        k += 1
    assert k == steps[0]
    err = rel_l2(b[0].download(), o.download())
    _say("batch CFL run vs oracle", err)
    assert np.all(err < 1e-5), err
    batch.close()
    for r in b:
        r.close()


def _step_alone(fg2, module, sizes, seed, params, dts, nsteps, seeds):
    out = []
    for k, (a, b) in enumerate(sizes):
        r = fg2.Run(module, a, b, params=params, seed=seeds[k])
        r.upload(state(module, a, b, seed=seed + k))
        for s in range(nsteps):
            assert r.step(s * dts[k], dts[k]) == 0
        out.append(r.download())
        r.close()
    return out


@pytest.mark.parametrize("module,params,sizes", [
    ("hydro-isot", FORCED, [(130, 200), (64, 300), (96, 96)]),
    ("disk", RADIATIVE, [(96, 80), (64, 128)]),
])
def test_batch_first_order_equals_alone(fg2, module, params, sizes):
    """Batched steps with an active first_order pre-step: eager steps, graph
    replays (the hydro-isot phases then drawn on the device) and the
    no-allocation entry point give every run the bits of the run stepped
    alone with fg2_step."""
    seeds = [11 + 7 * k for k in range(len(sizes))]
    probe = fg2.Run(module, *sizes[0], params=params)
    probe.upload(state(module, *sizes[0], seed=30))
    dts = [0.25 * probe.getdt_cfl()] * len(sizes)
    probe.close()
    _say(f"case batch first_order {module}")
    alone6 = _step_alone(fg2, module, sizes, 30, params, dts, 6, seeds)
    # eager: 2 steps, graph: 2 x 2 steps
    b = _runs(fg2, module, sizes, 30, params, seeds=seeds)
    batch = fg2.Batch(b, waves=4 if module != "disk" else 1)
    for _ in range(2):
        assert batch.step(dts) == 0
    batch.graph_run(dts, steps=2, replays=2)
    for k, r in enumerate(b):
        y = r.download()
        assert np.all(np.isfinite(y))
        np.testing.assert_array_equal(y, alone6[k], err_msg=f"run {k} {sizes[k]}")
    batch.close()
    # the no-allocation entry point: 3 steps
    alone3 = _step_alone(fg2, module, sizes, 30, params, dts, 3, seeds)
    c = _runs(fg2, module, sizes, 30, params, seeds=seeds)
    for _ in range(3):
        assert fg2.batch_step_static(c, dts, waves=4 if module != "disk" else 1) == 0
    for k, r in enumerate(c):
        np.testing.assert_array_equal(r.download(), alone3[k], err_msg=f"static run {k}")
    for r in b + c:
        r.close()


def test_batch_solve_forced_isot(fg2):
    """Forced hydro-isot runs under fg2_batch_solve: the phases drawn on the
    device each step are the host's, so the runs match fg2_solve."""
    sizes = [(96, 128), (128, 64)]
    seeds = [5, 6]
    a, b = [], []
    for lst in (a, b):
        for k, (n1, n2) in enumerate(sizes):
            r = fg2.Run("hydro-isot", n1, n2, params=FORCED, seed=seeds[k])
            r.set_dt_mode(fg2.DT_CFL)
            r.upload(state("hydro-isot", n1, n2, seed=100 + k))
            lst.append(r)
    t_final = [9.5 * r.getdt_cfl() for r in a]
    batch = fg2.Batch(b, waves=4)
    code, t, steps = batch.solve([0.0, 0.0], t_final, max_steps=4096)
    assert code == 0
    for k, r in enumerate(a):
        code, tk, mk = r.solve(0.0, t_final[k])
        assert code == 0 and tk == t[k] and mk == steps[k]
        np.testing.assert_array_equal(b[k].download(), r.download(), err_msg=f"run {k}")
    batch.close()
    for r in a + b:
        r.close()


def test_server_runs_like_configured(fg2):
    """Requests shaped by configured runs: hydro in CFL mode with bulk
    viscosity and density diffusion, forced hydro-isot (fixed dt, and CFL),
    default hydro-isot; mixed in one server.  Each result is the bits of
    its run stepped alone with fg2_step and dt = fg2_getdt or fg2_getdt_cfl
    before every step."""
    specs = [
        ("hydro", (96, 128), {"nub": 1e-4, "dd": 2e-4}, True, 0, 30),
        ("hydro-isot", (128, 96), FORCED, False, 17, 25),
        ("hydro-isot", (64, 200), FORCED, True, 23, 40),
        ("hydro-isot", (100, 100), None, False, 0, 12),
        ("hydro", (64, 64), None, False, 0, 20),
    ]
    S = fg2.Server(hbm_budget=64 << 20)
    ids, likes, inits = [], [], []
    for k, (m, (n1, n2), params, cfl, seed, steps) in enumerate(specs):
        like = fg2.Run(m, n1, n2, params=params, seed=seed)
        if cfl:
            like.set_dt_mode(fg2.DT_CFL)
        x = state(m, n1, n2, seed=200 + k)
        ids.append(S.submit_run(like, x, steps, now=0.0))
        likes.append(like)
        inits.append(x)
    for it in range(100000):
        a, q, f = S.tick(0.0)
        if f == len(specs) and a == 0 and q == 0:
            break
    for k, (m, (n1, n2), params, cfl, seed, steps) in enumerate(specs):
        st, _, y = S.result(ids[k])
        assert st == 0, (k, st)
        r = fg2.Run(m, n1, n2, params=params, seed=seed)
        if cfl:
            r.set_dt_mode(fg2.DT_CFL)
        r.upload(inits[k])
        t = 0.0
        for s in range(steps):
            dt = r.getdt_cfl() if cfl else r.getdt()
            assert r.step(t, dt) == 0
            t += dt
        np.testing.assert_array_equal(y, r.download(), err_msg=f"request {k} {m} {n1}x{n2}")
        r.close()
    S.close()
    for like in likes:
        like.close()


def test_batch_forced_bounded_isot_graph(fg2):
    """Forced hydro-isot runs with Neumann rows step by substages; their
    graph replays draw the phases on the device (the clock of a substage
    batch) and match the runs stepped alone."""
    sizes = [(96, 128), (64, 80)]
    seeds = [3, 4]
    geo = dict(l1=6.283185307179586, l2=6.283185307179586, p1=0, p2=1, c=0.5)
    alone = []
    for k, (n1, n2) in enumerate(sizes):
        r = fg2.Run("hydro-isot", n1, n2, params=FORCED, geometry=geo, seed=seeds[k])
        r.upload(state("hydro-isot", n1, n2, seed=60 + k))
        dt = 0.25 * r.getdt()
        for s in range(6):
            assert r.step(s * dt, dt) == 0
        alone.append((r.download(), dt))
        r.close()
    b = []
    for k, (n1, n2) in enumerate(sizes):
        r = fg2.Run("hydro-isot", n1, n2, params=FORCED, geometry=geo, seed=seeds[k])
        r.upload(state("hydro-isot", n1, n2, seed=60 + k))
        b.append(r)
    dts = [d for _, d in alone]
    batch = fg2.Batch(b, waves=1)
    for _ in range(2):
        assert batch.step(dts) == 0
    batch.graph_run(dts, steps=2, replays=2)
    for k, r in enumerate(b):
        np.testing.assert_array_equal(r.download(), alone[k][0], err_msg=f"run {k}")
    batch.close()
    for r in b:
        r.close()
