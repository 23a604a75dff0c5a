"""GPU: a periodic grid cut along dim 1 into slabs (one per GPU when
sharded), each stepped by the one-pass kernel from halo rows its neighbours
supply, steps to the same bits as the whole grid.

The slab path changes only where the kernel reads the rows outside its
slab (the halo buffers, filled by fg2_slab_exchange, instead of wrapped
rows) and the row index the right-hand side sees (the global one, which
the hydro-isot Kolmogorov forcing uses).  The per-cell arithmetic is that
of the whole-grid kernel, whose bits the other GPU tests tie to the
substage and plain paths and, within tolerance, to the FP64 oracle.  So
the bar here is bit identity with the whole grid, at every step.
"""
import sys

import numpy as np
import pytest

from helpers import DISK_PARAMS, state

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _slabs(fg2, module, x, heights, params=None, seed=0, cfl=False):
    n1g, n2 = x.shape[:2]
    assert sum(heights) == n1g
    runs, row0 = [], 0
    for h in heights:
        r = fg2.Run(module, h, n2, params=params, seed=seed)
        r.slab(row0, n1g)
        if cfl:
            r.set_dt_mode(fg2.DT_CFL)
        r.upload(x[row0:row0 + h])
        runs.append(r)
        row0 += h
    return runs


def _gather(runs):
    return np.concatenate([r.download() for r in runs], axis=0)


def _step_both(fg2, whole, runs, dt, steps):
    for k in range(steps):
        fg2.slab_exchange(runs)
        for r in runs:
            r.slab_step(dt)
        assert whole.step(k * dt, dt) == 0
        for r in runs:
            assert r.check() == 0
        np.testing.assert_array_equal(_gather(runs), whole.download(), err_msg=f"step {k}")


@pytest.mark.parametrize("module,n1,n2,heights", [
    ("hydro", 256, 300, [64, 50, 90, 52]),     # unequal slabs; 300 columns: idle waves
    ("hydro", 96, 61, [96]),                   # one slab: its own neighbour
    ("hydro-isot", 120, 130, [9, 40, 71]),     # a slab of exactly the halo depth
])
def test_slabs_equal_whole_grid(fg2, module, n1, n2, heights):
    _say(f"case slabs {module} {n1}x{n2} heights={heights}")
    x = state(module, n1, n2, seed=5)
    whole = fg2.Run(module, n1, n2)
    whole.upload(x)
    runs = _slabs(fg2, module, x, heights)
    dt = whole.getdt()
    assert all(r.getdt() == dt for r in runs)      # getdt sees the whole grid's rows
    _step_both(fg2, whole, runs, dt, 4)


def test_slabs_kolmogorov_forcing_uses_global_rows(fg2):
    """hydro-isot with f k < 0: the forcing f cos(k x1) depends on the
    global row, so a slab that saw its local row index would differ."""
    n1, n2 = 128, 96
    params = {"f": -0.3, "k": 3.0}
    _say(f"case slabs isot Kolmogorov {n1}x{n2}")
    x = state("hydro-isot", n1, n2, seed=6)
    whole = fg2.Run("hydro-isot", n1, n2, params=params)
    whole.upload(x)
    runs = _slabs(fg2, "hydro-isot", x, [40, 48, 40], params=params)
    dt = whole.getdt()
    _step_both(fg2, whole, runs, dt, 3)
    free = fg2.Run("hydro-isot", n1, n2)
    free.upload(x)
    for k in range(3):
        assert free.step(k * dt, dt) == 0
    assert np.max(np.abs(whole.download() - free.download())) > 1e-4   # the forcing acted


def test_slabs_cfl_least_equals_whole(fg2):
    """The FP64 CFL rate is a max over cells, so the least slab dt equals
    the whole grid's dt exactly."""
    n1, n2 = 192, 160
    _say(f"case slabs CFL hydro {n1}x{n2}")
    x = state("hydro", n1, n2, seed=7)
    whole = fg2.Run("hydro", n1, n2)
    whole.set_dt_mode(fg2.DT_CFL)
    whole.upload(x)
    runs = _slabs(fg2, "hydro", x, [64, 64, 64], cfl=True)
    for k in range(3):
        dt = whole.getdt_cfl()
        assert min(r.getdt_cfl() for r in runs) == dt
        fg2.slab_exchange(runs)
        for r in runs:
            r.slab_step(dt)
        assert whole.step(0.0, dt) == 0
    np.testing.assert_array_equal(_gather(runs), whole.download())


def test_slabs_full_size_share(fg2):
    """3584^2 Kelvin-Helmholtz cut into 8 slabs of 448 rows (the share of
    one GPU of 8): bit identity with the whole grid over 3 steps."""
    n = 3584
    _say(f"case slabs KH {n}^2 x 8 slabs")
    whole = fg2.Run("hydro", n, n)
    whole.init("KH", seed=1)
    x = whole.download()
    runs = _slabs(fg2, "hydro", x, [n // 8] * 8)
    dt = whole.getdt()
    for k in range(3):
        fg2.slab_exchange(runs)
        for r in runs:
            r.slab_step(dt)
        assert whole.step(k * dt, dt) == 0
    for r in runs:
        assert r.check() == 0
    np.testing.assert_array_equal(_gather(runs), whole.download())


def test_slab_autotune_keeps_bits(fg2):
    """The autotuner times a slab's own kernel (with its halo buffers) and
    leaves its state alone; the tuned slabs still step to the whole grid's
    bits."""
    n1, n2 = 192, 300
    _say(f"case slab autotune {n1}x{n2}")
    x = state("hydro", n1, n2, seed=8)
    whole = fg2.Run("hydro", n1, n2)
    whole.upload(x)
    runs = _slabs(fg2, "hydro", x, [96, 96])
    fg2.slab_exchange(runs)
    for r in runs:
        r.autotune()
    np.testing.assert_array_equal(_gather(runs), x.astype(np.float32))
    _step_both(fg2, whole, runs, whole.getdt(), 3)


def test_slab_refusals(fg2):
    """Host-side checks: what a slab cannot be, and the whole-grid entry
    points a slab refuses."""
    with pytest.raises(fg2.Fg2Error, match="not periodic"):
        fg2.Run("disk", 32, 32, params=DISK_PARAMS).slab(0, 64)
    r = fg2.Run("hydro", 8, 32)
    with pytest.raises(fg2.Fg2Error, match="at least 9 rows"):
        r.slab(0, 64)
    r = fg2.Run("hydro", 16, 32)
    with pytest.raises(fg2.Fg2Error, match="do not fit"):
        r.slab(0, 8)
    with pytest.raises(fg2.Fg2Error, match="not a slab"):
        r.slab_step(1e-3)
    r.slab(0, 32)
    with pytest.raises(fg2.Fg2Error, match="fg2_slab_step"):
        r.step_async(0.0, 1e-3)
    with pytest.raises(fg2.Fg2Error, match="slabs"):
        fg2.Batch([r])
    s = fg2.Run("hydro", 16, 32)
    s.slab(20, 32)                       # rows 20..35 wrap: 16 + 16 = 32, but 0 + 16 != 20
    with pytest.raises(fg2.Fg2Error, match="next one starts"):
        fg2.slab_exchange([r, s])
    f = fg2.Run("hydro-isot", 16, 32, params={"f": 0.3, "k": 5.0})
    f.slab(0, 16)
    with pytest.raises(fg2.Fg2Error, match="first_order"):
        f.slab_step(1e-3)
    r.slab(0, 0)                         # a whole grid again
    r.upload(state("hydro", 16, 32))
    r.step_async(0.0, 1e-3)
    assert r.check() == 0
