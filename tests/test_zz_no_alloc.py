"""fg2_batch_step_static -- the batched step that allocates nothing -- against
fg2_batch_step on the same runs: bit-identical states, on two calls in a row
with the second on fewer, smaller runs (the fixed device array keeps the
first call's records and counters, which the second must not see), for a
one-pass batch (periodic hydro-isot) and a substage batch (disk), and the
error for more runs than the array holds."""
import numpy as np
import pytest

from helpers import DISK_PARAMS, state

pytestmark = pytest.mark.gpu


def _runs(fg2, module, sizes, seed, params=None):
    out = []
    for k, (a, b) in enumerate(sizes):
        r = fg2.Run(module, a, b, params=params)
        r.upload(state(module, a, b, seed=seed + k))
        out.append(r)
    return out


def _dts(runs):
    return [0.25 * r.getdt() if r.module != "disk" else 0.5 * r.getdt_cfl() for r in runs]


@pytest.mark.parametrize("module,calls", [
    ("hydro-isot", [[(130, 200), (64, 300), (257, 96), (90, 90)], [(70, 80), (40, 150)]]),
    ("disk", [[(96, 80), (64, 128)], [(48, 40)]]),
])
@pytest.mark.parametrize("waves", [1, 4])
def test_static_batch_equals_batch(fg2, module, calls, waves):
    params = DISK_PARAMS if module == "disk" else None
    for c, sizes in enumerate(calls):
        print(f"case static batch {module} waves={waves} call {c} sizes {sizes}", flush=True)
        a = _runs(fg2, module, sizes, 80 + 10 * c, params)
        b = _runs(fg2, module, sizes, 80 + 10 * c, params)
        dts = _dts(a)
        batch = fg2.Batch(b, waves=waves, tile_h=0)
        for _ in range(3):
            assert fg2.batch_step_static(a, dts, waves=waves, tile_h=0) == 0
            assert batch.step(dts) == 0
        for k, (ra, rb) in enumerate(zip(a, b)):
            ya, yb = ra.download(), rb.download()
            assert np.all(np.isfinite(ya))
            np.testing.assert_array_equal(ya, yb, err_msg=f"call {c} run {k} {sizes[k]}")
        batch.close()
        for r in a + b:
            r.close()


def test_static_batch_limits(fg2):
    r = fg2.Run("hydro", 16, 16)
    r.upload(state("hydro", 16, 16))
    runs = [r] * (fg2.STATIC_MAX_RUNS + 1)
    with pytest.raises(fg2.Fg2Error) as e:
        fg2.batch_step_static(runs, [1e-4] * len(runs))
    assert e.value.code == fg2.ERR_SHAPE and "at most" in str(e.value)
    # a NaN is reported as by fg2_batch_step
    x = state("hydro", 16, 16)
    x[3, 4, 1] = np.nan
    r.upload(x)
    assert fg2.batch_step_static([r], [1e-4]) == fg2.ERR_NAN
    r.close()
