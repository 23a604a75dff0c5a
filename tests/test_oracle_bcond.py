"""Pin the oracle's ghost-zone fill (src/bcond.cu:22-86) by brute force with
numpy: periodic = wrap, Neumann1 = even mirror about the face, reflect2 =
mirror with the module's odd variable negated; dim 1 is filled first, so the
corners carry the dim-2 transform of the dim-1 image."""
import numpy as np
import pytest

from oracle.oracle import Oracle

ODD = {"hydro": 2, "hydro-isot": 2, "disk": 1}   # src/*/bcond.cuh


def expected(x, module, p1, p2):
    g = np.pad(x, ((3, 3), (0, 0), (0, 0)), mode="wrap" if p1 else "symmetric")
    g = np.pad(g, ((0, 0), (3, 3), (0, 0)), mode="wrap" if p2 else "symmetric")
    if not p2:
        g[:, :3, ODD[module]] *= -1
        g[:, -3:, ODD[module]] *= -1
    return g


@pytest.mark.parametrize("module", ["hydro", "hydro-isot", "disk"])
@pytest.mark.parametrize("p1,p2", [(1, 1), (0, 0), (1, 0), (0, 1)])
def test_bcond_brute_force(module, p1, p2):
    n1, n2 = 9, 7
    o = Oracle(module, n1, n2, config=False)
    o.set_geometry(l1=1.0, l2=1.0, p1=p1, p2=p2)
    rng = np.random.default_rng(p1 * 2 + p2)
    x = rng.standard_normal(o.shape)
    o.upload(x)
    o.bcond()
    g = o.download_ghosted()
    np.testing.assert_array_equal(g, expected(x, module, p1, p2))


def test_bcond_mirror_is_about_the_face():
    """Neumann1 copies row k to ghost row -(k+1): the mirror is about the
    cell face, not about cell 0 (which would copy row k+1)."""
    o = Oracle("hydro", 6, 6, config=False)
    o.set_geometry(p1=0, p2=0)
    x = np.arange(6 * 6 * 4, dtype=float).reshape(6, 6, 4)
    o.upload(x)
    o.bcond()
    g = o.download_ghosted()
    np.testing.assert_array_equal(g[2, 3:-3], x[0])      # ghost row -1 <- row 0
    np.testing.assert_array_equal(g[0, 3:-3], x[2])      # ghost row -3 <- row 2
    np.testing.assert_array_equal(g[3:-3, 2, 0], x[:, 0, 0])
    np.testing.assert_array_equal(g[3:-3, 2, 2], -x[:, 0, 2])
