"""Pin the oracle's module right-hand sides (src/hydro/eqns.cuh,
src/hydro-isot/eqns.cuh, src/disk/eqns.cuh).

  - manufactured smooth fields: the oracle's discrete F(u) against sympy's
    exact-derivative F (fixtures from tests/golden/make_manufactured.py),
    converging at sixth order between n = 32 and n = 64 (for disk the R, z
    derivatives come from sympy's chain rule, so the (log r, theta) -> (R, z)
    transforms of the oracle are checked independently);
  - a uniform state gives F = 0 (hydro, hydro-isot);
  - u1 <-> u2 with the grid transposed gives the transposed F;
  - disk: a uniform state at rest with Omega^2 R = GM sin(theta)/(r - rS)^2
    balances the radial equation exactly; the vertical one feels gravity only.
"""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "manufactured.npz"))

import importlib.util
_spec = importlib.util.spec_from_file_location(
    "make_manufactured", os.path.join(HERE, "golden", "make_manufactured.py"))
MM = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(MM)


def discrete_rhs(module, key, n, params, geometry=None):
    o = Oracle(module, n, n, params=params, geometry=geometry)
    o.upload(FIX[f"{key}_{n}_state"])
    return o.rhs()


@pytest.mark.parametrize("module,key,params,margin", [
    ("hydro", "hydro", MM.HYDRO_PAR, 0),
    ("hydro-isot", "isot", MM.ISOT_PAR, 0),
    ("disk", "disk", MM.DISK_PAR, 6),
])
def test_manufactured_sixth_order(module, key, params, margin):
    errs = []
    for n in (32, 64):
        F = discrete_rhs(module, key, n, params)
        E = FIX[f"{key}_{n}_rhs"]
        m = margin * n // 32
        sl = (slice(m, n - m), slice(m, n - m)) if m else (slice(None), slice(None))
        err = np.abs(F[sl] - E[sl]).max(axis=(0, 1))
        scale = np.abs(E[sl]).max(axis=(0, 1))
        errs.append(err / scale)
    e32, e64 = errs
    print(module, "rel err n=32", e32, "n=64", e64, "ratio", e32 / e64)
    assert np.all(e64 < 1e-4), e64
    # sixth order: the error drops by 2^6 = 64 from n = 32 to 64 (allow 30)
    assert np.all(e32 / e64 > 30), e32 / e64


@pytest.mark.parametrize("module", ["hydro", "hydro-isot"])
def test_uniform_state_has_zero_rhs(module):
    o = Oracle(module, 12, 10, params={"dd": 1e-3, "nub": 1e-3})
    x = np.zeros(o.shape)
    x[..., 0], x[..., 1], x[..., 2] = 0.3, 0.1, -0.2
    if module == "hydro":
        x[..., 3] = np.log(0.9)
    o.upload(x)
    assert np.abs(o.rhs()).max() < 1e-12


def test_transpose_symmetry_hydro():
    n = 24
    o = Oracle("hydro", n, n, params={"dd": 1e-3, "nub": 1e-3})
    x = MM.evaluate(MM.hydro_fields(), *np.meshgrid((np.arange(n) + .5) / n,
                                                    (np.arange(n) + .5) / n, indexing="ij"))
    xt = np.transpose(x, (1, 0, 2))[..., [0, 2, 1, 3]]
    o.upload(x)
    F = o.rhs()
    o.upload(xt)
    Ft = o.rhs()
    np.testing.assert_allclose(Ft, np.transpose(F, (1, 0, 2))[..., [0, 2, 1, 3]],
                               rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("eps", [0.0, 0.4])
def test_disk_rotational_balance(eps):
    """At rest with Omega^2 R = g sin(theta), g = M d / (d^2 + eps^2)^(3/2)
    and d = r - rS (eps = 0: the reference's M / d^2; eps > 0: the added
    Plummer softening), the radial equation balances and the vertical one
    feels gravity only."""
    n, M, rS = 16, 1.3, 0.5
    o = Oracle("disk", n, n, params={"M": M, "rS": rS, "ff": 0.0, "alpha": 0.0, "nu": 0.0,
                                     "eps": eps})
    g = o.geometry()
    lnr = g["l1"] * (np.arange(n) + 0.5) / n
    th = g["l2"] * (np.arange(n) + 0.5) / n
    sphr = 3.0 * np.exp(lnr)[:, None]
    R = sphr * np.sin(th)[None, :]
    x = np.zeros(o.shape)
    x[..., 0], x[..., 4], x[..., 5] = -1.0, np.log(0.2), np.log(0.05)
    d = sphr - rS
    grav = M * d / (d * d + eps * eps) ** 1.5
    x[..., 3] = np.sqrt(grav * np.sin(th)[None, :] / R)
    o.upload(x)
    F = o.rhs()
    assert np.abs(F[..., 1]).max() < 1e-12 * np.abs(grav).max()
    np.testing.assert_allclose(F[..., 2], -np.cos(th)[None, :] * grav * np.ones((n, 1)), rtol=1e-12, atol=1e-14)
    assert np.abs(F[..., 0]).max() < 1e-14 and np.abs(F[..., 3]).max() < 1e-14


def test_kolmogorov_forcing():
    """hydro-isot with f k < 0 on a state at rest: F_u2 = f cos(k i Delta1)
    with the cell index i (src/hydro-isot/eqns.cuh:81-84)."""
    n = 16
    o = Oracle("hydro-isot", n, n, params={"f": 0.2, "k": -3.0})
    o.upload(np.zeros(o.shape))
    F = o.rhs()
    d1 = o.geometry()["Delta1"]
    i = np.arange(n)[:, None] * np.ones((1, n))
    np.testing.assert_allclose(F[..., 2], 0.2 * np.cos(-3.0 * i * d1), rtol=1e-12)
    assert np.abs(F[..., :2]).max() == 0.0


def test_disk_flux_limiter_limits():
    """The radiative flux limiter of the disk module, Lambda(R) =
    (2 + R) / (6 + 3R + R^2) (src/disk/eqns.cuh:194-196), is pinned by its
    physical limits, not by a second restatement: the diffusion (Eddington)
    limit Lambda(0) = 1/3, the free-streaming limit R Lambda(R) -> 1, and
    Lambda <= 1/3 throughout.  Its derivative (src/disk/eqns.cuh:216) is
    checked against a central difference of Lambda and against its own
    limits, Lambda'(0) = 0 and R^2 Lambda'(R) -> -1."""
    from oracle.oracle import flux_limiter
    lam0, dlam0 = flux_limiter(0.0)
    assert lam0 == 1.0 / 3.0 and dlam0 == 0.0
    for R in (1e6, 1e8, 1e10):
        lam, dlam = flux_limiter(R)
        assert abs(R * lam - 1.0) < 2.0 / R          # R Lambda = 1 - 1/R + O(1/R^2)
        assert abs(R * R * dlam + 1.0) < 3.0 / R     # R^2 Lambda' = -1 + 2/R + O(1/R^2)
    for R in (1e-3, 0.1, 0.7, 1.0, 3.0, 25.0):
        lam, dlam = flux_limiter(R)
        assert 0.0 < lam <= 1.0 / 3.0 and dlam < 0.0
        h = 1e-5 * max(R, 1.0)
        fd = (flux_limiter(R + h)[0] - flux_limiter(R - h)[0]) / (2 * h)
        assert abs(fd - dlam) < 1e-8 * max(1.0, abs(dlam)), (R, fd, dlam)
    # the FP32 instantiation rounds the same expression
    lam32, _ = flux_limiter(0.5, fp64=False)
    assert abs(lam32 - flux_limiter(0.5)[0]) < 1e-7
