"""Pin the oracle's kick / drift / RK3 step / first_order / time steps.

  - drift and kick's v <- beta v + F by element-wise brute force;
  - the RK3 step (src/step.cpp:22-43) on linear modes, where the exact
    discrete answer is the matrix polynomial P(dt L) = I + dtL + (dtL)^2/2
    + (dtL)^3/6 applied per step, with L the semi-discrete operator built
    from the closed-form modified wavenumbers;
  - the closed-form getdt (src/getdt.cpp:26) and the CFL constants.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from test_oracle_operators import grid, modified_k1, modified_k2


def rk3_poly(L, dt, n):
    Z = dt * L
    P = np.eye(len(L)) + Z + Z @ Z / 2 + Z @ Z @ Z / 6
    return np.linalg.matrix_power(P, n)


def test_drift_brute_force():
    o = Oracle("hydro", 8, 10)
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal(o.shape), rng.standard_normal(o.shape)
    o.upload(u)
    o.upload_v(v)
    o.drift(0.37)
    np.testing.assert_array_equal(o.download(), u + 0.37 * v)
    assert o.nan_flag() == 0
    v[3, 4, 1] = np.nan
    o.upload_v(v)
    o.drift(0.1)
    assert o.nan_flag() == 1


@pytest.mark.parametrize("module", ["hydro", "hydro-isot", "disk"])
def test_kick_is_beta_v_plus_rhs(module):
    from helpers import state
    o = Oracle(module, 12, 14)
    x = state(module, 12, 14)
    o.upload(x)
    v = np.random.default_rng(1).standard_normal(o.shape)
    F = o.rhs()
    o.upload_v(v)
    o.kick(-5.0 / 9.0)
    np.testing.assert_array_equal(o.download_v(), -5.0 / 9.0 * v + F)


def test_rk3_isothermal_sound_wave():
    """hydro-isot, nus = 0: lnd = eps cos(a x), u1 = 0 is a discrete sound
    wave: (a, b) with lnd = a cos, u1 = b sin obeys a' = -k b, b' = cs2 k a,
    k = k1(a h)/h."""
    n, eps, cs2 = 32, 1e-7, 1.7
    o = Oracle("hydro-isot", n, n, params={"nus": 0.0, "cs2": cs2})
    L1 = o.geometry()["l1"]
    X, Y = grid(n, n, L1, L1)
    m = 3
    a = 2 * np.pi * m / L1
    h = L1 / n
    k = modified_k1(a * h) / h
    x = np.zeros(o.shape)
    x[..., 0] = eps * np.cos(a * X)
    o.upload(x)
    dt, steps = 0.02, 25
    assert o.run(0.0, dt, steps) == 0
    L = np.array([[0.0, -k], [cs2 * k, 0.0]])
    amp = rk3_poly(L, dt, steps) @ np.array([eps, 0.0])
    y = o.download()
    np.testing.assert_allclose(y[..., 0], amp[0] * np.cos(a * X), atol=1e-6 * eps)
    np.testing.assert_allclose(y[..., 1], amp[1] * np.sin(a * X), atol=1e-6 * eps)
    # a wrong stage coefficient would miss this by far more than the tolerance
    wrong = rk3_poly(L * 1.0, dt * 1.01, steps) @ np.array([eps, 0.0])
    assert np.abs(wrong - amp).max() > 1e-3 * eps


def test_rk3_adiabatic_sound_wave_hydro():
    """hydro, nus = kappa = 0: (lnd, u1, lne) amplitudes obey d' = -k b,
    b' = T k (d + e), e' = -(gamma-1) k b with T = (gamma-1) e0."""
    n, eps, gamma = 32, 1e-7, 5.0 / 3.0
    e0 = 0.9
    o = Oracle("hydro", n, n, params={"nus": 0.0, "kappa": 0.0})
    X, Y = grid(n, n)
    a, h = 2 * np.pi * 2, 1.0 / n
    k = modified_k1(a * h) / h
    T = (gamma - 1) * e0
    x = np.zeros(o.shape)
    x[..., 0] = eps * np.cos(a * X)
    x[..., 3] = np.log(e0)
    o.upload(x)
    dt, steps = 0.004, 30
    assert o.run(0.0, dt, steps) == 0
    L = np.array([[0, -k, 0], [T * k, 0, T * k], [0, -(gamma - 1) * k, 0]])
    amp = rk3_poly(L, dt, steps) @ np.array([eps, 0.0, 0.0])
    y = o.download()
    np.testing.assert_allclose(y[..., 0], amp[0] * np.cos(a * X), atol=1e-6 * eps)
    np.testing.assert_allclose(y[..., 1], amp[1] * np.sin(a * X), atol=1e-6 * eps)
    np.testing.assert_allclose(y[..., 3] - np.log(e0), amp[2] * np.cos(a * X), atol=1e-6 * eps)


def test_rk3_viscous_shear_decay():
    """hydro-isot shear wave u1 = eps sin(b y): exactly linear,
    u1' = -nu k2(b h)/h^2 u1, so each step multiplies by 1 + z + z^2/2 + z^3/6."""
    n, nu = 32, 0.05
    o = Oracle("hydro-isot", n, n, params={"nus": nu})
    L1 = o.geometry()["l1"]
    X, Y = grid(n, n, L1, L1)
    b, h = 2 * np.pi * 5 / L1, L1 / n
    x = np.zeros(o.shape)
    x[..., 1] = 0.3 * np.sin(b * Y)
    o.upload(x)
    dt, steps = 0.01, 20
    assert o.run(0.0, dt, steps) == 0
    z = -nu * modified_k2(b * h) / h ** 2 * dt
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6
    np.testing.assert_allclose(o.download()[..., 1], 0.3 * R ** steps * np.sin(b * Y), atol=1e-12)


def test_rk3_thermal_diffusion_hydro():
    """hydro thermal conduction, (gamma-1) kappa on ln e, coupled to sound."""
    n, eps, gamma, kappa = 32, 1e-7, 5.0 / 3.0, 2e-3
    e0 = 0.9
    o = Oracle("hydro", n, n, params={"nus": 0.0, "kappa": kappa})
    X, Y = grid(n, n)
    a, h = 2 * np.pi * 3, 1.0 / n
    k, k2 = modified_k1(a * h) / h, modified_k2(a * h) / h ** 2
    T = (gamma - 1) * e0
    x = np.zeros(o.shape)
    x[..., 3] = np.log(e0) + eps * np.cos(a * X)
    o.upload(x)
    dt, steps = 0.004, 25
    assert o.run(0.0, dt, steps) == 0
    D = (gamma - 1) * kappa
    L = np.array([[0, -k, 0], [T * k, 0, T * k], [0, -(gamma - 1) * k, -D * k2]])
    amp = rk3_poly(L, dt, steps) @ np.array([0.0, 0.0, eps])
    y = o.download()
    np.testing.assert_allclose(y[..., 3] - np.log(e0), amp[2] * np.cos(a * X), atol=1e-6 * eps)
    np.testing.assert_allclose(y[..., 0], amp[0] * np.cos(a * X), atol=1e-6 * eps)


def test_getdt_closed_form():
    """dt = c / (n1/l1 + n2/l2) (src/getdt.cpp:26)."""
    o = Oracle("hydro", 256, 256)
    assert o.getdt() == 0.5 / 512
    o = Oracle("hydro", 100, 50, geometry=dict(l1=2.0, l2=0.5, c=0.3))
    assert abs(o.getdt() - 0.3 / (50 + 100)) < 1e-15


def test_rk3_stability_constants():
    R = lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6
    assert abs(abs(R(1j * np.sqrt(3))) - 1) < 1e-14
    assert abs(abs(R(-2.5127453266183286)) - 1) < 1e-12
    assert abs(R(-2.6)) > 1 and abs(R(1j * 1.8)) > 1


def test_hydro_first_order_is_identity():
    o = Oracle("hydro", 8, 8)
    x = np.random.default_rng(2).standard_normal(o.shape)
    o.upload(x)
    o.euler(0.1)
    np.testing.assert_array_equal(o.download(), x)


def test_disk_first_order_quartic_root():
    """disk radiation equilibrium (src/disk/Euler.cuh:19-69): when it fires,
    e_eq solves [ar (gamma-1)^4/rho] e^4 + e = e + e_r, i.e. with
    x = e_eq * scale, x^4 + x - x0 = 0, and e + e_r is conserved."""
    n = 6
    o = Oracle("disk", n, n, params={"ff": 1e3, "ar": 1e2})
    rng = np.random.default_rng(3)
    x = np.zeros(o.shape)
    x[..., 0] = rng.uniform(-1, 1, (n, n))
    x[..., 4] = rng.uniform(-3, -1, (n, n))
    x[..., 5] = rng.uniform(-3, -1, (n, n))
    o.upload(x)
    o.euler(1.0)
    y = o.download()
    g1 = 5.0 / 3.0 - 1
    rho, e, er = np.exp(x[..., 0]), np.exp(x[..., 4]), np.exp(x[..., 5])
    scale = (1e2 * g1 ** 4 / rho) ** (1 / 3)
    xe = np.exp(y[..., 4]) * scale
    x0 = (e + er) * scale
    fired = y[..., 4] != x[..., 4]
    assert fired.sum() > n * n // 2
    np.testing.assert_allclose((xe ** 4 + xe)[fired], x0[fired], rtol=1e-10)
    np.testing.assert_allclose((np.exp(y[..., 4]) + np.exp(y[..., 5]))[fired],
                               (e + er)[fired], rtol=1e-12)


def test_isot_random_phase_sequence():
    """hydro-isot draws two phases per Euler call from SplitMix64 (the seeded
    stand-in for rand() in src/hydro-isot/Euler.cuh:39-44)."""
    o = Oracle("hydro-isot", 4, 4, seed=42)
    o.upload(np.zeros(o.shape))
    o.euler(0.1)
    r = o.rand()
    M = (1 << 64) - 1

    def sm(s):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return s, z ^ (z >> 31)

    s, a = sm(42)
    s, b = sm(s)
    np.testing.assert_allclose(r, [6.2831853071795865 * (a >> 11) * 2.0 ** -53,
                                   6.2831853071795865 * (b >> 11) * 2.0 ** -53], rtol=1e-15)
    # forcing off (f = 0): the state is untouched
    np.testing.assert_array_equal(o.download(), 0.0)


def _forced(seed, dt, lnd=0.0, f=0.3, k=5.0, n1=24, n2=20):
    o = Oracle("hydro-isot", n1, n2, params={"f": f, "k": k}, seed=seed)
    x = np.zeros(o.shape)
    x[..., 0] = lnd
    o.upload(x)
    o.euler(dt)
    return o.download(), o.rand(), o.geometry()


def test_isot_forcing_pinned():
    """The stochastic forcing of src/hydro-isot/Euler.cuh:24-36, for the
    phases (r1, r2) the oracle drew: a kick along the fixed direction
    (cos r1, sin r1); sqrt(dt) cos(r2) cos(r1) f at the origin cell; exact
    sqrt(dt) and 1/rho scaling; a plane wave whose integer wavevector
    truncates k (cos r1, sin r1) + 1/2 toward zero, as the (Z) cast does."""
    f, k, dt = 0.3, 5.0, 0.01
    for seed in range(12):
        y, (r1, r2), g = _forced(seed, dt, f=f, k=k)
        du1, du2 = y[..., 1], y[..., 2]
        # direction: du is parallel to (cos r1, sin r1)
        np.testing.assert_allclose(du1 * np.sin(r1) - du2 * np.cos(r1), 0.0, atol=1e-15)
        # origin cell: phi = 0
        assert du1[0, 0] == pytest.approx(np.sqrt(dt) * f * np.cos(r2) * np.cos(r1), rel=1e-13, abs=1e-16)
        # sqrt(dt) and 1/rho scaling
        y4, _, _ = _forced(seed, 4 * dt, f=f, k=k)
        np.testing.assert_allclose(y4[..., 1:], 2.0 * y[..., 1:], rtol=1e-14, atol=1e-17)
        y2, _, _ = _forced(seed, dt, lnd=np.log(2.0), f=f, k=k)
        np.testing.assert_allclose(y2[..., 1:], 0.5 * y[..., 1:], rtol=1e-13, atol=1e-17)
        # wavevector: truncation toward zero of k n + 1/2
        m1, m2 = k * np.cos(r1), k * np.sin(r1)
        K1, K2 = int(m1 + 0.5), int(m2 + 0.5)             # Python int() truncates
        i = np.arange(y.shape[0])[:, None]
        j = np.arange(y.shape[1])[None, :]
        phi = 2 * np.pi * (K1 * i * g["Delta1"] + K2 * j * g["Delta2"])
        amp = np.sqrt(dt) * f * np.cos(phi + r2) / k
        np.testing.assert_allclose(du1, amp * m1, rtol=1e-10, atol=1e-15)
        if np.floor(m1 + 0.5) != K1 or np.floor(m2 + 0.5) != K2:
            F1, F2 = int(np.floor(m1 + 0.5)), int(np.floor(m2 + 0.5))
            phf = 2 * np.pi * (F1 * i * g["Delta1"] + F2 * j * g["Delta2"])
            assert not np.allclose(du1, np.sqrt(dt) * f * np.cos(phf + r2) / k * m1, atol=1e-12)
