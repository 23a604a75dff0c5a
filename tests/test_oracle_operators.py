"""Pin the oracle's 6th-order operators (src/deriv.h:41-63) against what the
mathematics fixes: exactness on polynomials, the closed-form modified
wavenumbers of Fourier modes, and the mixed derivative at (pi, pi)."""
import numpy as np
import pytest

from oracle.oracle import Oracle


def grid(n1, n2, l1=1.0, l2=1.0):
    x = l1 * (np.arange(n1) + 0.5) / n1
    y = l2 * (np.arange(n2) + 0.5) / n2
    return np.meshgrid(x, y, indexing="ij")


def interior(a, m=3):
    return a[m:-m, m:-m]


@pytest.mark.parametrize("p,q", [(0, 0), (1, 0), (3, 2), (6, 0), (0, 6), (2, 4), (5, 1)])
def test_first_derivative_polynomial_exact(p, q):
    """D1, D2 are exact for degree <= 6 (away from the boundary)."""
    n = 24
    o = Oracle("hydro", n, n, geometry=dict(l1=1.3, l2=0.9, p1=0, p2=0), config=False)
    o.config()
    X, Y = grid(n, n, 1.3, 0.9)
    f = np.zeros(o.shape)
    f[..., 0] = X ** p * Y ** q
    o.upload(f)
    d1 = o.deriv("d1")[..., 0]
    d2 = o.deriv("d2")[..., 0]
    e1 = p * X ** max(p - 1, 0) * Y ** q if p else 0 * X
    e2 = q * X ** p * Y ** max(q - 1, 0) if q else 0 * X
    np.testing.assert_allclose(interior(d1), interior(e1), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(interior(d2), interior(e2), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("p,q", [(2, 0), (7, 0), (0, 7), (4, 3), (1, 1), (3, 4), (6, 1), (2, 5)])
def test_second_derivatives_polynomial_exact(p, q):
    """D11, D22 exact to degree 7; D12 exact for every x^p y^q with p+q <= 7."""
    n = 24
    o = Oracle("hydro", n, n, geometry=dict(l1=1.0, l2=1.2, p1=0, p2=0))
    X, Y = grid(n, n, 1.0, 1.2)
    f = np.zeros(o.shape)
    f[..., 1] = X ** p * Y ** q
    o.upload(f)

    def dpow(k, m):
        return k * (k - 1) if k >= 2 else 0

    e11 = dpow(p, 0) * X ** max(p - 2, 0) * Y ** q
    e22 = dpow(q, 0) * X ** p * Y ** max(q - 2, 0)
    e12 = p * q * X ** max(p - 1, 0) * Y ** max(q - 1, 0)
    np.testing.assert_allclose(interior(o.deriv("d11")[..., 1]), interior(e11), rtol=1e-7, atol=1e-6)
    np.testing.assert_allclose(interior(o.deriv("d22")[..., 1]), interior(e22), rtol=1e-7, atol=1e-6)
    np.testing.assert_allclose(interior(o.deriv("d12")[..., 1]), interior(e12), rtol=1e-7, atol=1e-6)


def test_mixed_derivative_fails_at_degree_8():
    """The D12 stencil uses only diagonal points; x^1 y^7 is the first
    monomial it does not differentiate exactly (a plausible-mistake guard:
    a full 7x7 stencil would pass this)."""
    n = 24
    o = Oracle("hydro", n, n, geometry=dict(l1=1.0, l2=1.0, p1=0, p2=0))
    X, Y = grid(n, n)
    f = np.zeros(o.shape)
    f[..., 1] = X * Y ** 7
    o.upload(f)
    err = np.abs(interior(o.deriv("d12")[..., 1]) - interior(7 * Y ** 6)).max()
    assert err > 1e-9   # exact cases stay below ~1e-12 here


def modified_k1(kh):
    return (45 * np.sin(kh) - 9 * np.sin(2 * kh) + np.sin(3 * kh)) / 30.0


def modified_k2(kh):
    return (490 - 540 * np.cos(kh) + 54 * np.cos(2 * kh) - 4 * np.cos(3 * kh)) / 180.0


@pytest.mark.parametrize("m1,m2", [(1, 0), (3, 5), (7, 2), (16, 16), (31, 1)])
def test_fourier_modified_wavenumbers(m1, m2):
    """On a periodic grid f = sin(a x + b y): D1 f = k1(a h1)/h1 cos, D11 f =
    -k2(a h1)/h1^2 f, D12 f = sum_k c_k (-4 sin(a k h1) sin(b k h2))/(h1 h2) f."""
    n1, n2 = 64, 48
    l1, l2 = 1.0, 0.75
    o = Oracle("hydro", n1, n2, geometry=dict(l1=l1, l2=l2, p1=1, p2=1))
    X, Y = grid(n1, n2, l1, l2)
    a, b = 2 * np.pi * m1 / l1, 2 * np.pi * (m2 % (n2 // 2)) / l2
    h1, h2 = l1 / n1, l2 / n2
    phi = a * X + b * Y
    f = np.zeros(o.shape)
    f[..., 0] = np.sin(phi)
    o.upload(f)
    np.testing.assert_allclose(o.deriv("d1")[..., 0], modified_k1(a * h1) / h1 * np.cos(phi), atol=1e-9)
    np.testing.assert_allclose(o.deriv("d2")[..., 0], modified_k1(b * h2) / h2 * np.cos(phi), atol=1e-9)
    np.testing.assert_allclose(o.deriv("d11")[..., 0], -modified_k2(a * h1) / h1 ** 2 * np.sin(phi), atol=1e-7)
    np.testing.assert_allclose(o.deriv("d22")[..., 0], -modified_k2(b * h2) / h2 ** 2 * np.sin(phi), atol=1e-7)
    c = np.array([270.0, -27.0, 2.0]) / 720.0
    k = np.arange(1, 4)
    m12 = np.sum(c * (-4) * np.sin(a * k * h1) * np.sin(b * k * h2)) / (h1 * h2)
    np.testing.assert_allclose(o.deriv("d12")[..., 0], m12 * np.sin(phi), atol=1e-7)


def test_nyquist_values():
    """D11 at Nyquist is -1088/180/h^2; D12 at (pi, pi) is 0."""
    n = 16
    o = Oracle("hydro", n, n, geometry=dict(l1=1.0, l2=1.0, p1=1, p2=1))
    I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    f = np.zeros(o.shape)
    f[..., 0] = (-1.0) ** (I + J)
    o.upload(f)
    h = 1.0 / n
    np.testing.assert_allclose(o.deriv("d11")[..., 0], -1088 / 180 / h ** 2 * f[..., 0], rtol=1e-12)
    np.testing.assert_allclose(o.deriv("d12")[..., 0], 0.0, atol=1e-9)


def test_kmax_constant():
    """The CFL constant: max over kh of the D1 modified wavenumber."""
    kh = np.linspace(0, np.pi, 2_000_001)
    assert abs(modified_k1(kh).max() - 1.5859783962656393) < 1e-11
    assert abs(modified_k2(np.pi) - 1088 / 180) < 1e-12
