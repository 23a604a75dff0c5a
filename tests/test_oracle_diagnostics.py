"""Pinning of the spectral-diagnostics oracle (oracle/diagnostics.py) by a
direct DFT, closed forms of single Fourier modes, Parseval and brute-force
bin edges."""
import numpy as np
import pytest

from oracle import diagnostics as D


def direct_dft(u):
    """u^(k) = (1/N) sum_x u(x) exp(-2 pi i (k1 i / n1 + k2 j / n2)) by brute force."""
    n1, n2 = u.shape
    i = np.arange(n1)[:, None, None, None]
    j = np.arange(n2)[None, :, None, None]
    k1 = np.arange(n1)[None, None, :, None]
    k2 = np.arange(n2)[None, None, None, :]
    ph = np.exp(-2j * np.pi * (k1 * i / n1 + k2 * j / n2))
    return np.einsum("ij,ijkl->kl", u, ph) / (n1 * n2)


@pytest.mark.parametrize("n1,n2", [(6, 10), (7, 5), (8, 8)])
def test_against_direct_dft(n1, n2):
    rng = np.random.default_rng(n1 * 100 + n2)
    u1, u2 = rng.standard_normal((2, n1, n2))
    a, b = direct_dft(u1), direct_dft(u2)
    E = (np.abs(a) ** 2 + np.abs(b) ** 2) / 2
    k1, k2 = D.getk(n1, n2)
    # brute-force shells: every wavevector by hand
    kG = min(n1, n2) // 2 + 0.99
    n = int(np.ceil(kG))
    Eb = np.zeros(n)
    for p in range(n1):
        for q in range(n2):
            kk = np.float32(np.sqrt(np.float32(k1[p, q] ** 2 + k2[p, q] ** 2)))
            if kk == 0:
                continue
            for s in range(n):
                lo, hi = s + 0.5, s + 1.5
                if (s == 0 and 0 < kk < hi) or (s == n - 1 and s > 0 and lo <= kk <= hi) or \
                   (0 < s < n - 1 and lo <= kk < hi):
                    Eb[s] += E[p, q]
    got = D.spectrum(u1, u2)
    np.testing.assert_allclose(got["E"], Eb, rtol=1e-12, atol=1e-15)


def test_getk_layout():
    k1, k2 = D.getk(6, 5)
    assert list(k1[:, 0]) == [0, 1, 2, -3, -2, -1]
    assert list(k2[0]) == [0, 1, 2, -2, -1]
    z1, z2 = D.getk(6, 5, zeronyquist=True)
    assert list(z1[:, 0]) == [0, 1, 2, 0, -2, -1] and list(z2[0]) == [0, 1, 2, -2, -1]


@pytest.mark.parametrize("a,b", [(3, 4), (0, 5), (7, 1)])
def test_single_mode_closed_forms(a, b):
    n1, n2 = 64, 48
    i = np.arange(n1)[:, None]
    j = np.arange(n2)[None, :]
    th = 2 * np.pi * (a * i / n1 + b * j / n2)
    u1 = np.cos(th)
    u2 = np.zeros_like(u1)
    s = D.spectrum(u1, u2)
    shell = int(np.floor(np.hypot(a, b) - 0.5))
    # u1^ = 1/2 at +-(a, b): E = 2 x (1/2)(1/4) = 1/4, Z = 2 x (1/2)(b^2/4) = b^2/4
    assert s["E"][shell] == pytest.approx(0.25, rel=1e-12)
    assert s["Z"][shell] == pytest.approx(b * b / 4, rel=1e-12)
    assert np.sum(s["E"]) == pytest.approx(0.25, rel=1e-12)
    # vorticity = -d u1 / d y in integer-wavenumber units = b sin(theta)
    np.testing.assert_allclose(D.vorticity(u1, u2), b * np.sin(th), atol=1e-12)


def test_parseval_band_limited():
    n1, n2 = 96, 80
    rng = np.random.default_rng(5)
    k1, k2 = D.getk(n1, n2)
    mask = (np.sqrt(k1 ** 2 + k2 ** 2) < 20) & (np.sqrt(k1 ** 2 + k2 ** 2) > 0)
    f = (rng.standard_normal((n1, n2)) + 1j * rng.standard_normal((n1, n2))) * mask
    u1 = np.real(np.fft.ifft2(f))
    u2 = np.imag(np.fft.ifft2(f))
    s = D.spectrum(u1, u2)
    assert np.sum(s["E"]) == pytest.approx(0.5 * np.mean(u1 ** 2 + u2 ** 2), rel=1e-12)
    assert np.sum(s["counts"]) == np.count_nonzero(
        (np.sqrt(k1 ** 2 + k2 ** 2) > 0) & (np.sqrt(k1 ** 2 + k2 ** 2) <= len(s["E"]) + 0.5))
