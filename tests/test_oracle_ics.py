"""Pinning of the oracle's initial conditions (src/init.cpp:26-42 and the
modules' init.h) by what the physics and the published RNG fix, not by
restating the formulas:

  SplitMix64      published outputs for seeds 1234567 and 0
  draw order      KH consumes one draw per cell and Kf two, row-major (the
                  reference's rand() call order), mapped to ((z>>11)+1/2) 2^-53
  hydro           uniform pressure of KH (2.5) and bow (0.01); blast's
                  pressure jump at r^2 = 0.01; Sod's stripes; implos' and
                  bow's limiting values; unknown names give zeros
  hydro-isot Kf   Box-Muller radius q1^2 + q2^2 = -2 ln r1; the Kolmogorov
                  shear f / (nus k^2) cos(k x)
  disk (Hawley)   polytrope p = K rho^gamma everywhere; the Bernoulli
                  constant gamma e - M/r + l^2 / (2 R^2) inside the torus;
                  constant specific angular momentum inside, the sin^2
                  scaled background outside; density 1 at the pressure
                  maximum r0 = 16 on the midplane
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, splitmix64

G53 = 2.0 ** -53


def unit(z):
    return ((z >> 11) + 0.5) * G53


def centres(o):
    g = o.geometry()
    x = np.float32(g["l1"]) * (np.arange(o.n1) + 0.5) / o.n1
    y = np.float32(g["l2"]) * (np.arange(o.n2) + 0.5) / o.n2
    X, Y = np.meshgrid(x.astype(np.float32).astype(np.float64),
                       y.astype(np.float32).astype(np.float64), indexing="ij")
    return g, X, Y


def test_splitmix64_published_outputs():
    # Vigna's splitmix64.c, seed 1234567; and the first output from state 0
    assert splitmix64(1234567, 5) == [6457827717110365317, 3203168211198807973,
                                      9817491932198370423, 4593380528125082431,
                                      16408922859458223821]
    assert splitmix64(0, 1) == [0xE220A8397B1DCDAF]


GAMMA = float(np.float32(5.0 / 3.0))   # para_gamma is R-typed


def pressure(x):
    return (GAMMA - 1.0) * np.exp(x[..., 0] + x[..., 3])


def test_hydro_kh():
    n1, n2, seed = 24, 40, 77
    o = Oracle("hydro", n1, n2)
    o.init("KH", seed)
    x = o.download()
    g, X, Y = centres(o)
    inside = np.abs(np.float32(Y - 0.5 * g["l2"])) < 0.25
    assert np.allclose(pressure(x), 2.5, rtol=1e-14)
    assert np.allclose(np.exp(x[..., 0]), np.where(inside, 2.0, 1.0), rtol=1e-14)
    # one draw per cell, row-major over (i, j)
    draws = np.array([unit(z) for z in splitmix64(seed, n1 * n2)]).reshape(n1, n2)
    expect = np.where(inside, 0.5, -0.5) + 0.01 * draws - 0.005
    assert np.allclose(x[..., 1], expect, rtol=0, atol=1e-15)
    assert np.all(x[..., 2] == 0.0)
    # another seed gives other perturbations, the same mean flow
    o2 = Oracle("hydro", n1, n2)
    o2.init("KH", seed + 1)
    x2 = o2.download()
    assert not np.allclose(x2[..., 1], x[..., 1], atol=1e-9)
    assert np.max(np.abs(x2[..., 1] - x[..., 1])) <= 0.01


def test_hydro_blast_sod_and_zeros():
    n = 64
    o = Oracle("hydro", n, n)
    g, X, Y = centres(o)
    xx = np.float32(X - 0.5).astype(np.float64)
    yy = np.float32(Y - 0.5).astype(np.float64)

    o.init("blast")
    x = o.download()
    hot = xx * xx + yy * yy < 0.01
    assert hot.sum() == np.sum((np.arange(n) + 0.5 - n / 2)[:, None] ** 2 +
                               (np.arange(n) + 0.5 - n / 2)[None, :] ** 2 < 0.01 * n * n)
    assert np.all(x[..., 0] == 0.0) and np.all(x[..., 1] == 0.0)
    assert np.allclose(pressure(x), np.where(hot, 1.0, 0.01), rtol=1e-14)

    o.init("Sod")
    x = o.download()
    left = np.fmod(xx - 2 * yy + 1.5, 1.0) < 0.5
    assert np.allclose(np.exp(x[..., 0]), np.where(left, 1.0, 0.125), rtol=1e-14)
    assert np.allclose(pressure(x), np.where(left, 1.0, 0.1), rtol=1e-14)
    # stripes of slope 1/2 in (x, y): shifting x by 2 dy maps the pattern onto itself
    assert left[8, 3] == left[8 + 2 * 5, 3 + 5]

    for name in ("zeros", "default", "no-such-ic"):
        o.init(name)
        assert np.all(o.download() == 0.0)


def test_hydro_bow_and_implos_limits():
    n = 512
    o = Oracle("hydro", n, n)
    g, X, Y = centres(o)
    r = np.hypot(X - 0.5, Y - 0.5)

    o.init("bow")
    x = o.download()
    d = np.exp(x[..., 0])
    assert np.allclose(pressure(x), 0.01, rtol=1e-13)          # uniform pressure
    far = r > 0.1
    assert np.allclose(d[far], 0.1, rtol=1e-12)                 # ambient density
    assert np.allclose(x[..., 1][far], 1.0, rtol=1e-12)         # (10/0.1 - 1)/99
    core = r < 0.005
    assert core.any() and np.allclose(d[core], 10.0, rtol=1e-2)
    assert np.all(np.abs(x[..., 1][core]) < 2e-3)

    o.init("implos")
    x = o.download()
    d, p = np.exp(x[..., 0]), pressure(x)
    dist = np.abs(X - 0.5) + np.abs(Y - 0.5)
    # the low-pressure diamond |x| + |y| < 0.15 inside a dense, hot surrounding
    assert np.allclose(d[dist > 0.2], 1.0, rtol=1e-12)
    assert np.allclose(p[dist > 0.2], 1.0, rtol=1e-12)
    assert np.allclose(d[dist < 0.1], 0.125, rtol=1e-12)
    assert np.allclose(p[dist < 0.1], 0.14, rtol=1e-12)
    # density and pressure share one Fermi-Dirac profile
    assert np.allclose((d - 0.125) / 0.875, (p - 0.14) / 0.86, rtol=1e-10, atol=1e-12)


def test_isot_kf():
    n1, n2, seed = 64, 48, 5
    f, k = 0.02, 4.0
    o = Oracle("hydro-isot", n1, n2, params={"f": f, "k": k})
    o.init("Kf", seed)
    x = o.download()
    g, X, Y = centres(o)
    nus = float(np.float32(2e-4))
    a = float(np.float32(f)) / (nus * k * k)
    assert np.all(x[..., 0] == 0.0)
    shear = a * np.cos(k * X)
    q1 = x[..., 1] * 1e9
    q2 = (x[..., 2] - shear) * 1e9
    # the noise is 1e-9 N(0, 1): recover the normal pair and its draws
    draws = np.array([unit(z) for z in splitmix64(seed, 2 * n1 * n2)]).reshape(n1, n2, 2)
    assert np.allclose(q1 ** 2 + q2 ** 2, -2.0 * np.log(draws[..., 0]), rtol=1e-5, atol=1e-5)
    assert np.allclose(np.arctan2(q1, q2) % (2 * np.pi), (2 * np.pi * draws[..., 1]) % (2 * np.pi),
                       atol=1e-4)
    assert abs(np.std(q1) - 1.0) < 0.1 and abs(np.mean(q1)) < 0.1
    # unknown names give zeros
    o.init("zeros")
    assert np.all(o.download() == 0.0)


def test_disk_hawley_torus():
    n = 256
    o = Oracle("disk", n, n)
    o.init("anything")
    x = o.download()
    g, X, Y = centres(o)
    M, rin, ar = 1.0, 3.0, 1.0e8
    r = rin * np.exp(X)
    R = r * np.sin(Y)
    den, Omg, eng = np.exp(x[..., 0]), x[..., 3], np.exp(x[..., 4])
    e0, r0, bg = 0.01, 16.0, 1e-4
    K = (GAMMA - 1.0) * e0
    lK = np.sqrt(M * r0)
    # polytrope everywhere
    assert np.allclose((GAMMA - 1.0) * eng * den, K * den ** GAMMA, rtol=1e-10)
    torus = den > bg * (1 + 1e-9)
    assert 100 < torus.sum() < n * n // 2
    # Bernoulli constant inside the torus
    bern = GAMMA * eng - M / r + lK ** 2 / (2.0 * R ** 2)
    assert np.allclose(bern[torus], GAMMA * e0 - M / r0 + lK ** 2 / (2.0 * r0 ** 2), rtol=1e-9)
    # constant specific angular momentum inside, sin^2-scaled outside
    assert np.allclose((Omg * R ** 2)[torus], lK, rtol=1e-9)
    assert np.allclose((Omg * R ** 2)[~torus], (lK * np.sin(Y) ** 2)[~torus], rtol=1e-9)
    assert np.allclose(den[~torus], bg, rtol=1e-12)
    # pressure maximum at r0 on the midplane, where rho = d0 = 1
    i0 = np.argmin(np.abs(r[:, 0] - r0))
    j0 = np.argmin(np.abs(Y[0] - np.pi / 2))
    assert abs(den[i0, j0] - 1.0) < 0.02
    assert den.max() <= 1.0 + 1e-12
    assert np.all(x[..., 1] == 0.0) and np.all(x[..., 2] == 0.0)
    tbg = K * bg ** (GAMMA - 1.0)
    assert np.allclose(x[..., 5], np.log(ar * tbg ** 4 / bg), rtol=1e-12)
