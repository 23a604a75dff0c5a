"""Seeded synthetic initial states for the benchmark and test workloads.

Host-side set-up (numpy), not part of the time-step path.  Coordinates are
cell centred, x = l1 (i + 1/2) / n1 along dim 1 and y = l2 (j + 1/2) / n2
along dim 2 (src/init.cpp:31-37).  The reference draws its random
perturbations with libc rand() (src/hydro/init.h:68-71,
src/hydro-isot/init.h:23-24); these generators use numpy's PCG64 with an
explicit seed so every workload is reproducible.
"""
import numpy as np

GAMMA = 5.0 / 3.0


def gaussian_random_field(n1, n2, kmax, rms, seed, ncomp=1, dtype=np.float64):
    """Real random field(s) on a periodic n1 x n2 grid built from every
    integer wavevector 0 < |k| <= kmax with independent complex normal
    amplitudes, scaled so the root-mean-square (over cells, of the vector
    norm when ncomp > 1) equals rms.  Synthesised as Re(E1 A E2^T) with
    E1[i, a] = exp(2 pi i k_a i / n1), E2[j, b] = exp(2 pi i k_b j / n2) and
    A the (2 kmax + 1)^2 amplitude matrix: one real matrix product of inner
    size 2 (2 kmax + 1) per component."""
    rng = np.random.default_rng(seed)
    K = int(np.floor(kmax))
    ks = np.arange(-K, K + 1)
    mask = (ks[:, None] ** 2 + ks[None, :] ** 2 <= kmax ** 2) & \
           ((ks[:, None] != 0) | (ks[None, :] != 0))
    E1 = np.exp(2j * np.pi * np.outer(np.arange(n1), ks) / n1)
    E2 = np.exp(2j * np.pi * np.outer(np.arange(n2), ks) / n2)
    right = np.concatenate([E2.real, E2.imag], axis=1).T.astype(dtype)   # 2(2K+1) x n2
    f = np.empty((n1, n2, ncomp), dtype)
    for q in range(ncomp):
        A = (rng.standard_normal(mask.shape) + 1j * rng.standard_normal(mask.shape)) * mask
        Pm = E1 @ A                                                        # n1 x (2K+1)
        left = np.concatenate([Pm.real, -Pm.imag], axis=1).astype(dtype)
        f[..., q] = left @ right
    norm = np.sqrt(np.mean(np.sum(f.astype(np.float64) ** 2, axis=-1))) if f.size < 2 ** 22 \
        else np.sqrt(np.mean(np.sum(f ** 2, axis=-1, dtype=np.float64)))
    f *= (rms / norm if norm > 0 else 0.0)
    return f


def hydro_random(n1, n2, seed=1, kmax=4, rms=0.01, gamma=GAMMA, dtype=np.float64):
    """hydro: rho = 1, p = 1/gamma (so c_s = 1), velocity a Gaussian random
    field with |k| <= kmax and rms |u| = rms * c_s.  Returns (n1, n2, 4)."""
    u = gaussian_random_field(n1, n2, kmax, rms, seed, ncomp=2, dtype=dtype)
    x = np.zeros((n1, n2, 4), dtype)
    x[..., 1:3] = u
    x[..., 3] = np.log((1.0 / gamma) / (gamma - 1.0))   # e = p / (rho (gamma - 1))
    return x


def hydro_kh(n1, n2, seed=2, amp=0.01, gamma=GAMMA, l1=1.0, l2=1.0):
    """Kelvin-Helmholtz shear layer on [-l/2, l/2]^2 (as src/hydro/init.h:62-63):
    rho = 2, u1 = +0.5 for |y| < 0.25, rho = 1, u1 = -0.5 elsewhere, p = 2.5,
    u2 = amp * N(0, 1) white noise.  Returns (n1, n2, 4)."""
    rng = np.random.default_rng(seed)
    y = l2 * (np.arange(n2) + 0.5) / n2 - 0.5 * l2
    inner = np.abs(y) < 0.25
    rho = np.where(inner, 2.0, 1.0)[None, :].repeat(n1, 0)
    x = np.zeros((n1, n2, 4))
    x[..., 0] = np.log(rho)
    x[..., 1] = np.where(inner, 0.5, -0.5)[None, :]
    x[..., 2] = amp * rng.standard_normal((n1, n2))
    x[..., 3] = np.log(2.5 / rho / (gamma - 1.0))
    return x


def isot_density(n1, n2, seed=3, kmax=8, rms=0.05, dtype=np.float64):
    """hydro-isot: lnd = ln(1 + delta), delta a Gaussian random field with
    |k| <= kmax and rms 0.05, u = 0.  Returns (n1, n2, 3)."""
    d = gaussian_random_field(n1, n2, kmax, rms, seed, dtype=dtype)[..., 0]
    x = np.zeros((n1, n2, 3), dtype)
    x[..., 0] = np.log1p(d)
    return x


def disk_kepler(n1, n2, seed=4, l1=None, rin=3.0, gm=1.0, h_over_r=0.05,
                noise=0.01, floor=1e-3, rad_frac=1e-3, gamma=GAMMA, eps=0.0):
    """disk: Keplerian disc around a central mass on the (log r, theta) grid
    of src/disk/config.h:21-34.  Sigma ~ R^-1 (rho_mid ~ R^-2) with a
    Gaussian vertical profile of aspect ratio h_over_r above a density floor
    of `floor` times the local midplane density, times (1 + noise N(0,1)).
    Omega = sqrt(GM / (r^2 + eps^2)^(3/2)) on spheres, the Keplerian value of
    a point mass with Plummer softening length eps (the disk parameter eps;
    the centrifugal force balances the cylindrical-radial gravity
    everywhere), isothermal sound speed h_over_r * v_K(R), ur = uz = 0, and
    radiation energy rad_frac times the thermal energy.  Returns (n1, n2, 6)."""
    rng = np.random.default_rng(seed)
    l2 = np.pi
    if l1 is None:
        hd2 = 0.5 * l2 / n2
        l1 = 2.0 * np.arcsinh(hd2) * n1
    lnr = l1 * (np.arange(n1) + 0.5) / n1
    th = l2 * (np.arange(n2) + 0.5) / n2
    sphr = rin * np.exp(lnr)[:, None]
    R = sphr * np.sin(th)[None, :]
    z = sphr * np.cos(th)[None, :]
    H = h_over_r * R
    rho_mid = R ** -2.0
    rho = rho_mid * (np.exp(-0.5 * (z / H) ** 2) + floor)
    rho = rho * (1.0 + noise * rng.standard_normal((n1, n2)))
    cs2 = h_over_r ** 2 * gm / R
    e = cs2 / (gamma - 1.0)                      # (gamma - 1) e = c_s^2
    x = np.zeros((n1, n2, 6))
    x[..., 0] = np.log(rho)
    x[..., 3] = np.sqrt(gm / (sphr ** 2 + eps ** 2) ** 1.5) * np.ones_like(R)
    x[..., 4] = np.log(e)
    x[..., 5] = np.log(rad_frac * e)
    return x


def log_uniform_sizes(rng, count, lo, hi, mult):
    """Grid sides drawn log-uniform on [lo, hi], rounded to a multiple of mult."""
    v = np.exp(rng.uniform(np.log(lo), np.log(hi), count))
    return np.clip((np.round(v / mult) * mult).astype(int), lo, hi)


def batched_workload(nruns=64, seed=7):
    """The batched workload: nruns hydro-isot runs with sides n1, n2 drawn
    independently log-uniform on [128, 1024] as multiples of 32, each
    starting from a seeded density perturbation (isot_density, rms 0.05,
    |k| <= 8, u = 0).  Returns (n1s, n2s, states)."""
    rng = np.random.default_rng(seed)
    n1s = log_uniform_sizes(rng, nruns, 128, 1024, 32)
    n2s = log_uniform_sizes(rng, nruns, 128, 1024, 32)
    xs = [isot_density(int(a), int(b), seed=seed + 1 + k) for k, (a, b) in enumerate(zip(n1s, n2s))]
    return n1s, n2s, xs


def serving_workload(nreq=1000, rate=100.0, seed=11):
    """The served request stream: Poisson arrivals at `rate` per second,
    module hydro or hydro-isot with equal odds, sides log-uniform on
    [64, 2048], step counts uniform on [50, 2000].  Returns (arrivals,
    modules, n1s, n2s, steps, state) where state(k) builds request k's
    seeded initial state (hydro_random or isot_density, FP32)."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    mods = ["hydro" if m == 0 else "hydro-isot" for m in rng.integers(0, 2, nreq)]
    n1s = log_uniform_sizes(rng, nreq, 64, 2048, 1)
    n2s = log_uniform_sizes(rng, nreq, 64, 2048, 1)
    steps = rng.integers(50, 2001, nreq)

    def state(k):
        a, b = int(n1s[k]), int(n2s[k])
        if mods[k] == "hydro":
            return hydro_random(a, b, seed=1000 + k, dtype=np.float32)
        return isot_density(a, b, seed=1000 + k, dtype=np.float32)
    return arrivals, mods, n1s, n2s, steps, state
