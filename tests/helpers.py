"""Shared helpers of the test suite: error norms and workload states."""
import numpy as np

from fg2_amd import ics


def rel_l2(a, b):
    """Relative L2 error of a against reference b, per field (last axis).
    The normaliser is the L2 norm of the reference field; fields whose
    reference norm is below 1e-30 are normalised by 1."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    num = np.sqrt(np.sum((a - b) ** 2, axis=(0, 1)))
    den = np.sqrt(np.sum(b ** 2, axis=(0, 1)))
    return num / np.where(den > 1e-30, den, 1.0)


def rel_linf(a, b):
    """max |a - b| / max |b| per field."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    num = np.max(np.abs(a - b), axis=(0, 1))
    den = np.max(np.abs(b), axis=(0, 1))
    return num / np.where(den > 1e-30, den, 1.0)


# disk parameters of the test states: Newtonian point mass (rS = 0) and a
# large electron-scattering opacity, so the flux-limited radiative diffusion
# stays non-stiff in the low-density atmosphere (with the default es the
# radiation-energy equation needs dt ~ 1e-6 at these resolutions)
DISK_PARAMS = {"es": 1.0e6, "rS": 0.0}


def params_for(module):
    return dict(DISK_PARAMS) if module == "disk" else None


def state(module, n1, n2, seed=1):
    """A smooth seeded state of each module at any size."""
    if module == "hydro":
        return ics.hydro_random(n1, n2, seed=seed)
    if module == "hydro-isot":
        x = ics.isot_density(n1, n2, seed=seed)
        x[..., 1:3] = ics.gaussian_random_field(n1, n2, 4, 0.05, seed + 100, ncomp=2)
        return x
    return ics.disk_kepler(n1, n2, seed=seed, h_over_r=0.2)
