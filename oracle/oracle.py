"""ctypes wrapper of the host oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module; it is what GPU results are compared with and
the CPU baseline, never the thing measured or shipped.  The product package
(fg2_amd) does not import it.

The oracle restates the reference's time-step path on the host, templated on
the storage type: ``fp64=True`` is the FP64 oracle, ``fp64=False`` the FP32 CPU
baseline.  Each function of oracle.cpp cites the reference file:line it follows.

Pinning (which test fixes each function against something other than itself):

  operators D1 D2 D11 D22 D12 (src/deriv.h:41-63)
      polynomial exactness to degree 6/7, Fourier modified wavenumbers in
      closed form, D12 at (pi, pi) = 0                  tests/test_oracle_operators.py
  bcond (src/bcond.cu:22-86)
      brute force against numpy roll / mirror, corners    tests/test_oracle_bcond.py
  drift (src/drift.cu:23-32), kick's beta*v + F order (src/kick.cu:41-57)
      element-wise brute force                            tests/test_oracle_step.py
  RK3 step (src/step.cpp:22-43)
      linear-mode amplification 1 + z + z^2/2 + z^3/6     tests/test_oracle_step.py
  hydro, hydro-isot RHS (src/hydro/eqns.cuh, src/hydro-isot/eqns.cuh)
      uniform state -> 0, transpose symmetry, manufactured smooth fields
      against sympy's continuous RHS (committed fixtures), sound speed,
      viscous and thermal decay rates, Kolmogorov forcing  tests/test_oracle_eqns.py
  disk RHS (src/disk/eqns.cuh)
      uniform state -> gravity only, Keplerian balance of the radial
      equation, manufactured smooth fields against sympy's continuous RHS
      (chain rule from (log r, theta) to (R, z))           tests/test_oracle_eqns.py
  disk flux limiter Lambda(R), Lambda'(R) (src/disk/eqns.cuh:194-196,216)
      Lambda(0) = 1/3 (Eddington), R Lambda(R) -> 1 (free streaming),
      Lambda' against a central difference and its closed-form limits
      0 and -1/R^2                                     tests/test_oracle_eqns.py
  first_order / Euler (src/*/Euler.cuh)
      hydro: identity.  disk: the quartic root x^4 + x - x0 = 0 by
      substitution.  hydro-isot forcing, for the phases drawn: kick along
      (cos r1, sin r1), its value at the origin cell, exact sqrt(dt) and
      1/rho scaling, the wavevector's truncation toward zero; the phase
      stream itself is SplitMix64 in place of libc rand() (the reference's
      own rand() sequence is not reproducible)         tests/test_oracle_step.py
  getdt closed form (src/getdt.cpp:26)
      closed form values                                  tests/test_oracle_step.py
  CFL time step (new; src/getdt.cpp:26-28 is a TODO)
      stability constants: |R(i sqrt3)| = 1, |R(-2.5127)| = 1, max of the
      D1 modified wavenumber                              tests/test_oracle_step.py
  SplitMix64 (the seeded stream in place of libc rand())
      the published reference outputs of seed 1234567     tests/test_oracle_ics.py
  initial conditions (src/init.cpp:26-42, src/*/init.h)
      KH: the two-layer density, pressure and shear, the noise's bound;
      blast, Sod, zeros: the piecewise states (hot-cell count by brute
      force); bow, implos: their limits away from and at the centre; Kf:
      Box-Muller velocities from the stream's draws (radius and angle);
      Hawley torus: the polytrope, the Bernoulli constant and the constant
      specific angular momentum inside the torus      tests/test_oracle_ics.py

The spectral diagnostics have their own restatement and pins in
oracle/diagnostics.py.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MODULES = {"hydro": 0, "hydro-isot": 1, "disk": 2}
NVAR = {"hydro": 4, "hydro-isot": 3, "disk": 6}


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        DP = ctypes.POINTER(ctypes.c_double)
        sig = {
            "orc_create": (P, [I, I, I, I]),
            "orc_destroy": (None, [P]),
            "orc_nvar": (I, [P]),
            "orc_set_param": (I, [P, ctypes.c_char_p, D]),
            "orc_get_param": (D, [P, ctypes.c_char_p]),
            "orc_set_geometry": (None, [P, D, D, I, I, D]),
            "orc_get_geometry": (None, [P, DP]),
            "orc_config": (None, [P]),
            "orc_seed": (None, [P, ctypes.c_ulonglong]),
            "orc_get_rand": (None, [P, DP]),
            "orc_upload": (None, [P, DP]),
            "orc_download": (None, [P, DP]),
            "orc_download_ghosted": (None, [P, DP]),
            "orc_upload_v": (None, [P, DP]),
            "orc_download_v": (None, [P, DP]),
            "orc_bcond": (None, [P]),
            "orc_kick": (None, [P, D]),
            "orc_drift": (None, [P, D]),
            "orc_euler": (None, [P, D]),
            "orc_rhs": (None, [P, DP]),
            "orc_deriv": (None, [P, I, DP]),
            "orc_step": (I, [P, D, D]),
            "orc_run": (I, [P, D, D, I]),
            "orc_getdt": (D, [P]),
            "orc_getdt_cfl": (D, [P]),
            "orc_nan_flag": (I, [P]),
            "orc_init": (None, [P, ctypes.c_char_p, ctypes.c_ulonglong]),
            "orc_splitmix64": (ctypes.c_ulonglong, [ctypes.POINTER(ctypes.c_ulonglong)]),
            "orc_flux_limiter": (None, [D, I, DP]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One simulation on the host: state u, swap v, geometry and parameters."""

    def __init__(self, module, n1, n2, fp64=True, params=None, geometry=None,
                 config=True, seed=0):
        self.module = module
        self.n1, self.n2 = n1, n2
        self.nv = NVAR[module]
        self.h = lib().orc_create(MODULES[module], n1, n2, 1 if fp64 else 0)
        if not self.h:
            raise ValueError("bad oracle shape")
        for k, v in (params or {}).items():
            self.set_param(k, v)
        if geometry is not None:
            self.set_geometry(**geometry)
        if config:
            self.config()
        lib().orc_seed(self.h, seed)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @property
    def shape(self):
        return (self.n1, self.n2, self.nv)

    def set_param(self, name, val):
        if lib().orc_set_param(self.h, name.encode(), float(val)) != 0:
            raise KeyError(name)

    def get_param(self, name):
        return lib().orc_get_param(self.h, name.encode())

    def set_geometry(self, l1=1.0, l2=1.0, p1=1, p2=1, c=0.5):
        lib().orc_set_geometry(self.h, l1, l2, int(p1), int(p2), c)

    def geometry(self):
        out = np.zeros(7)
        lib().orc_get_geometry(self.h, _dp(out))
        keys = ("l1", "l2", "p1", "p2", "c", "Delta1", "Delta2")
        return dict(zip(keys, out.tolist()))

    def config(self):
        lib().orc_config(self.h)

    def init(self, name, seed=0):
        """Initial condition `name` at the cell centres (the reference's
        init(), src/init.cpp:26-42), rand() replaced by SplitMix64(seed)."""
        lib().orc_init(self.h, name.encode(), int(seed))

    def rand(self):
        out = np.zeros(2)
        lib().orc_get_rand(self.h, _dp(out))
        return out

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.shape)
        lib().orc_upload(self.h, _dp(x))

    def download(self):
        x = np.empty(self.shape)
        lib().orc_download(self.h, _dp(x))
        return x

    def upload_v(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.shape)
        lib().orc_upload_v(self.h, _dp(x))

    def download_v(self):
        x = np.empty(self.shape)
        lib().orc_download_v(self.h, _dp(x))
        return x

    def download_ghosted(self):
        x = np.empty((self.n1 + 6, self.n2 + 6, self.nv))
        lib().orc_download_ghosted(self.h, _dp(x))
        return x

    def bcond(self):
        lib().orc_bcond(self.h)

    def kick(self, beta):
        lib().orc_kick(self.h, beta)

    def drift(self, dt):
        lib().orc_drift(self.h, dt)

    def euler(self, dt):
        lib().orc_euler(self.h, dt)

    def rhs(self):
        out = np.empty(self.shape)
        lib().orc_rhs(self.h, _dp(out))
        return out

    def deriv(self, op):
        """op in 'd1', 'd2', 'd11', 'd22', 'd12': the operator of every
        variable over the interior after a boundary fill."""
        code = {"d1": 0, "d2": 1, "d11": 2, "d22": 3, "d12": 4}[op]
        out = np.empty(self.shape)
        lib().orc_deriv(self.h, code, _dp(out))
        return out

    def step(self, t, dt):
        return lib().orc_step(self.h, t, dt)

    def run(self, t, dt, n):
        return lib().orc_run(self.h, t, dt, n)

    def getdt(self):
        return lib().orc_getdt(self.h)

    def getdt_cfl(self):
        return lib().orc_getdt_cfl(self.h)

    def nan_flag(self):
        return lib().orc_nan_flag(self.h)


def splitmix64(state, count):
    """The first `count` outputs of SplitMix64 started at `state` (the
    oracle's stand-in for libc rand())."""
    st = ctypes.c_ulonglong(state)
    return [lib().orc_splitmix64(ctypes.byref(st)) for _ in range(count)]


def flux_limiter(bigR, fp64=True):
    """(Lambda, Lambda') of the disk's radiative flux limiter at R, as the
    oracle's disk right-hand side evaluates them."""
    out = (ctypes.c_double * 2)()
    lib().orc_flux_limiter(float(bigR), 1 if fp64 else 0, out)
    return out[0], out[1]
