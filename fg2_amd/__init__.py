"""fg2_amd -- MI355X (gfx950) implementation of fg2's time-step path.

Python host mirror of the reference's driver API (src/fg2.h:70-84) over the
C ABI of libfg2_amd.so (include/fg2_amd.h).  The reference keeps one
simulation in `namespace global`; here a `Run` holds it:

    setup(n1, n2)           -> Run(module, n1, n2)        src/setup.cpp:32
    config()                -> Run.config()               src/config.cpp:24
    para("name=val")        -> Run.para("name=val")       src/para.cpp:25
    getdt()                 -> Run.getdt()                src/getdt.cpp:22
    step(t, dt)             -> Run.step(t, dt)            src/step.cpp:22
    solve(t, t_final, ...)  -> Run.solve(t, t_final)      src/solve.cpp:25
    Euler/bcond/kick/drift  -> Run.euler/bcond/kick/drift src/fg2.h:81-84
    init(name)              -> Run.init(name, seed)       src/init.cpp:26
    load(name) / dump(...)  -> Run.load(path) / Run.dump  src/io.cpp:24,56

`fg2_amd.raw` reads and writes .raw snapshots on the host, and the
command-line driver `fg2_amd/fg2run` mirrors the reference's main().

All arithmetic runs in the HIP library; this module only moves host arrays
and calls the C ABI.  The library is opened on first use, which raises if
it is not built: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfg2_amd.so")



class _Modules(dict):
    """The library's physics modules (fg2_module_count, fg2_module), read on
    first use: name -> value(nvar, fields)."""

    def __init__(self, value):
        super().__init__()
        self._value = value

    def _fill(self):
        if not dict.__len__(self):
            for k in range(lib.fg2_module_count()):
                n, nv, f = ctypes.c_char_p(), ctypes.c_int(), ctypes.c_char_p()
                _check(lib.fg2_module(k, ctypes.byref(n), ctypes.byref(nv), ctypes.byref(f)))
                dict.__setitem__(self, n.value.decode(),
                                 self._value(nv.value, tuple(f.value.decode().split(","))))

    def __getitem__(self, k):
        self._fill()
        return dict.__getitem__(self, k)

    def __contains__(self, k):
        self._fill()
        return dict.__contains__(self, k)

    def __iter__(self):
        self._fill()
        return dict.__iter__(self)

    def __len__(self):
        self._fill()
        return dict.__len__(self)

    def keys(self):
        self._fill()
        return dict.keys(self)

    def items(self):
        self._fill()
        return dict.items(self)


NVAR = _Modules(lambda nv, fields: nv)          # module name -> NVAR
FIELDS = _Modules(lambda nv, fields: fields)    # module name -> state variable names

OK, ERR_NAN, ERR_HIP, ERR_SHAPE, ERR_ARG, ERR_NOMEM = range(6)
PATH_FUSED, PATH_PLAIN, PATH_SUBSTAGE = 0, 1, 2
KERNEL_AUTO, KERNEL_ONE_PASS, KERNEL_SUBSTAGE = 0, 1, 2
DT_CLOSED, DT_CFL = 0, 1
SERVE_BATCHED, SERVE_PER_RUN, SERVE_PER_RUN_PLAIN = 0, 1, 2


class Fg2Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fg2 error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m fg2_amd.build` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P, I, D, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float
    PP = ctypes.POINTER(ctypes.c_void_p)
    DP = ctypes.POINTER(ctypes.c_double)
    FP = ctypes.POINTER(ctypes.c_float)
    IP = ctypes.POINTER(ctypes.c_int)
    LP = ctypes.POINTER(ctypes.c_long)
    SP = ctypes.POINTER(ctypes.c_size_t)
    sig = {
        "fg2_last_error": (ctypes.c_char_p, []),
        "fg2_module_count": (I, []),
        "fg2_module": (I, [I, ctypes.POINTER(ctypes.c_char_p), IP,
                           ctypes.POINTER(ctypes.c_char_p)]),
        "fg2_setup": (I, [PP, ctypes.c_char_p, I, I, SP]),
        "fg2_free": (None, [P]),
        "fg2_set_geometry": (I, [P, D, D, I, I, D]),
        "fg2_get_geometry": (I, [P, DP]),
        "fg2_config": (I, [P]),
        "fg2_para": (ctypes.c_char_p, [P, ctypes.c_char_p]),
        "fg2_get_param": (I, [P, ctypes.c_char_p, DP]),
        "fg2_seed": (I, [P, ctypes.c_ulonglong]),
        "fg2_layout": (I, [P, IP, LP]),
        "fg2_u": (P, [P]),
        "fg2_v": (P, [P]),
        "fg2_stream": (P, [P]),
        "fg2_upload": (I, [P, FP]),
        "fg2_download": (I, [P, FP]),
        "fg2_upload_v": (I, [P, FP]),
        "fg2_download_v": (I, [P, FP]),
        "fg2_euler": (I, [P, P, F]),
        "fg2_bcond": (I, [P, P, I, I]),
        "fg2_kick": (I, [P, P, P, F, F]),
        "fg2_drift": (I, [P, P, P, F]),
        "fg2_getdt": (D, [P]),
        "fg2_getdt_cfl": (I, [P, DP]),
        "fg2_set_path": (I, [P, I]),
        "fg2_set_dt_mode": (I, [P, I]),
        "fg2_set_tile": (I, [P, I, I]),
        "fg2_get_tile": (I, [P, IP, IP]),
        "fg2_set_kernel": (I, [P, I]),
        "fg2_get_kernel": (I, [P, IP]),
        "fg2_autotune": (I, [P, IP, IP]),
        "fg2_step": (I, [P, D, D]),
        "fg2_step_async": (I, [P, D, D]),
        "fg2_substage": (I, [P, I, D]),
        "fg2_check": (I, [P]),
        "fg2_solve": (I, [P, DP, D, IP]),
        "fg2_slab": (I, [P, I, I]),
        "fg2_slab_halo": (I, [P, PP, PP, PP, PP, SP]),
        "fg2_slab_exchange": (I, [PP, I]),
        "fg2_slab_step": (I, [P, D]),
        "fg2_sums": (I, [P, DP]),
        "fg2_perf_model": (I, [P, DP, DP, DP]),
        "fg2_init": (I, [P, ctypes.c_char_p, ctypes.c_ulonglong]),
        "fg2_spectrum": (I, [P, DP, DP, ctypes.POINTER(ctypes.c_ulonglong), IP]),
        "fg2_vorticity": (I, [P, FP]),
        "fg2_dump": (I, [P, ctypes.c_char_p, D]),
        "fg2_load": (I, [P, ctypes.c_char_p, DP]),
        "fg2_batch_create": (I, [PP, PP, I, I, I]),
        "fg2_batch_free": (None, [P]),
        "fg2_batch_step_async": (I, [P, DP]),
        "fg2_batch_step": (I, [P, DP]),
        "fg2_batch_graph_run": (I, [P, DP, I, I]),
        "fg2_batch_stream": (P, [P]),
        "fg2_batch_autotune": (I, [P, DP, IP, IP]),
        "fg2_batch_step_static": (I, [PP, I, DP, I, I]),
        "fg2_batch_solve": (I, [P, DP, DP, I, IP]),
        "fg2_server_create": (I, [PP, ctypes.c_size_t, I, I]),
        "fg2_server_free": (None, [P]),
        "fg2_server_submit": (I, [P, ctypes.c_char_p, I, I, I, FP, D,
                                  ctypes.POINTER(ctypes.c_longlong)]),
        "fg2_server_submit_run": (I, [P, P, I, FP, D, ctypes.POINTER(ctypes.c_longlong)]),
        "fg2_server_tick": (I, [P, D, IP, IP, IP]),
        "fg2_server_result": (I, [P, ctypes.c_longlong, FP, DP, IP]),
        "fg2_server_stats": (I, [P, ctypes.POINTER(ctypes.c_longlong), SP]),
        "fg2_server_set_mode": (I, [P, I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


class _LazyLib:
    """The loaded library, opened on first use (so that the build module of
    this package can be imported before the library exists).  Any use while
    the library is missing raises ImportError."""

    _L = None

    def __getattr__(self, name):
        if _LazyLib._L is None:
            _LazyLib._L = _load()
        return getattr(_LazyLib._L, name)


lib = _LazyLib()


def _check(code):
    if code != OK:
        raise Fg2Error(code, lib.fg2_last_error().decode())


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Run:
    """One simulation on the current HIP device (the reference's global state)."""

    def __init__(self, module, n1, n2, params=None, geometry=None, config=True,
                 seed=0, path=PATH_FUSED, tile=None):
        if module not in NVAR:
            raise ValueError(f"unknown module {module!r}")
        self.module, self.n1, self.n2 = module, int(n1), int(n2)
        self.nv = NVAR[module]
        h = ctypes.c_void_p()
        nbytes = ctypes.c_size_t()
        _check(lib.fg2_setup(ctypes.byref(h), module.encode(), self.n1, self.n2,
                             ctypes.byref(nbytes)))
        self.h = h
        self.bytes = nbytes.value
        if geometry is not None:
            self.set_geometry(**geometry)
        for k, v in (params or {}).items():
            self.para(f"{k}={v!r}")
        if config:
            self.config()
        _check(lib.fg2_seed(self.h, seed))
        self.set_path(path)
        if tile is not None:
            self.set_tile(*tile)

    def close(self):
        if getattr(self, "h", None):
            lib.fg2_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def shape(self):
        return (self.n1, self.n2, self.nv)

    # -- set-up -----------------------------------------------------------
    def set_geometry(self, l1=1.0, l2=1.0, p1=1, p2=1, c=0.5):
        _check(lib.fg2_set_geometry(self.h, l1, l2, int(p1), int(p2), c))

    def geometry(self):
        out = np.zeros(7)
        _check(lib.fg2_get_geometry(self.h, _dp(out)))
        return dict(zip(("l1", "l2", "p1", "p2", "c", "Delta1", "Delta2"), out.tolist()))

    def config(self):
        _check(lib.fg2_config(self.h))

    def para(self, assignment):
        r = lib.fg2_para(self.h, assignment.encode())
        if r is None:
            raise KeyError(lib.fg2_last_error().decode())
        return r.decode()

    def get_param(self, name):
        v = ctypes.c_double()
        _check(lib.fg2_get_param(self.h, name.encode(), ctypes.byref(v)))
        return v.value

    def layout(self):
        nv, s = ctypes.c_int(), ctypes.c_long()
        _check(lib.fg2_layout(self.h, ctypes.byref(nv), ctypes.byref(s)))
        return nv.value, s.value

    def u_ptr(self):
        return lib.fg2_u(self.h)

    def v_ptr(self):
        return lib.fg2_v(self.h)

    def set_path(self, path):
        _check(lib.fg2_set_path(self.h, path))

    def set_dt_mode(self, mode):
        _check(lib.fg2_set_dt_mode(self.h, mode))

    def set_tile(self, tile_h, waves=1):
        _check(lib.fg2_set_tile(self.h, int(tile_h), int(waves)))

    def set_kernel(self, kernel):
        """Fix the fused path's kernel (KERNEL_AUTO, KERNEL_ONE_PASS,
        KERNEL_SUBSTAGE) of a hydro or hydro-isot run."""
        _check(lib.fg2_set_kernel(self.h, int(kernel)))

    def kernel(self):
        """The kernel a step takes now: KERNEL_ONE_PASS, KERNEL_SUBSTAGE, or
        0 on the plain path."""
        k = ctypes.c_int()
        _check(lib.fg2_get_kernel(self.h, ctypes.byref(k)))
        return k.value

    def autotune(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib.fg2_autotune(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def tile(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib.fg2_get_tile(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # -- state ------------------------------------------------------------
    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.shape)
        _check(lib.fg2_upload(self.h, _fp(x)))

    def download(self):
        x = np.empty(self.shape, np.float32)
        _check(lib.fg2_download(self.h, _fp(x)))
        return x

    def upload_v(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.shape)
        _check(lib.fg2_upload_v(self.h, _fp(x)))

    def download_v(self):
        x = np.empty(self.shape, np.float32)
        _check(lib.fg2_download_v(self.h, _fp(x)))
        return x

    # -- initial conditions and snapshots ----------------------------------
    def init(self, name, seed=0):
        """Initial condition `name` evaluated on the device (init(),
        src/init.cpp:26-42); KH and Kf draw from SplitMix64(seed)."""
        _check(lib.fg2_init(self.h, name.encode(), int(seed)))

    def dump(self, path, time):
        """Write a .raw snapshot (dump(), src/io.cpp:56-76)."""
        _check(lib.fg2_dump(self.h, os.fsencode(path), float(time)))

    def load(self, path):
        """Read a .raw snapshot, taking l1, l2, p1, p2, c from it, and
        return its time (load(), src/io.cpp:24-54).  Call config() after."""
        t = ctypes.c_double(0.0)
        _check(lib.fg2_load(self.h, os.fsencode(path), ctypes.byref(t)))
        return t.value

    # -- spectral diagnostics (tools/spec.pro, tools/vis.pro) ---------------
    def spectrum(self):
        """Shell-integrated energy and vorticity spectra of the velocity:
        dict with k (shell centres 1..n), E, Z and counts per shell."""
        n = ctypes.c_int(0)
        _check(lib.fg2_spectrum(self.h, None, None, None, ctypes.byref(n)))
        E = np.zeros(n.value)
        Z = np.zeros(n.value)
        c = np.zeros(n.value, np.uint64)
        _check(lib.fg2_spectrum(self.h, _dp(E), _dp(Z),
                                c.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)),
                                ctypes.byref(n)))
        return {"k": np.arange(1, n.value + 1, dtype=np.float64), "E": E, "Z": Z, "counts": c}

    def vorticity(self):
        """Spectral vorticity field (n1, n2) of the velocity (tools/vis.pro)."""
        out = np.empty((self.n1, self.n2), np.float32)
        _check(lib.fg2_vorticity(self.h, _fp(out)))
        return out

    # -- device operations (src/fg2.h:81-84) ------------------------------
    def euler(self, dt):
        _check(lib.fg2_euler(self.h, self.u_ptr(), dt))

    def bcond(self, p1=None, p2=None):
        g = self.geometry()
        p1 = int(g["p1"]) if p1 is None else p1
        p2 = int(g["p2"]) if p2 is None else p2
        _check(lib.fg2_bcond(self.h, self.u_ptr(), p1, p2))

    def kick(self, beta, t=0.0):
        _check(lib.fg2_kick(self.h, self.v_ptr(), self.u_ptr(), t, beta))

    def drift(self, dt):
        _check(lib.fg2_drift(self.h, self.u_ptr(), self.v_ptr(), dt))

    # -- integration (src/fg2.h:77-79) ------------------------------------
    def getdt(self):
        return lib.fg2_getdt(self.h)

    def getdt_cfl(self):
        d = ctypes.c_double()
        _check(lib.fg2_getdt_cfl(self.h, ctypes.byref(d)))
        return d.value

    def step(self, t, dt):
        """One RK3 step; returns the status code (0, or ERR_NAN) like step()."""
        code = lib.fg2_step(self.h, t, dt)
        if code not in (OK, ERR_NAN):
            _check(code)
        return code

    def step_async(self, t, dt):
        _check(lib.fg2_step_async(self.h, t, dt))

    def substage(self, stage, dt):
        _check(lib.fg2_substage(self.h, stage, dt))

    def check(self):
        code = lib.fg2_check(self.h)
        if code not in (OK, ERR_NAN):
            _check(code)
        return code

    def run(self, t, dt, n):
        """n steps of fixed dt, one NaN check at the end."""
        for k in range(n):
            self.step_async(t + k * dt, dt)
        return self.check()

    def solve(self, t, t_final):
        tt = ctypes.c_double(t)
        m = ctypes.c_int()
        code = lib.fg2_solve(self.h, ctypes.byref(tt), t_final, ctypes.byref(m))
        if code not in (OK, ERR_NAN):
            _check(code)
        return code, tt.value, m.value

    def sums(self):
        out = np.zeros(4)
        _check(lib.fg2_sums(self.h, _dp(out)))
        return out

    def perf_model(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib.fg2_perf_model(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"flops": a.value, "bits": b.value, "alg_bytes_per_cell": c.value}

    def stream(self):
        return lib.fg2_stream(self.h)

    # -- slabs of a grid cut along dim 1 ----------------------------------
    def slab(self, row0, n1_global):
        """Make this run rows [row0, row0 + n1) of a periodic grid of
        n1_global rows (0: a whole grid again)."""
        _check(lib.fg2_slab(self.h, row0, n1_global))

    def slab_halo(self):
        """Device addresses of one halo exchange: recv_lo, recv_hi,
        send_first, send_last and the byte count of each range."""
        p = [ctypes.c_void_p() for _ in range(4)]
        n = ctypes.c_size_t()
        _check(lib.fg2_slab_halo(self.h, *[ctypes.byref(x) for x in p], ctypes.byref(n)))
        return {"recv_lo": p[0].value, "recv_hi": p[1].value, "send_first": p[2].value,
                "send_last": p[3].value, "bytes": n.value}

    def slab_step(self, dt):
        """One RK3 step of the slab from its current halos (asynchronous)."""
        _check(lib.fg2_slab_step(self.h, dt))


def slab_exchange(runs):
    """Fill the halos of consecutive slabs of one grid from their neighbours."""
    arr = (ctypes.c_void_p * len(runs))(*[r.h for r in runs])
    _check(lib.fg2_slab_exchange(arr, len(runs)))


class Batch:
    """Runs of one module advanced together by persistent launches."""

    def __init__(self, runs, waves=1, tile_h=0):
        self.runs = list(runs)
        arr = (ctypes.c_void_p * len(self.runs))(*[r.h.value for r in self.runs])
        h = ctypes.c_void_p()
        _check(lib.fg2_batch_create(ctypes.byref(h), arr, len(self.runs), waves, tile_h))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.fg2_batch_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, dts):
        d = np.ascontiguousarray(dts, dtype=np.float64)
        code = lib.fg2_batch_step(self.h, _dp(d))
        if code not in (OK, ERR_NAN):
            _check(code)
        return code

    def step_async(self, dts):
        d = np.ascontiguousarray(dts, dtype=np.float64)
        _check(lib.fg2_batch_step_async(self.h, _dp(d)))

    def graph_run(self, dts, steps, replays=1):
        d = np.ascontiguousarray(dts, dtype=np.float64)
        _check(lib.fg2_batch_graph_run(self.h, _dp(d), steps, replays))

    def solve(self, t, t_final, max_steps=1 << 30):
        """Advance run k from t[k] to t_final[k] with device-chosen steps
        (each run's dt mode), as fg2_solve advances one run; returns (status,
        times reached, steps taken)."""
        tt = np.ascontiguousarray(t, dtype=np.float64).copy()
        tf = np.ascontiguousarray(t_final, dtype=np.float64)
        st = np.zeros(len(self.runs), dtype=np.int32)
        code = lib.fg2_batch_solve(self.h, _dp(tt), _dp(tf), int(max_steps),
                                   st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        if code not in (OK, ERR_NAN):
            _check(code)
        return code, tt, st

    def stream(self):
        return lib.fg2_batch_stream(self.h)

    def autotune(self, dts):
        d = np.ascontiguousarray(dts, dtype=np.float64)
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib.fg2_batch_autotune(self.h, _dp(d), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def batch_step_static(runs, dts, waves=1, tile_h=0):
    """One batched step of `runs` through fg2_batch_step_static (no device
    allocation; at most FG2_STATIC_MAX_RUNS runs); returns OK or ERR_NAN."""
    arr = (ctypes.c_void_p * len(runs))(*[r.h.value for r in runs])
    d = np.ascontiguousarray(dts, dtype=np.float64)
    code = lib.fg2_batch_step_static(arr, len(runs), _dp(d), int(waves), int(tile_h))
    if code not in (OK, ERR_NAN):
        _check(code)
    return code


STATIC_MAX_RUNS = 512


class Server:
    """Continuous-batching request loop over the C ABI (fg2_server_*)."""

    def __init__(self, hbm_budget, waves=0, tile_h=0):
        h = ctypes.c_void_p()
        _check(lib.fg2_server_create(ctypes.byref(h), int(hbm_budget), waves, tile_h))
        self.h = h
        self.shapes = {}

    def close(self):
        if getattr(self, "h", None):
            lib.fg2_server_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, module, init, steps, now):
        x = np.ascontiguousarray(init, dtype=np.float32)
        n1, n2 = x.shape[0], x.shape[1]
        i = ctypes.c_longlong()
        _check(lib.fg2_server_submit(self.h, module.encode(), n1, n2, int(steps), _fp(x),
                                     float(now), ctypes.byref(i)))
        self.shapes[i.value] = x.shape
        return i.value

    def submit_run(self, like, init, steps, now):
        """A request shaped like the configured Run `like` (its module, grid,
        geometry, parameters, dt mode and forcing seed)."""
        x = np.ascontiguousarray(init, dtype=np.float32)
        i = ctypes.c_longlong()
        _check(lib.fg2_server_submit_run(self.h, like.h, int(steps), _fp(x), float(now),
                                         ctypes.byref(i)))
        self.shapes[i.value] = x.shape
        return i.value

    def tick(self, now):
        a, q, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib.fg2_server_tick(self.h, float(now), ctypes.byref(a), ctypes.byref(q),
                                   ctypes.byref(f)))
        return a.value, q.value, f.value

    def result(self, rid, fetch=True):
        """(status, (t_arrive, t_admit, t_done), state or None)."""
        t = np.zeros(3)
        st = ctypes.c_int()
        out = np.empty(self.shapes[rid], np.float32) if fetch else None
        _check(lib.fg2_server_result(self.h, rid, _fp(out) if fetch else None, _dp(t),
                                     ctypes.byref(st)))
        return st.value, tuple(t.tolist()), (out if st.value >= 0 else None)

    def set_mode(self, mode):
        """SERVE_BATCHED (one batched launch per module per tick),
        SERVE_PER_RUN (each run's own step launch) or SERVE_PER_RUN_PLAIN
        (each run's plain kernels); the per-run modes are the baselines."""
        _check(lib.fg2_server_set_mode(self.h, int(mode)))

    def stats(self):
        c, b = ctypes.c_longlong(), ctypes.c_size_t()
        _check(lib.fg2_server_stats(self.h, ctypes.byref(c), ctypes.byref(b)))
        return c.value, b.value
