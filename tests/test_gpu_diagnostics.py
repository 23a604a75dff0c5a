"""GPU checks of the spectral diagnostics (fg2_spectrum, fg2_vorticity)
against the numpy restatement of tools/spec.pro and tools/vis.pro.

Bars: shell counts exact; E and Z per shell within 1e-5 of the FP64 oracle
relative to the shell's value plus 1e-6 of the spectrum's largest shell (the
device FFT is FP32, whose error is relative to the whole field); the
vorticity field within 1e-5 relative L2.  Single Fourier modes against
their closed forms."""
import sys

import numpy as np
import pytest

from helpers import state

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _velocity_state(fg2, n1, n2, seed):
    x = state("hydro", n1, n2, seed=seed)
    r = fg2.Run("hydro", n1, n2)
    r.upload(x)
    return r, x


def _close(got, ref):
    return np.abs(got - ref) <= 1e-5 * np.abs(ref) + 1e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("n1,n2", [(64, 48), (63, 81), (200, 136), (256, 256), (3, 7)])
def test_spectrum_vs_oracle(fg2, n1, n2):
    from oracle import diagnostics as D
    _say(f"case spectrum {n1}x{n2}")
    r, x = _velocity_state(fg2, n1, n2, seed=n1 + n2)
    got = r.spectrum()
    ref = D.spectrum(x[..., 1], x[..., 2])
    assert len(got["E"]) == len(ref["E"])
    np.testing.assert_array_equal(got["counts"], ref["counts"])
    for key in ("E", "Z"):
        ok = _close(got[key], ref[key])
        assert ok.all(), (key, np.argwhere(~ok)[:5].ravel(), got[key][~ok][:5], ref[key][~ok][:5])
    w = r.vorticity()
    wr = D.vorticity(x[..., 1], x[..., 2])
    err = np.linalg.norm(w - wr) / np.linalg.norm(wr)
    assert err < 1e-5, err
    # the state is not changed
    np.testing.assert_array_equal(r.download(), x.astype(np.float32))


@pytest.mark.parametrize("a,b", [(3, 4), (0, 5), (11, 2)])
def test_single_mode(fg2, a, b):
    n1, n2 = 128, 96
    _say(f"case single mode ({a},{b})")
    i = np.arange(n1)[:, None]
    j = np.arange(n2)[None, :]
    th = 2 * np.pi * (a * i / n1 + b * j / n2)
    x = np.zeros((n1, n2, 4), np.float32)
    x[..., 1] = np.cos(th)
    r = fg2.Run("hydro", n1, n2)
    r.upload(x)
    s = r.spectrum()
    shell = int(np.floor(np.hypot(a, b) - 0.5))
    assert s["E"][shell] == pytest.approx(0.25, rel=1e-5)
    assert s["Z"][shell] == pytest.approx(b * b / 4, rel=1e-5)
    assert np.sum(s["E"]) == pytest.approx(0.25, rel=1e-5)
    np.testing.assert_allclose(r.vorticity(), b * np.sin(th), atol=2e-5 * max(b, 1))


def test_spectrum_full_size_kh(fg2):
    from oracle import diagnostics as D
    n = 3584
    _say(f"case spectrum KH {n}x{n}")
    r = fg2.Run("hydro", n, n)
    r.init("KH", 5)
    x = r.download()
    got = r.spectrum()
    ref = D.spectrum(x[..., 1], x[..., 2])
    np.testing.assert_array_equal(got["counts"], ref["counts"])
    for key in ("E", "Z"):
        ok = _close(got[key], ref[key])
        assert ok.all(), (key, np.argwhere(~ok)[:5].ravel(), got[key][~ok][:5], ref[key][~ok][:5])


def test_spectrum_arguments(fg2):
    r = fg2.Run("hydro", 32, 40)
    import ctypes
    n = ctypes.c_int(0)
    assert fg2.lib.fg2_spectrum(r.h, None, None, None, ctypes.byref(n)) == 0 and n.value == 17
    E = np.zeros(5)
    n = ctypes.c_int(5)
    code = fg2.lib.fg2_spectrum(r.h, E.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, None,
                                ctypes.byref(n))
    assert code == fg2.ERR_ARG and n.value == 17
    d = fg2.Run("disk", 64, 64)
    d.init("Hawley")
    assert np.all(d.spectrum()["E"] == 0.0)            # the torus starts at rest
