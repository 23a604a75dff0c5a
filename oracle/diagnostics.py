"""numpy restatement of the reference's IDL spectral diagnostics (FP64).

TEST INFRASTRUCTURE ONLY.  Only tests/ and bench.py's checks may import this
module; it is what fg2_spectrum / fg2_vorticity are compared with, never part of the
product.

    getk(n1, n2, zeronyquist)   tools/getk.pro:18-33
    oned(k, E)                  tools/oned.pro:18-44
    spectrum(u1, u2)            tools/spec.pro:25-31 (E) and :58-63 (Z)
    vorticity(u1, u2)           tools/vis.pro:57-63

Arrays are (n1, n2): axis 0 is the reference's first dimension (tools/
load.pro:46-52 transposes the .raw data to d[i1, i2, var]).  IDL's forward
fft divides by N = n1 n2 and its inverse does not.  Wavenumbers and |k| are
FP32, as IDL's findgen and sqrt of a float array; the transforms and the
shell sums are FP64 here (IDL's fft of a float array is single precision and
total() of a float array accumulates in single precision).

Pinning (tests/test_oracle_diagnostics.py): a direct O(N^2) DFT on small
grids, closed forms of single Fourier modes (E = 1/4 and Z = b^2/4 in the
shell of |(a, b)|, vorticity b sin(theta)), Parseval for band-limited fields,
and the bin edges of oned against brute force.
"""
import numpy as np


def getk(n1, n2, zeronyquist=False):
    """Integer wavenumbers [0, 1, ..., -2, -1] on each axis, as (n1, n2)
    FP32 arrays k1, k2 (tools/getk.pro:24-31)."""
    def axis(n):
        k = np.concatenate([np.arange(n - n // 2), -(np.arange(n // 2)[::-1] + 1)])
        k = k.astype(np.float32)
        if zeronyquist and n % 2 == 0:
            k[n // 2] = 0.0
        return k
    k1, k2 = axis(n1), axis(n2)
    return np.repeat(k1[:, None], n2, axis=1), np.repeat(k2[None, :], n1, axis=0)


def oned(k, E):
    """Shell sums of E over |k| bins (tools/oned.pro:18-44): dict with E,
    k (bin centres 1..n), n (wavevectors per bin), b (edges) and c (kG)."""
    kG = min(k.shape) // 2 + 0.99
    n = int(np.ceil(kG))
    kb = (np.arange(n + 1) + 0.5).astype(np.float32)
    Ec = np.zeros(n)
    nb = np.zeros(n, np.int64)
    for i in range(n):
        if i == 0:
            j = (0.0 < k) & (k < kb[i + 1])
        elif i == n - 1:
            j = (kb[i] <= k) & (k <= kb[i + 1])
        else:
            j = (kb[i] <= k) & (k < kb[i + 1])
        nb[i] = int(np.count_nonzero(j))
        Ec[i] = float(np.sum(E[j], dtype=np.float64)) if nb[i] else 0.0
    return {"E": Ec, "k": np.arange(1, n + 1, dtype=np.float64), "n": nb, "b": kb, "c": kG}


def _hat(u):
    u = np.asarray(u, np.float64)
    return np.fft.fft2(u) / u.size           # IDL forward fft: divided by N


def spectrum(u1, u2):
    """E and Z shell spectra of a velocity field (tools/spec.pro)."""
    a, b = _hat(u1), _hat(u2)
    E = (np.abs(a) ** 2 + np.abs(b) ** 2) / 2
    k1, k2 = getk(*a.shape)
    k = np.sqrt(k1 * k1 + k2 * k2)           # FP32, as IDL
    s = oned(k, E)
    z1, z2 = getk(*a.shape, zeronyquist=True)
    w = 1j * z1.astype(np.float64) * b - 1j * z2.astype(np.float64) * a
    Z = np.abs(w) ** 2 / 2
    return {"k": s["k"], "E": s["E"], "Z": oned(k, Z)["E"], "counts": s["n"]}


def vorticity(u1, u2):
    """Real part of the inverse fft of omega^ (tools/vis.pro:57-63)."""
    a, b = _hat(u1), _hat(u2)
    z1, z2 = getk(*a.shape, zeronyquist=True)
    w = 1j * z1.astype(np.float64) * b - 1j * z2.astype(np.float64) * a
    return np.real(np.fft.ifft2(w) * w.size)  # IDL's inverse is not normalised
