"""Host reader and writer of fg2's .raw snapshots (src/io.cpp:24-76).

Layout, native little-endian:
    int32[6]   n1, n2, NVAR, sizeof(R) (4 for FP32), p1, p2
    float64    time                      (E = double when R = float)
    R[3]       l1, l2, c
    R[n1*n2*NVAR]   the interior, row-major over (i, j), AoS over variables

The state is the only thing a run needs to continue: the swap array v is
dead at the start of a step (the first RK3 substage has beta = 0).
"""
import numpy as np

HEADER_BYTES = 6 * 4 + 8 + 3 * 4


class Snapshot:
    """A .raw file: `data` (n1, n2, NVAR) plus time and geometry."""

    def __init__(self, data, time=0.0, l1=1.0, l2=1.0, c=0.5, p1=1, p2=1):
        self.data = np.asarray(data)
        if self.data.ndim != 3:
            raise ValueError("data must be (n1, n2, NVAR)")
        self.time = float(time)
        self.l1, self.l2, self.c = l1, l2, c
        self.p1, self.p2 = int(p1), int(p2)

    @property
    def shape(self):
        return self.data.shape


def write(path, snap):
    """Write `snap` as .raw (dump(), src/io.cpp:56-76); R follows the data
    dtype (float32 or float64)."""
    data = np.ascontiguousarray(snap.data)
    if data.dtype not in (np.float32, np.float64):
        data = data.astype(np.float32)
    n1, n2, nv = data.shape
    word = data.dtype.itemsize
    etype = np.float64 if word == 4 else np.longdouble
    with open(path, "wb") as f:
        f.write(np.array([n1, n2, nv, word, snap.p1, snap.p2], "<i4").tobytes())
        f.write(np.array([snap.time], etype).tobytes())
        f.write(np.array([snap.l1, snap.l2, snap.c], data.dtype).tobytes())
        f.write(data.tobytes())


def read(path):
    """Read a .raw file (load(), src/io.cpp:24-54) into a Snapshot."""
    with open(path, "rb") as f:
        size = np.frombuffer(f.read(24), "<i4")
        if size.size != 6:
            raise ValueError(f"{path}: short header")
        n1, n2, nv, word, p1, p2 = (int(x) for x in size)
        if word == 4:
            rtype, etype = np.float32, np.float64
        elif word == 8:
            rtype, etype = np.float64, np.longdouble
        else:
            raise ValueError(f"{path}: word size {word}")
        time = float(np.frombuffer(f.read(np.dtype(etype).itemsize), etype)[0])
        info = np.frombuffer(f.read(3 * word), rtype)
        count = n1 * n2 * nv
        data = np.frombuffer(f.read(count * word), rtype)
        if data.size != count:
            raise ValueError(f"{path}: short data ({data.size} of {count} values)")
    return Snapshot(data.reshape(n1, n2, nv).copy(), time, float(info[0]), float(info[1]),
                    float(info[2]), p1, p2)
