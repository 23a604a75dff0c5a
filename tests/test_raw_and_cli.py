"""CPU checks of the .raw snapshot format (src/io.cpp:24-76) and of the
command-line driver's argument handling (src/main.cpp:29-60, src/usage.cpp)
that needs no device."""
import os
import struct
import subprocess

import numpy as np
import pytest

from fg2_amd import raw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FG2RUN = os.path.join(ROOT, "fg2_amd", "fg2run")


def test_raw_layout_matches_the_reference_fwrite_order(tmp_path):
    """int32[6] {n1, n2, NVAR, sizeof(R), p1, p2}, E time, R[3] {l1, l2, c},
    then the interior: a file packed field by field with struct reads back."""
    n1, n2, nv = 3, 5, 4
    data = np.arange(n1 * n2 * nv, dtype=np.float32).reshape(n1, n2, nv) * 0.25 - 3
    blob = struct.pack("<6i", n1, n2, nv, 4, 1, 0) + struct.pack("<d", 0.125) \
        + struct.pack("<3f", 1.0, 2.0, 0.5) + data.tobytes()
    p = tmp_path / "0007.raw"
    p.write_bytes(blob)
    s = raw.read(p)
    assert s.shape == (n1, n2, nv) and s.data.dtype == np.float32
    assert np.array_equal(s.data, data)
    assert (s.time, s.l1, s.l2, s.c, s.p1, s.p2) == (0.125, 1.0, 2.0, 0.5, 1, 0)
    assert raw.HEADER_BYTES == 44
    q = tmp_path / "copy.raw"
    raw.write(q, s)
    assert q.read_bytes() == blob


def test_raw_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(3)
    s = raw.Snapshot(rng.standard_normal((7, 9, 3)).astype(np.float32), 1.5, 6.25, 6.5, 0.3, 0, 1)
    p = tmp_path / "a.raw"
    raw.write(p, s)
    assert os.path.getsize(p) == 44 + 7 * 9 * 3 * 4
    t = raw.read(p)
    assert np.array_equal(t.data, s.data) and t.time == 1.5 and (t.p1, t.p2) == (0, 1)
    assert np.float32(t.l1) == np.float32(6.25) and np.float32(t.c) == np.float32(0.3)
    p.write_bytes(p.read_bytes()[:-4])
    with pytest.raises(ValueError, match="short data"):
        raw.read(p)


needs_driver = pytest.mark.skipif(not os.path.exists(FG2RUN), reason="fg2run not built")


@needs_driver
def test_cli_help_and_illegal_options():
    r = subprocess.run([FG2RUN, "hydro", "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "Usage: fg2run MODULE" in r.stdout
    for bad in (["hydro", "-x"], ["hydro", "-n"], ["hydro", "-c", "-b"]):
        r = subprocess.run([FG2RUN, *bad], capture_output=True, text=True)
        assert r.returncode == 255 and "illegal option" in r.stderr, bad
    r = subprocess.run([FG2RUN, "mhd"], capture_output=True, text=True)
    assert r.returncode == 255 and "unknown module" in r.stderr
