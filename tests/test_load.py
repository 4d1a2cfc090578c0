"""The N-GenIC / Gadget-1 reader (tg_ngenic_load; src/load.cpp:13-96) on files written here with
the format's own framing (Fortran record markers around the 256-byte header and each block).
Expected values are worked out from the format and the reference's rules, not by the reader:
N = 8 particles have mean spacing 1/2, so the shift is 1/4; a coordinate past 1 after the shift
wraps once (0.9 -> 1.15 -> 0.15 in float); velocities scale by (1 + z)^(-3/2) = 1/8 at z = 3."""
import struct

import numpy as np
import pytest

from tigergrav_amd import ics


def write_ngenic(path, x, v, redshift, mass=1e-3, marker_fix=True):
    n = x.shape[0]
    hdr = struct.pack("<6I6d2d2i6I2i4d3i84s", 0, n, 0, 0, 0, 0, 0.0, mass, 0, 0, 0, 0, 1.0 / (1 + redshift), redshift,
                      0, 0, 0, n, 0, 0, 0, 0, 0, 1, 1.0, 0.3, 0.7, 0.7, 0, 0, 0, b"")
    assert len(hdr) == 256

    def block(b):
        return struct.pack("<i", len(b)) + b + struct.pack("<i", len(b))

    with open(path, "wb") as f:
        f.write(block(hdr))
        f.write(block(np.ascontiguousarray(x, "<f4").tobytes()))
        f.write(block(np.ascontiguousarray(v, "<f4").tobytes()))


def test_header_positions_velocities(tmp_path):
    x = np.array([[0.25, 0.0, 0.5], [0.9, 0.75, 0.1], [0.7499, 0.0625, 0.999]] + [[0.1, 0.2, 0.3]] * 5, np.float32)
    v = np.array([[8.0, -16.0, 0.5]] * 8, np.float32)
    p = tmp_path / "ics.dat"
    write_ngenic(p, x, v, redshift=3.0)
    xl, vl, h = ics.load_ngenic(p)
    assert h["npart"][1] == 8 and h["redshift"] == 3.0 and h["Omega0"] == 0.3
    # shift 0.5 * 8^(-1/3) = 0.25, applied in double and stored back as float
    want = np.float32(x.astype(np.float64) + 0.25)
    want = np.where(want > 1.0, np.float32(want.astype(np.float64) - 1.0), want)
    assert want[1, 0] == np.float32(np.float64(np.float32(1.15)) - 1.0)  # 0.9 wrapped once
    assert np.array_equal(xl, ics.double_to_pos(want.astype(np.float64)))
    assert xl[0, 0] == 0  # 0.25 + 0.25 = 1/2: the box centre is position 0
    np.testing.assert_array_equal(vl, np.float32(v.astype(np.float64) / 8.0))


def test_size_query_and_errors(tmp_path):
    import ctypes as C

    from tigergrav_amd._lib import TG_EINVAL, lib

    p = tmp_path / "ics.dat"
    write_ngenic(p, np.zeros((27, 3), np.float32), np.zeros((27, 3), np.float32), 0.0)
    L = lib()
    L.tg_ngenic_load.argtypes = [C.c_char_p, C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
    n = C.c_int64()
    assert L.tg_ngenic_load(str(p).encode(), 0, None, None, C.byref(n)) == 0 and n.value == 27
    x = np.zeros((26, 3), np.int32)
    v = np.zeros((26, 3), np.float32)
    assert L.tg_ngenic_load(str(p).encode(), 26, x.ctypes.data, v.ctypes.data, C.byref(n)) == TG_EINVAL
    with pytest.raises(Exception, match="unable to open"):
        ics.load_ngenic(tmp_path / "missing.dat")
    q = tmp_path / "short.dat"
    q.write_bytes(p.read_bytes()[:400])
    with pytest.raises(Exception, match="truncated"):
        ics.load_ngenic(q)
