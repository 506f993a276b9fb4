"""NAL packaging of the slices (host/h264_host.c): emulation prevention
against the reference's own bs_rbsp_to_ebsp (src/bitstream.c:232-250,
compiled into oracle/_ref/libref_cavlc.so), its defining properties, and the
slice NAL layout of src/encoder_h264.c:3440-3470."""
import ctypes as C
import os

import numpy as np
import pytest

import orc
from conftest import load_pkg


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


def unescape(ebsp):
    out, zeros, i = bytearray(), 0, 0
    while i < len(ebsp):
        b = ebsp[i]
        if zeros >= 2 and b == 3:
            zeros, i = 0, i + 1
            continue
        out.append(b)
        zeros = zeros + 1 if b == 0 else 0
        i += 1
    return bytes(out)


def zero_rich(rs, n):
    a = rs.integers(0, 256, n).astype(np.uint8)
    a[rs.random(n) < 0.6] = 0
    a[rs.random(n) < 0.1] = rs.integers(1, 4)
    return a.tobytes()


def test_ebsp_properties(pkg):
    rs = np.random.default_rng(3)
    for n in list(range(0, 12)) + [100, 1000, 5000]:
        for _ in range(20):
            r = zero_rich(rs, n)
            e = pkg.rbsp_to_ebsp(r)
            assert unescape(e) == r
            # 7.4.1: 00 00 00, 00 00 01 and 00 00 02 never occur; 00 00 03 is the escape
            for i in range(len(e) - 2):
                assert not (e[i] == 0 and e[i + 1] == 0 and e[i + 2] <= 2)
    assert pkg.rbsp_to_ebsp(b"\x00\x00\x00\x00\x01") == b"\x00\x00\x03\x00\x00\x03\x01"
    assert pkg.rbsp_to_ebsp(b"\x00\x00\x04") == b"\x00\x00\x04"


@pytest.mark.skipif(not os.path.exists(orc.REF_CAVLC), reason="oracle/_ref not built (reference tree absent)")
def test_ebsp_matches_reference(pkg):
    ref = C.CDLL(orc.REF_CAVLC)
    ref.bs_rbsp_to_ebsp.restype = C.c_size_t
    ref.bs_rbsp_to_ebsp.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
    rs = np.random.default_rng(4)
    for n in [1, 2, 3, 7, 64, 777, 4096]:
        for _ in range(30):
            r = np.frombuffer(zero_rich(rs, n), np.uint8)
            d = np.zeros(2 * n + 4, np.uint8)
            k = ref.bs_rbsp_to_ebsp(d.ctypes.data_as(C.c_void_p), d.size, r.ctypes.data_as(C.c_void_p), n)
            assert pkg.rbsp_to_ebsp(r.tobytes()) == d[:k].tobytes()


def test_frame_nals_layout(pkg):
    rs = np.random.default_rng(5)
    from test_oracle_cavlc import synth_frame
    for idr in (1, 0):
        lv, dc, modes, mvs = synth_frame(rs, 6, 4, idr)
        slices = orc.h264_cavlc_slices(lv, dc, modes, mvs, orc.slice_cfg(6, 4, 27, idr, 3))
        out = pkg.h264_frame_nals(slices, idr)
        parts = out.split(b"\x00\x00\x00\x01")
        assert parts[0] == b"" and len(parts) == 4
        for p, s in zip(parts[1:], slices):
            assert p[0] == (0x65 if idr else 0x41)
            assert unescape(p[1:]) == s
