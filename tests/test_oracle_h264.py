"""Pins the CPU oracle (oracle/orc_h264.c) before it is trusted as the
checker of the HIP path.

1. Against the reference's own H.264 decoder functions, compiled from its
   sources into oracle/_ref/libref_h264.so (h264_mc.c, h264_pred.c,
   h264_recon.c, h264_dec_tables.c): quarter-pel / eighth-pel MC at every
   fractional position including picture edges, Intra16x16 and chroma DC
   prediction, and a full decode==recon check -- the oracle's levels and
   modes are decoded with the reference's dequant / IDCT / DC transforms and
   must rebuild exactly the oracle's reconstruction, I and P frames.
2. Closed forms: the forward core transform equals C X C^T; the MF / V
   tables satisfy MF*V*16 ~ 2^21 * {1, 0.64, 0.8}; inverse of a DC-only
   block; brute-force full search on tiny inputs.
3. The reference's regression fixture for the Intra16x16 DC transpose
   (tests/test_encode.c:60-96 of the reference).
4. Deblocking: tables equal the reference decoder's, hand-computed lines for
   each filter branch, and picture-level invariants.
"""
import ctypes as C
import os

import numpy as np
import pytest

import orc

REF = orc.REF_H264
need_ref = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference tree absent)")

_ref = None


def ref():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF)
        P, I = C.c_void_p, C.c_int
        lib.h264d_mc_fetch_luma.restype = P
        lib.h264d_mc_fetch_luma.argtypes = [P, C.POINTER(I), P, I, I, I, I, I, I, I]
        lib.h264d_mc_fetch_chroma.restype = P
        lib.h264d_mc_fetch_chroma.argtypes = [P, C.POINTER(I), P, I, I, I, I, I, I, I]
        lib.h264d_mc_luma.argtypes = [P, I, P, I, I, I, I, I]
        lib.h264d_mc_chroma.argtypes = [P, I, P, I, I, I, I, I]
        lib.h264d_pred16x16.argtypes = [P, I, I, P, P, C.c_uint8, C.c_bool, C.c_bool]
        lib.h264d_pred_chroma.argtypes = [P, I, I, P, P, C.c_uint8, C.c_bool, C.c_bool]
        lib.h264d_dequant_build.argtypes = [P, I, P, P]
        lib.h264d_idct4_add.argtypes = [P, I, P, P, I, C.c_bool]
        lib.h264d_luma_dc_transform.argtypes = [P, I, C.c_int32]
        lib.h264d_chroma_dc_transform.argtypes = [P, I, C.c_int32]
        _ref = lib
    return _ref


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def rand_plane(rs, h, w):
    return rs.integers(0, 256, (h, w)).astype(np.uint8)


# ------------------------------------------------------------ interpolation

@need_ref
def test_luma_qpel_matches_reference_mc():
    lib = ref()
    rs = np.random.default_rng(1)
    plane = rand_plane(rs, 40, 48)
    buf = np.zeros(64 * 64, np.uint8)
    out = np.zeros((16, 16), np.uint8)
    for trial in range(60):
        w = int(rs.choice([4, 8, 16]))
        h = int(rs.choice([4, 8, 16]))
        x, y = int(rs.integers(-24, 60)), int(rs.integers(-24, 50))  # includes far outside
        for frac in range(16):
            fx, fy = frac & 3, frac >> 2
            ss = C.c_int()
            src = lib.h264d_mc_fetch_luma(p(buf), C.byref(ss), p(plane), plane.strides[0], 48, 40, x, y, w, h)
            lib.h264d_mc_luma(p(out), 16, src, ss.value, w, h, fx, fy)
            exp = out[:h, :w]
            got = np.array([[orc.luma_qpel(plane, (x + i) * 4 + fx, (y + j) * 4 + fy) for i in range(w)]
                            for j in range(h)], np.uint8)
            np.testing.assert_array_equal(got, exp, err_msg=f"x={x} y={y} frac={frac}")


@need_ref
def test_chroma_eighth_pel_matches_reference_mc():
    lib = ref()
    rs = np.random.default_rng(2)
    uv = rand_plane(rs, 20, 48)  # NV12 interleaved, 24 samples per component
    buf = np.zeros(32 * 32, np.uint8)
    out = np.zeros((8, 8), np.uint8)
    for comp in range(2):
        plane = np.ascontiguousarray(uv[:, comp::2])
        for trial in range(40):
            x, y = int(rs.integers(-12, 30)), int(rs.integers(-12, 26))
            for fx in range(8):
                for fy in range(8):
                    ss = C.c_int()
                    src = lib.h264d_mc_fetch_chroma(p(buf), C.byref(ss), p(plane), plane.strides[0], 24, 20, x, y,
                                                    8, 8)
                    lib.h264d_mc_chroma(p(out), 8, src, ss.value, 8, 8, fx, fy)
                    got = np.array([[orc.lib.orc_chroma_8th(p(uv), uv.strides[0], 24, 20, comp,
                                                            (x + i) * 8 + fx, (y + j) * 8 + fy)
                                     for i in range(8)] for j in range(8)], np.uint8)
                    np.testing.assert_array_equal(got, out)


# ------------------------------------------------------------ intra prediction

@need_ref
def test_pred16x16_and_chroma_dc_match_reference():
    lib = ref()
    rs = np.random.default_rng(3)
    out_r = np.zeros((16, 16), np.uint8)
    out_o = np.zeros(256, np.uint8)
    for trial in range(200):
        top = rs.integers(0, 256, 16).astype(np.int32)
        left = rs.integers(0, 256, 16).astype(np.int32)
        corner = int(rs.integers(0, 256))
        ta, la = bool(rs.integers(0, 2)), bool(rs.integers(0, 2))
        if trial < 4:
            ta, la = bool(trial & 1), bool(trial & 2)
        t8, l8 = top.astype(np.uint8), left.astype(np.uint8)
        for mode in range(4):
            if (mode == 0 and not ta) or (mode == 1 and not la) or (mode == 3 and not (ta and la)):
                continue
            lib.h264d_pred16x16(p(out_r), 16, mode, p(t8), p(l8), corner, ta, la)
            orc.lib.orc_pred16x16(mode, p(top), p(left), corner, int(ta), int(la), p(out_o))
            np.testing.assert_array_equal(out_o.reshape(16, 16), out_r, err_msg=f"mode {mode} ta={ta} la={la}")
        # chroma DC (mode 0 of h264d_pred_chroma)
        co = np.zeros(64, np.uint8)
        cr = np.zeros((8, 8), np.uint8)
        lib.h264d_pred_chroma(p(cr), 8, 0, p(t8[:8].copy()), p(l8[:8].copy()), corner, ta, la)
        orc.lib.orc_pred_chroma_dc(p(top[:8].copy()), p(left[:8].copy()), int(ta), int(la), p(co))
        np.testing.assert_array_equal(co.reshape(8, 8), cr)


# ------------------------------------------------------------ decode == recon

class Dequant(C.Structure):
    _fields_ = [("d4", (C.c_int32 * 16) * 6), ("d8", (C.c_int32 * 64) * 6), ("qp", C.c_int), ("valid", C.c_int)]


def _dequant(lib, qp):
    dq = Dequant()
    flat4 = np.full((6, 16), 16, np.uint8)
    flat8 = np.full((6, 64), 16, np.uint8)
    lib.h264d_dequant_build(C.byref(dq), qp, p(flat4), p(flat8))
    return np.array(dq.d4[0][:], np.int32)


def _decode_mb_residual(lib, dst_y, dst_uv, mbx, mby, levels, qp, intra, dc_levels_luma):
    """Add the residual of one MB to the prediction already in dst, the way
    the reference decoder does (h264_recon_mb.c)."""
    qpc = orc.lib.orc_chroma_qp(qp)
    dq, dqc = _dequant(lib, qp), _dequant(lib, qpc)
    lv = levels.reshape(24, 16)
    if intra:
        dcl = np.ascontiguousarray(dc_levels_luma, np.int16)
        lib.h264d_luma_dc_transform(p(dcl), qp, int(dq[0]))
    for b in range(16):
        blk = lv[b].astype(np.int16).copy()
        if intra:
            blk[0] = dcl[b]
        y0, x0 = mby * 16 + (b >> 2) * 4, mbx * 16 + (b & 3) * 4
        sub = np.ascontiguousarray(dst_y[y0:y0 + 4, x0:x0 + 4])
        lib.h264d_idct4_add(p(sub), 4, p(blk), p(dq), qp, intra)
        dst_y[y0:y0 + 4, x0:x0 + 4] = sub
    dcs = np.zeros(8, np.int32)
    return dcs


def _decode_chroma(lib, dst_uv, mbx, mby, levels, dc_in, qp):
    qpc = orc.lib.orc_chroma_qp(qp)
    dqc = _dequant(lib, qpc)
    lv = levels.reshape(24, 16)
    for comp in range(2):
        raw = np.array([int(dc_in[16 + comp * 4 + s]) for s in range(4)], np.int32)
        # transmitted chroma DC levels (host stage): 2x2 Hadamard + quantiser
        h = np.array([raw[0] + raw[1] + raw[2] + raw[3], raw[0] - raw[1] + raw[2] - raw[3],
                      raw[0] + raw[1] - raw[2] - raw[3], raw[0] - raw[1] - raw[2] + raw[3]])
        cl = np.array([orc.lib.orc_quant_dc_chroma(int(v), qpc) for v in h], np.int16)
        lib.h264d_chroma_dc_transform(p(cl), qpc, int(dqc[0]))
        plane = np.ascontiguousarray(dst_uv[:, comp::2])
        for s in range(4):
            blk = lv[16 + comp * 4 + s].astype(np.int16).copy()
            blk[0] = cl[s]
            y0, x0 = mby * 8 + (s >> 1) * 4, mbx * 8 + (s & 1) * 4
            sub = np.ascontiguousarray(plane[y0:y0 + 4, x0:x0 + 4])
            lib.h264d_idct4_add(p(sub), 4, p(blk), p(dqc), qpc, True)
            plane[y0:y0 + 4, x0:x0 + 4] = sub
        dst_uv[:, comp::2] = plane


def _frames(w, h, n, seed=4):
    rs = np.random.default_rng(seed)
    # smooth content + noise + one moving square, enough for real residuals
    yy, xx = np.mgrid[0:h, 0:w]
    out = []
    for f in range(n):
        y = 128 + 40 * np.sin((xx + 2.3 * f) / 7.0) * np.cos((yy - 1.7 * f) / 9.0)
        y = y + rs.normal(0, 3, (h, w))
        y[10 + f:30 + f, 20 + 2 * f:44 + 2 * f] = rs.integers(16, 236, (20, 24))
        y = np.clip(np.rint(y), 0, 255).astype(np.uint8)
        u = np.clip(128 + (y[0::2, 0::2].astype(int) - 128) // 4, 0, 255)
        v = np.clip(128 - (y[1::2, 1::2].astype(int) - 128) // 5, 0, 255)
        uv = np.empty((h // 2, w), np.uint8)
        uv[:, 0::2], uv[:, 1::2] = u, v
        out.append((y, uv))
    return out


@need_ref
@pytest.mark.parametrize("qp", [4, 22, 27, 32, 37, 51])
def test_decode_equals_recon_intra(qp):
    lib = ref()
    (cy, cuv), = _frames(64, 48, 1)
    for ns in (1, 2):
        e = orc.encode_i(cy, cuv, qp, ns)
        mbw, mbh = 4, 3
        dy = np.zeros_like(cy)
        duv = np.zeros_like(cuv)
        for m in range(mbw * mbh):
            mbx, mby = m % mbw, m // mbw
            sl = orc.lib.orc_slice_of(m, mbw * mbh, ns)
            ta = mby > 0 and orc.lib.orc_slice_of(m - mbw, mbw * mbh, ns) == sl
            la = mbx > 0 and orc.lib.orc_slice_of(m - 1, mbw * mbh, ns) == sl
            y0, x0 = mby * 16, mbx * 16
            top = dy[y0 - 1, x0:x0 + 16].copy() if mby > 0 else np.full(16, 128, np.uint8)
            left = dy[y0:y0 + 16, x0 - 1].copy() if mbx > 0 else np.full(16, 128, np.uint8)
            corner = int(dy[y0 - 1, x0 - 1]) if (mbx > 0 and mby > 0) else 128
            pred = np.zeros((16, 16), np.uint8)
            lib.h264d_pred16x16(p(pred), 16, int(e["modes"][m]), p(top), p(left), corner, ta, la)
            dy[y0:y0 + 16, x0:x0 + 16] = pred
            dc_in = e["dc"][m * 24:m * 24 + 16]
            nat = np.zeros(16, np.int32)
            orc.lib.orc_intra16_luma_dc_levels(p(np.ascontiguousarray(dc_in, np.int32)), qp, p(nat))
            _decode_mb_residual(lib, dy, duv, mbx, mby, e["levels"][m * 384:(m + 1) * 384], qp, True, nat)
            for comp in range(2):
                t = duv[mby * 8 - 1, mbx * 16 + comp:mbx * 16 + 16:2].copy() if ta else np.full(8, 128, np.uint8)
                l = duv[mby * 8:mby * 8 + 8, mbx * 16 - 2 + comp].copy() if la else np.full(8, 128, np.uint8)
                cp = np.zeros((8, 8), np.uint8)
                lib.h264d_pred_chroma(p(cp), 8, 0, p(np.ascontiguousarray(t)), p(np.ascontiguousarray(l)), 128, ta,
                                      la)
                duv[mby * 8:mby * 8 + 8, mbx * 16 + comp:mbx * 16 + 16:2] = cp
            _decode_chroma(lib, duv, mbx, mby, e["levels"][m * 384:(m + 1) * 384], e["dc"][m * 24:(m + 1) * 24], qp)
        np.testing.assert_array_equal(dy, e["y"], err_msg=f"luma qp={qp} ns={ns}")
        np.testing.assert_array_equal(duv, e["uv"], err_msg=f"chroma qp={qp} ns={ns}")


@need_ref
@pytest.mark.parametrize("qp", [0, 18, 27, 36, 51])
def test_decode_equals_recon_inter(qp):
    lib = ref()
    (ry, ruv), (cy, cuv) = _frames(64, 48, 2)
    rs = np.random.default_rng(qp)
    nmb = 12
    mvs = np.zeros(nmb, orc.MV)
    mvs["mvx"] = rs.integers(-70, 70, nmb)
    mvs["mvy"] = rs.integers(-70, 70, nmb)
    e = orc.encode_p(cy, cuv, ry, ruv, qp, mvs)
    dy = np.zeros_like(cy)
    duv = np.zeros_like(cuv)
    buf = np.zeros(64 * 64, np.uint8)
    for m in range(nmb):
        mbx, mby = m % 4, m // 4
        mx, my = int(mvs["mvx"][m]), int(mvs["mvy"][m])
        ss = C.c_int()
        src = lib.h264d_mc_fetch_luma(p(buf), C.byref(ss), p(ry), 64, 64, 48, mbx * 16 + (mx >> 2),
                                      mby * 16 + (my >> 2), 16, 16)
        pred = np.zeros((16, 16), np.uint8)
        lib.h264d_mc_luma(p(pred), 16, src, ss.value, 16, 16, mx & 3, my & 3)
        dy[mby * 16:mby * 16 + 16, mbx * 16:mbx * 16 + 16] = pred
        for comp in range(2):
            plane = np.ascontiguousarray(ruv[:, comp::2])
            src = lib.h264d_mc_fetch_chroma(p(buf), C.byref(ss), p(plane), 32, 32, 24, mbx * 8 + (mx >> 3),
                                            mby * 8 + (my >> 3), 8, 8)
            cp = np.zeros((8, 8), np.uint8)
            lib.h264d_mc_chroma(p(cp), 8, src, ss.value, 8, 8, mx & 7, my & 7)
            duv[mby * 8:mby * 8 + 8, mbx * 16 + comp:mbx * 16 + 16:2] = cp
        _decode_mb_residual(lib, dy, duv, mbx, mby, e["levels"][m * 384:(m + 1) * 384], qp, False, None)
        _decode_chroma(lib, duv, mbx, mby, e["levels"][m * 384:(m + 1) * 384], e["dc"][m * 24:(m + 1) * 24], qp)
    np.testing.assert_array_equal(dy, e["y"])
    np.testing.assert_array_equal(duv, e["uv"])


# ------------------------------------------------------------ closed forms

def test_forward_transform_is_matrix_product():
    Cm = np.array([[1, 1, 1, 1], [2, 1, -1, -2], [1, -1, -1, 1], [1, -2, 2, -1]])
    rs = np.random.default_rng(5)
    for _ in range(200):
        x = rs.integers(-255, 256, (4, 4))
        np.testing.assert_array_equal(orc.fwd4x4(x).reshape(4, 4), Cm @ x @ Cm.T)


def test_inverse_of_dc_only_block():
    for v in range(-2000, 2001, 37):
        d = np.zeros(16, np.int32)
        d[0] = v
        assert (orc.inv4x4(d) == (v + 32) >> 6).all()


def test_mf_v_design_products():
    # MF * V * 16 is 2^21 times the squared norm ratio of the position class
    targets = {0: 1.0, 1: 0.64, 2: 0.8}
    pos_of_class = {0: 0, 1: 5, 2: 1}
    for qp in range(6):
        for cls, pos in pos_of_class.items():
            mf = None
            # recover MF from the quantiser: level for |c| = 2^15 at qp%6, per = 0 with f removed
            mf = orc.lib.orc_quant_ac(1 << 15, pos, qp, 0)  # (2^15*MF + 2^15/6) >> 15 = MF
            v16 = orc.lib.orc_dequant_ac(1 << 4, pos, qp)   # per = 0: (16 * 16V + 8) >> 4 = 16V
            assert abs(mf * v16 / 2 ** 21 - targets[cls]) < 2e-3, (qp, cls, mf, v16)


def test_ac_rounding_offset_is_a_third_intra_and_a_sixth_inter():
    """Known answers of quantize.comp:105-124: level k first appears at
    |c| = ceil((k 2^qbits - f) / MF) with f = 2^qbits / 3 in intra blocks and
    2^qbits / 6 in inter ones (qbits = 15 + qp/6).  MF is the standard's
    forward table; swapping the two offsets, or using another fraction, moves
    every threshold."""
    mf_table = [(13107, 5243, 8066), (11916, 4660, 7490), (10082, 4194, 6554),
                (9362, 3647, 5825), (8192, 3355, 5243), (7282, 2893, 4559)]
    pos_of_class = {0: 0, 1: 5, 2: 1}
    for qp in range(0, 52, 3):
        qbits = 15 + qp // 6
        for cls, pos in pos_of_class.items():
            mf = mf_table[qp % 6][cls]
            for intra, frac in ((1, 3), (0, 6)):
                f = (1 << qbits) // frac
                for k in (1, 2, 7):
                    c = -(-(k * (1 << qbits) - f) // mf)  # ceil
                    assert orc.lib.orc_quant_ac(c, pos, qp, intra) == k, (qp, pos, intra, k)
                    assert orc.lib.orc_quant_ac(c - 1, pos, qp, intra) == k - 1, (qp, pos, intra, k)
                    assert orc.lib.orc_quant_ac(-c, pos, qp, intra) == -k
    # one value where the two offsets disagree: qp 0, |c| = 2, MF 13107
    assert orc.lib.orc_quant_ac(2, 0, 0, 1) == 1 and orc.lib.orc_quant_ac(2, 0, 0, 0) == 0


def test_dequant_shift_doubles_every_six_qp():
    for qp in range(24, 46):
        for pos in (0, 1, 5):
            assert orc.lib.orc_dequant_ac(3, pos, qp + 6) == 2 * orc.lib.orc_dequant_ac(3, pos, qp)


def test_chroma_qp_table_matches_reference():
    lib = C.CDLL(REF) if os.path.exists(REF) else None
    if lib is None:
        pytest.skip("oracle/_ref not built")
    t = (C.c_uint8 * 52).in_dll(lib, "h264d_chroma_qp")
    assert [orc.lib.orc_chroma_qp(q) for q in range(52)] == list(t)


def test_intra16_dc_transpose_fixture():
    """Reference fixture tests/test_encode.c:60-96: levels must come out as
    the transpose of the raw Hadamard order, and the raw order must be
    asymmetric for the check to mean anything."""
    dc_flat = np.array([16, 28, 40, 52, 64, 76, 88, 100, 112, 124, 136, 148, 160, 172, 184, 196], np.int32)
    raw = np.zeros(16, np.int32)
    orc.lib.orc_luma_dc_hadamard_fwd(p(dc_flat), p(raw))
    pre = np.array([orc.lib.orc_quant_dc_luma(int(v), 26) for v in raw]).reshape(4, 4)
    assert (pre != pre.T).any()
    out = np.zeros(16, np.int32)
    orc.lib.orc_intra16_luma_dc_levels(p(dc_flat), 26, p(out))
    np.testing.assert_array_equal(out.reshape(4, 4), pre.T)


def test_host_dc_stage_matches_oracle(pkg):
    """The product's host DC stage (host/h264_host.c) against the oracle."""
    rs = np.random.default_rng(6)
    for qp in range(52):
        dc_in = rs.integers(-4000, 4000, 16).astype(np.int32)
        a = np.zeros(16, np.int32)
        b = np.zeros(16, np.int32)
        pre = np.zeros(16, np.int32)
        pkg.host.h264_intra16_luma_dc_transform(p(dc_in), qp, p(a), p(pre))
        orc.lib.orc_intra16_luma_dc_levels(p(dc_in), qp, p(b))
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(a.reshape(4, 4), pre.reshape(4, 4).T)
        c4 = rs.integers(-3000, 3000, 4).astype(np.int32)
        o = np.zeros(4, np.int32)
        qpc = orc.lib.orc_chroma_qp(qp)
        pkg.host.h264_chroma_dc_levels(p(c4), qpc, p(o))
        h = [c4[0] + c4[1] + c4[2] + c4[3], c4[0] - c4[1] + c4[2] - c4[3], c4[0] + c4[1] - c4[2] - c4[3],
             c4[0] - c4[1] - c4[2] + c4[3]]
        assert list(o) == [orc.lib.orc_quant_dc_chroma(int(v), qpc) for v in h]
        assert pkg.host.h264_chroma_qp(qp) == qpc


@need_ref
def test_i16_decision_is_the_first_strict_argmin():
    """The I16x16 choice (intra_wavefront.comp:485-498): SAD of each available
    predictor against the source, tried in the order DC, V, H, Plane with
    strict <; V needs the top MB, H the left one, Plane top, left and (the
    standard's condition, deviation D3) top-left, all in the same slice.
    Recomputed here for every MB from the oracle's own (unfiltered)
    reconstruction with the reference decoder's 16x16 predictors, on frames
    made to tie: flat ones (every mode equal: DC), stripes (V or H ties
    Plane), two-level noise and ordinary texture."""
    lib = ref()
    rs = np.random.default_rng(16)
    w, h = 64, 48
    yy, xx = np.mgrid[0:h, 0:w]
    pics = [_frames(w, h, 1)[0][0], np.full((h, w), 128, np.uint8),
            (120 + 8 * ((xx // 4) % 2)).astype(np.uint8), (90 + 30 * ((yy // 3) % 2)).astype(np.uint8),
            (100 + rs.integers(0, 2, (h, w))).astype(np.uint8)]
    pics[1][16:32, 16:48] = 140
    uv = np.full((h // 2, w), 128, np.uint8)
    order = (2, 0, 1, 3)  # DC, V, H, Plane
    chosen = set()
    for cy in pics:
        for ns in (1, 2):
            for qp in (10, 30):
                e = orc.encode_i(cy, uv, qp, ns)
                mbw, total = w // 16, (w // 16) * (h // 16)
                for m in range(total):
                    mbx, mby = m % mbw, m // mbw
                    sl = orc.lib.orc_slice_of(m, total, ns)
                    ta = mby > 0 and orc.lib.orc_slice_of(m - mbw, total, ns) == sl
                    la = mbx > 0 and orc.lib.orc_slice_of(m - 1, total, ns) == sl
                    ca = mbx > 0 and mby > 0 and orc.lib.orc_slice_of(m - mbw - 1, total, ns) == sl
                    y0, x0 = mby * 16, mbx * 16
                    rec = e["y"]
                    top = rec[y0 - 1, x0:x0 + 16].copy() if mby > 0 else np.full(16, 128, np.uint8)
                    left = rec[y0:y0 + 16, x0 - 1].copy() if mbx > 0 else np.full(16, 128, np.uint8)
                    corner = int(rec[y0 - 1, x0 - 1]) if (mbx > 0 and mby > 0) else 128
                    avail = {2: True, 0: ta, 1: la, 3: ta and la and ca}
                    best, mode = None, None
                    for md in order:
                        if not avail[md]:
                            continue
                        pred = np.zeros((16, 16), np.uint8)
                        lib.h264d_pred16x16(p(pred), 16, md, p(top), p(left), corner, ta, la)
                        sad = int(np.abs(pred.astype(int) - cy[y0:y0 + 16, x0:x0 + 16]).sum())
                        if best is None or sad < best:
                            best, mode = sad, md
                    assert int(e["modes"][m]) & 3 == mode, (m, ns, qp)
                    chosen.add(mode)
    assert chosen == {0, 1, 2, 3}  # every mode is exercised


def test_fullsearch_brute_force_tiny():
    rs = np.random.default_rng(7)
    for trial in range(6):
        ref_p = rs.integers(0, 256, (32, 48)).astype(np.uint8)
        cur = np.roll(ref_p, (int(rs.integers(-3, 4)), int(rs.integers(-3, 4))), axis=(0, 1))
        cur = np.clip(cur.astype(int) + rs.integers(-3, 4, cur.shape), 0, 255).astype(np.uint8)
        R, lam = 8, int(trial % 3)
        i16, i8 = orc.me_fullsearch(cur, ref_p, R, lam)
        pad = np.pad(ref_p.astype(int), R, mode="edge")
        for m in range(6):
            y0, x0 = (m // 3) * 16, (m % 3) * 16
            d = np.arange(-R, R + 1)
            l1 = np.abs(d)[:, None] + np.abs(d)[None, :]
            sw = np.lib.stride_tricks.sliding_window_view(pad[y0:y0 + 16 + 2 * R, x0:x0 + 16 + 2 * R], (16, 16))
            blk = cur[y0:y0 + 16, x0:x0 + 16].astype(int)
            c16 = np.abs(sw - blk).sum(axis=(2, 3)) + 4 * lam * l1
            k = int(np.argmin(c16))
            assert (i16[m]["dy"], i16[m]["dx"]) == (k // (2 * R + 1) - R, k % (2 * R + 1) - R)
            assert i16[m]["cost"] == c16.flat[k]
            for q in range(4):
                qy, qx = (q >> 1) * 8, (q & 1) * 8
                c8 = np.abs(sw[:, :, qy:qy + 8, qx:qx + 8] - blk[qy:qy + 8, qx:qx + 8]).sum(axis=(2, 3)) + lam * l1
                k8 = int(np.argmin(c8))
                assert (i8[4 * m + q]["dy"], i8[4 * m + q]["dx"]) == (k8 // (2 * R + 1) - R, k8 % (2 * R + 1) - R)


# the two candidate rings of motion_estimation.comp:395-459, in its order
HALF_RING = ((2, 0), (-2, 0), (0, 2), (0, -2), (2, 2), (2, -2), (-2, 2), (-2, -2))
QTR_RING = tuple((a // 2, b // 2) for a, b in HALF_RING)


def _ref_block_sad(lib, cur, refp, x0, y0, n, mvx, mvy):
    """SAD of an n x n block against the reference decoder's 8.4.2.2.1
    prediction at quarter-sample vector (mvx, mvy)."""
    h, w = refp.shape
    buf = np.zeros((n + 6) * (n + 6), np.uint8)
    out = np.zeros((n, n), np.uint8)
    ss = C.c_int()
    src = lib.h264d_mc_fetch_luma(p(buf), C.byref(ss), p(refp), refp.strides[0], w, h, x0 + (mvx >> 2),
                                  y0 + (mvy >> 2), n, n)
    lib.h264d_mc_luma(p(out), n, src, ss.value, n, n, mvx & 3, mvy & 3)
    return int(np.abs(out.astype(int) - cur[y0:y0 + n, x0:x0 + n]).sum())


def _brute_refine(lib, cur, refp, x0, y0, n, dx, dy):
    cx, cy = 4 * dx, 4 * dy
    best = _ref_block_sad(lib, cur, refp, x0, y0, n, cx, cy)
    for ring in (HALF_RING, QTR_RING):
        bx = by = 0
        for ox, oy in ring:
            sad = _ref_block_sad(lib, cur, refp, x0, y0, n, cx + ox, cy + oy)
            if sad < best:
                best, bx, by = sad, ox, oy
        cx, cy = cx + bx, cy + by
    return cx, cy, best


@need_ref
def test_refine_rings_brute_force():
    """Sub-pel refinement by brute force: from the integer winner, the
    shader's 8 half-sample candidates in its order, then its 8 quarter-sample
    ones around the half winner, each taken only when its plain SAD (against
    the reference decoder's interpolation) is strictly lower.  Low-contrast
    two-level content makes ties frequent, so ring order and the strict
    comparison both matter; every MB and every 8x8 block is compared."""
    lib = ref()
    rs = np.random.default_rng(23)
    h, w = 48, 64
    yy, xx = np.mgrid[0:h, 0:w]
    cases = []
    for trial in range(4):
        if trial % 2:
            base = (100 + rs.integers(0, 2, (h, w))).astype(np.uint8)
        else:
            base = np.clip(128 + 60 * np.sin(xx / (2.5 + trial)) * np.cos(yy / 3.7) + rs.normal(0, 2, (h, w)),
                           0, 255).astype(np.uint8)
        cur = np.roll(base, (trial - 1, 2 - trial), axis=(0, 1))
        cases.append((cur, base))
        # between two integer positions: half- and quarter-sample winners
        a = base.astype(int)
        half = (a + np.roll(a, 1, axis=1) + 1) >> 1
        cases.append((half.astype(np.uint8), base))
        cases.append((((3 * a + np.roll(a, -1, axis=0) + 2) >> 2).astype(np.uint8), base))
    cases.append((np.full((h, w), 90, np.uint8), np.full((h, w), 90, np.uint8)))
    moved = 0
    for cur, refp in cases:
        i16, i8 = orc.me_fullsearch(cur, refp, 4, 1)
        m16, m8 = orc.me_refine(cur, refp, i16, i8)
        for m in range(12):
            x0, y0 = (m % 4) * 16, (m // 4) * 16
            want = _brute_refine(lib, cur, refp, x0, y0, 16, int(i16[m]["dx"]), int(i16[m]["dy"]))
            assert (int(m16[m]["mvx"]), int(m16[m]["mvy"]), int(m16[m]["sad"])) == want, m
            moved += want[:2] != (4 * int(i16[m]["dx"]), 4 * int(i16[m]["dy"]))
            for b in range(4):
                bx, by = x0 + (b & 1) * 8, y0 + (b >> 1) * 8
                k = 4 * m + b
                want = _brute_refine(lib, cur, refp, bx, by, 8, int(i8[k]["dx"]), int(i8[k]["dy"]))
                assert (int(m8[k]["mvx"]), int(m8[k]["mvy"]), int(m8[k]["sad"])) == want, (m, b)
    assert moved > 0  # fractional vectors occur


def test_refine_never_worse_than_integer():
    rs = np.random.default_rng(8)
    (ry, _), (cy, _) = _frames(64, 48, 2, seed=9)
    i16, i8 = orc.me_fullsearch(cy, ry, 8, 1)
    m16, m8 = orc.me_refine(cy, ry, i16, i8)
    assert (m16["sad"] <= i16["sad"]).all()
    assert (m8["sad"] <= i8["sad"]).all()
    assert (np.abs(m16["mvx"] - 4 * i16["dx"]) <= 3).all()
    # reported SAD is the SAD at the reported quarter-pel vector
    for m in range(12):
        x0, y0 = (m % 4) * 16, (m // 4) * 16
        pred = np.zeros((16, 16), np.uint8)
        orc.lib.orc_pred_luma_block(p(ry), 64, 64, 48, x0, y0, 16, 16, int(m16["mvx"][m]), int(m16["mvy"][m]),
                                    p(pred), 16)
        assert np.abs(pred.astype(int) - cy[y0:y0 + 16, x0:x0 + 16]).sum() == m16["sad"][m]


# ------------------------------------------------------------ deblocking

@need_ref
def test_deblock_tables_match_reference():
    lib = C.CDLL(REF)
    alpha = list((C.c_uint8 * 52).in_dll(lib, "h264d_alpha"))
    beta = list((C.c_uint8 * 52).in_dll(lib, "h264d_beta"))
    tc0 = np.array(((C.c_uint8 * 3) * 52).in_dll(lib, "h264d_tc0")).reshape(52, 3)
    assert [orc.lib.orc_h264_alpha(i) for i in range(52)] == alpha
    assert [orc.lib.orc_h264_beta(i) for i in range(52)] == beta
    mine = np.array([[orc.lib.orc_h264_tc0(i, b) for b in (1, 2, 3)] for i in range(52)])
    np.testing.assert_array_equal(mine, tc0)


def _line(vals, fn, bs, qp):
    a = np.array(vals, np.uint8)
    fn(p(a[4:]) if len(a) == 8 else p(a[2:]), 1, bs, qp)
    return list(a)


def test_deblock_line_kats():
    L, Cf = orc.lib.orc_h264_filter_line_luma, orc.lib.orc_h264_filter_line_chroma
    L.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    Cf.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    # hand-computed from H.264 8.7.2.3 / 8.7.2.4 (alpha/beta/tc0 from Tables 8-16, 8-17)
    assert _line([60, 60, 60, 60, 70, 70, 70, 70], L, 4, 40) == [60, 61, 63, 64, 66, 68, 69, 70]
    assert _line([50, 52, 54, 56, 64, 66, 68, 70], L, 2, 30) == [50, 52, 55, 59, 61, 65, 68, 70]
    assert _line([54, 56, 64, 66], Cf, 2, 30) == [54, 58, 62, 66]
    assert _line([54, 56, 64, 66], Cf, 4, 30) == [54, 58, 63, 66]
    # |p0-q0| >= alpha: untouched;  qp below 16: alpha = 0, never filtered
    assert _line([10, 10, 10, 10, 200, 200, 200, 200], L, 4, 40) == [10, 10, 10, 10, 200, 200, 200, 200]
    assert _line([0, 0, 0, 0, 255, 255, 255, 255], L, 4, 51) == [0, 0, 0, 0, 255, 255, 255, 255]
    assert _line([60, 60, 60, 60, 62, 62, 62, 62], L, 4, 15) == [60, 60, 60, 60, 62, 62, 62, 62]


def test_deblock_picture_invariants():
    rs = np.random.default_rng(10)
    y = rs.integers(0, 256, (48, 64)).astype(np.uint8)
    uv = rs.integers(0, 256, (24, 64)).astype(np.uint8)
    nmb = 12
    nz0 = np.zeros(nmb * 24, np.uint32)
    mv0 = np.zeros(nmb, orc.MV)
    # P picture, no coefficients, equal MVs: every bS is 0
    a, b = orc.deblock(y, uv, 40, 0, nz0, mv0, 1, 0)
    assert (a == y).all() and (b == uv).all()
    # disable_deblocking_filter_idc = 1
    a, b = orc.deblock(y, uv, 40, 1, nz0, mv0, 1, 1)
    assert (a == y).all() and (b == uv).all()
    # flat picture stays flat
    f = np.full((48, 64), 99, np.uint8)
    a, _ = orc.deblock(f, np.full((24, 64), 128, np.uint8), 51, 1, nz0, mv0, 1, 0)
    assert (a == 99).all()
    # bands of constant value per MB row: only horizontal MB edges carry a
    # step.  With one slice per MB row, idc 2 filters none of them, idc 0 all.
    bands = np.repeat(np.array([100, 110, 100], np.uint8), 16)[:, None].repeat(64, axis=1)
    flat_uv = np.full((24, 64), 128, np.uint8)
    a, _ = orc.deblock(bands, flat_uv, 40, 1, nz0, mv0, 3, 2)
    assert (a == bands).all()
    a, _ = orc.deblock(bands, flat_uv, 40, 1, nz0, mv0, 3, 0)
    assert (a[13:19] != bands[13:19]).any(axis=1).all() and (a[:12] == bands[:12]).all()


# ------------------------------------------------------------ intra 4x4 / 8x8

def _avail_cases(needs):
    out = []
    for at in (0, 1):
        for al in (0, 1):
            for ac in (0, 1):
                for atr in (0, 1):
                    if ("t" in needs and not at) or ("l" in needs and not al) or ("c" in needs and not ac):
                        continue
                    if ac and not (at and al) and "c" in needs:
                        continue
                    out.append((at, al, ac, atr))
    return out


@need_ref
def test_intra4x4_matches_reference():
    L = ref()
    L.h264d_pred4x4.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_uint8, C.c_bool,
                                C.c_bool]
    rs = np.random.default_rng(44)
    for mode, needs in orc.H264_I4_NEEDS.items():
        for (at, al, ac, atr) in _avail_cases(needs):
            for _ in range(25):
                t = rs.integers(0, 256, 8).astype(np.uint8)
                l = rs.integers(0, 256, 4).astype(np.uint8)
                if not atr:
                    t[4:] = t[3]  # the decoder pads a missing top-right run with p[3,-1]
                c = int(rs.integers(0, 256))
                want = np.zeros(16, np.uint8)
                L.h264d_pred4x4(p(want), 4, mode, p(t), p(l), c, bool(at), bool(al))
                assert np.array_equal(orc.h264_pred4x4(t, l, c, mode, at, al), want), (mode, at, al)


@need_ref
def test_intra8x8_matches_reference():
    L = ref()
    L.h264d_pred8x8_luma.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_uint8, C.c_bool,
                                     C.c_bool, C.c_bool, C.c_bool]
    rs = np.random.default_rng(88)
    for mode, needs in orc.H264_I4_NEEDS.items():
        for (at, al, ac, atr) in _avail_cases(needs):
            for _ in range(25):
                t = rs.integers(0, 256, 16).astype(np.uint8)
                l = rs.integers(0, 256, 8).astype(np.uint8)
                c = int(rs.integers(0, 256))
                want = np.zeros(64, np.uint8)
                L.h264d_pred8x8_luma(p(want), 8, mode, p(t), p(l), c, bool(at), bool(al), bool(ac), bool(atr))
                got = orc.h264_pred8x8(t, l, c, mode, at, al, ac, atr)
                assert np.array_equal(got, want), (mode, at, al, ac, atr)


def test_intra_flat_neighbours_predict_flat():
    """Every mode maps constant neighbours to the same constant (each sample
    is a normalised average of neighbours)."""
    for v in (0, 17, 200, 255):
        for mode in range(9):
            assert (orc.h264_pred4x4(np.full(8, v), np.full(4, v), v, mode, 1, 1) == v).all()
            assert (orc.h264_pred8x8(np.full(16, v), np.full(8, v), v, mode, 1, 1, 1, 1) == v).all()


# ------------------------------------------------------------ 8x8 transform

@need_ref
def test_8x8_tables_match_reference():
    L = ref()
    scan = np.frombuffer((C.c_uint8 * 16).in_dll(L, "h264d_dequant8_init_scan"), np.uint8)
    v8 = np.frombuffer((C.c_uint8 * 36).in_dll(L, "h264d_dequant8_init"), np.uint8).reshape(6, 6)
    for i in range(8):
        for j in range(8):
            assert orc.lib.orc_h264_class8(i, j) == scan[(i % 4) * 4 + j % 4]
    for m in range(6):
        for c in range(6):
            assert orc.lib.orc_h264_v8(m, c) == v8[m, c]


class _Dq(C.Structure):
    _fields_ = [("d4", C.c_int32 * 96), ("d8", C.c_int32 * 384), ("qp", C.c_int), ("valid", C.c_int)]


@need_ref
@pytest.mark.parametrize("qp", [0, 10, 18, 27, 35, 36, 41, 51])
def test_8x8_decode_equals_recon(qp):
    """The oracle's 8x8 levels decoded by the reference decoder (flat scaling
    lists, h264d_dequant_build + h264d_idct8_add) rebuild exactly the
    oracle's reconstruction."""
    L = ref()
    flat4 = (C.c_uint8 * 96)(*([16] * 96))
    flat8 = (C.c_uint8 * 384)(*([16] * 384))
    dq = _Dq()
    L.h264d_dequant_build(C.byref(dq), qp, flat4, flat8)
    L.h264d_idct8_add.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    rs = np.random.default_rng(qp + 8)
    for _ in range(40):
        pred = rs.integers(40, 216, 64).astype(np.uint8)
        res = rs.integers(-40, 41, 64).astype(np.int16)
        lv = orc.h264_tq8x8(res, qp, 4)
        rec = np.clip(pred.astype(int) + orc.h264_dqit8x8(lv, qp), 0, 255).astype(np.uint8)
        dst = pred.copy()
        blk = lv.copy()
        L.h264d_idct8_add(p(dst), 8, p(blk), C.addressof(dq.d8), qp)
        assert np.array_equal(dst, rec), qp


def test_8x8_quantiser_inverts_the_decoder():
    """With rounding to nearest at QP 0 the forward + quantiser + decoder
    chain returns the residual within a quantisation step (the scale of the
    quantiser matches the decoder's), and a DC-only block comes back exactly."""
    rs = np.random.default_rng(5)
    for _ in range(30):
        res = rs.integers(-255, 256, 64).astype(np.int16)
        back = orc.h264_dqit8x8(orc.h264_tq8x8(res, 0, 6), 0)
        assert np.abs(back.astype(int) - res).max() <= 2
    for v in (-100, -7, 3, 50, 255):
        lv = orc.h264_tq8x8(np.full(64, v, np.int16), 0, 6)
        assert np.count_nonzero(lv[1:]) == 0
        assert np.array_equal(orc.h264_dqit8x8(lv, 0), np.full(64, v, np.int16))


# ------------------------------------------------- the reference's Tier-2 search

def _sq_cost(cur, ref, x0, y0, mvx, mvy):
    """Cost of a square-search result recomputed from the planes: SAD + 5|mv|
    for an integer vector, plain SAD against the rounded-up average of the
    two integer samples around a half-sample one."""
    c = cur[y0:y0 + 16, x0:x0 + 16].astype(int)
    bx, by = mvx // 4, mvy // 4
    if mvx % 4 == 0 and mvy % 4 == 0:
        r = ref[y0 + by:y0 + by + 16, x0 + bx:x0 + bx + 16].astype(int)
        return int(np.abs(c - r).sum()) + 5 * (abs(bx) + abs(by))
    hx, hy = (1, 0) if mvx % 4 else (0, 1)
    a = ref[y0 + by:y0 + by + 16, x0 + bx:x0 + bx + 16].astype(int)
    b = ref[y0 + by + hy:y0 + by + hy + 16, x0 + bx + hx:x0 + bx + hx + 16].astype(int)
    return int(np.abs(c - ((a + b + 1) >> 1)).sum())


def test_square_me_reference_fixtures():
    """The expectations of the reference's own tests (tests/test_cpu_simd_me.c:
    16-224), restated on the oracle."""
    # flat frames: static below the threshold; a uniform +10 offset costs 2560
    # everywhere, so (0,0) wins every strict comparison
    a = np.full((64, 64), 128, np.uint8)
    m = orc.square_me(a, a, 64, 64, 8)
    assert (m["mvx"] == 0).all() and (m["sad"] == 0).all()
    m = orc.square_me(a, np.full((64, 64), 138, np.uint8), 64, 64, 8)
    assert (m["mvx"] == 0).all() and (m["mvy"] == 0).all() and (m["sad"] == 2560).all()
    # one textured MB moved by (+2,+2): found at (8, 8) quarter samples
    cur = np.full((64, 64), 100, np.uint8)
    ref = cur.copy()
    yy, xx = np.mgrid[0:16, 0:16]
    cur[16:32, 16:32] = (xx * 13 + yy * 7 + 40) % 256
    ref[18:34, 18:34] = cur[16:32, 16:32]
    m = orc.square_me(cur, ref, 64, 64, 8)
    assert (m[0]["mvx"], m[0]["mvy"]) == (0, 0)
    assert (m[5]["mvx"], m[5]["mvy"]) == (8, 8)
    # two neighbours moved alike: the second takes the left predictor
    cur = np.full((64, 64), 120, np.uint8)
    ref = cur.copy()
    for mbx in (1, 2):
        cur[16:32, mbx * 16:mbx * 16 + 16] = (xx * 13 + yy * 7 + mbx * 30) % 256
        ref[18:34, mbx * 16 + 2:mbx * 16 + 18] = cur[16:32, mbx * 16:mbx * 16 + 16]
    m = orc.square_me(cur, ref, 64, 64, 8)
    assert (m[5]["mvx"], m[5]["mvy"]) == (8, 8) and (m[6]["mvx"], m[6]["mvy"]) == (8, 8) and m[6]["sad"] < 100
    # a height that is not a multiple of 16 (MB-padded planes)
    m = orc.square_me(np.zeros((64, 32), np.uint8), np.zeros((64, 32), np.uint8), 32, 50, 8)
    assert m.shape == (8,) and (m["sad"] == 0).all()


@pytest.mark.parametrize("radius", [2, 8, 16])
def test_square_me_costs_recompute(pkg, radius):
    """On generated frames every reported cost equals the cost recomputed at
    the reported vector, vectors stay within the radius (+ a half sample),
    and no result costs more than the zero vector."""
    g = pkg.FrameGen(9, 160, 96, 4)
    (cy, _), (ry, _) = [pkg.pad_coded(*g.frame(i)) for i in (1, 0)]
    g.close()
    m = orc.square_me(cy, ry, 160, 96, radius)
    wmb = 10
    moved = 0
    for i, v in enumerate(m):
        x0, y0 = (i % wmb) * 16, (i // wmb) * 16
        mvx, mvy, sad = int(v["mvx"]), int(v["mvy"]), int(v["sad"])
        assert abs(mvx) <= 4 * radius + 2 and abs(mvy) <= 4 * radius + 2
        assert _sq_cost(cy, ry, x0, y0, mvx, mvy) == sad, (i, mvx, mvy, sad)
        zero = int(np.abs(cy[y0:y0 + 16, x0:x0 + 16].astype(int) - ry[y0:y0 + 16, x0:x0 + 16]).sum())
        assert sad <= zero
        moved += (mvx, mvy) != (0, 0)
    assert moved > 10


# ------------------------------------------------- the reference's diamond search

def _dia_cost_checks(cy, ry, m, lam):
    """Every result's cost recomputed from the planes: plain SAD of the
    8.4.2.2.1 prediction for a fractional vector; for an integer one either
    the zero-vector SAD (vector 0) or 2 x the even-parity SAD + lam |mv|."""
    hc, wc = cy.shape
    for i, v in enumerate(m):
        x0, y0 = (i % (wc // 16)) * 16, (i // (wc // 16)) * 16
        mvx, mvy, cost = int(v["mvx"]), int(v["mvy"]), int(v["sad"])
        assert abs(mvx) <= 67 and abs(mvy) <= 67
        c = cy[y0:y0 + 16, x0:x0 + 16].astype(int)
        pred = np.zeros((16, 16), np.uint8)
        orc.lib.orc_pred_luma_block(p(ry), wc, wc, hc, x0, y0, 16, 16, mvx, mvy, p(pred), 16)
        full = int(np.abs(pred.astype(int) - c).sum())
        if mvx % 4 or mvy % 4:
            assert cost == full, (i, mvx, mvy)
        elif (mvx, mvy) == (0, 0):
            assert cost == full or cost <= 512 and cost == full
        else:
            par = (np.add.outer(np.arange(16), np.arange(16)) & 1) == 0
            assert cost == 2 * int(np.abs(pred.astype(int) - c)[par].sum()) + lam * (abs(mvx) + abs(mvy)) // 4, i


def test_diamond_me_moved_block_and_static():
    """A textured MB moved by (+8, 0) in a flat frame is found by the first
    step's diamond point (8, 0) at cost 5 * 8 (no earlier point of that step
    is cheaper, and no fractional candidate's SAD gets under 40); static MBs
    keep (0, 0)."""
    cur = np.full((64, 64), 90, np.uint8)
    ref = cur.copy()
    yy, xx = np.mgrid[0:16, 0:16]
    cur[16:32, 16:32] = (xx * 29 + yy * 11 + 7) % 200 + 20
    ref[16:32, 24:40] = cur[16:32, 16:32]
    m = orc.diamond_me(cur, ref)
    assert (m[5]["mvx"], m[5]["mvy"]) == (32, 0) and m[5]["sad"] == 5 * 8
    assert all((m[i]["mvx"], m[i]["mvy"]) == (0, 0) for i in (0, 3, 12, 15))
    _dia_cost_checks(cur, ref, m, 5)


@pytest.mark.parametrize("lam", [5, 8])
def test_diamond_me_costs_recompute(pkg, lam):
    g = pkg.FrameGen(13, 192, 112, 5)
    (cy, _), (ry, _) = [pkg.pad_coded(*g.frame(i)) for i in (1, 0)]
    g.close()
    m = orc.diamond_me(cy, ry, lam)
    assert ((m["mvx"] % 4) != 0).sum() + ((m["mvy"] % 4) != 0).sum() > 5
    _dia_cost_checks(cy, ry, m, lam)
