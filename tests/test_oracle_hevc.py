"""Pins the HEVC half of the CPU oracle (oracle/orc_hevc.c) before it checks
the HIP kernels.

1. Against the reference's own functions, compiled from its sources into
   oracle/_ref/libref_hevc.so (hevc_intra.c, hevc_mc.c, hevc_dec_tables.c):
   4x4 DST-VII / DCT-II and 8x8 DCT-II with quantisation and
   dequantisation + inverse, at 8 and 10 bits, every QP and rounding;
   uni-prediction MC (8-tap luma, 4-tap chroma) at every fraction,
   including blocks hanging over the picture edge; 4x4 intra prediction and
   the SAD mode decision at every position class of the CTU16 / CU8 / PU4
   z-scan, slice tops included.
2. Whole I and P pictures: the oracle's frame coding against the same
   reference functions driven in the reference's coding order
   (hevc_enc_template.c: encode_ctu -> intra_trial; predict_cu ->
   inter_residual with the 8x8 transform).
3. Closed forms: DC-only round trips, the transforms' matrix products, the
   quantiser as the inverse of the dequantiser, refinement never worse than
   its start.
4. CTU32 (no reference counterpart): the z-scan rank against the standard's
   MinTbAddrZs formula (6.5.2), CTU32-row slices, and CTU32 coding equal to
   CTU16 coding where both orders and neighbourhoods coincide.
"""
import ctypes as C
import os

import numpy as np
import pytest

import orc

need_ref = pytest.mark.skipif(not os.path.exists(orc.REF_HEVC), reason="oracle/_ref not built (reference tree absent)")
_ref = None
P, I = C.c_void_p, C.c_int


def ref():
    global _ref
    if _ref is None:
        # lazy binding: the decoder file also references decoder-only
        # functions that these entry points never call
        lib = C.CDLL(orc.REF_HEVC, mode=os.RTLD_LAZY)
        for n in ("hevc_transform_quant_4x4", "hevc_transform_quant_4x4_10"):
            getattr(lib, n).argtypes = [P, I, I, I, P]
        for n in ("hevc_dequant_itransform_4x4", "hevc_dequant_itransform_4x4_10"):
            getattr(lib, n).argtypes = [P, I, I, P]
        lib.hevc_transform_quant_8x8.argtypes = [P, I, I, I, P]
        lib.hevc_dequant_itransform_8x8.argtypes = [P, I, I, P]
        for n in ("hevc_mc_uni_8", "hevc_mc_uni_10"):
            getattr(lib, n).argtypes = [P, I, I, I, I, I, I, I, I, I, I, P, I]
        for n in ("hevc_predict_4x4", "hevc_predict_4x4_10"):
            getattr(lib, n).argtypes = [P, I, I, I, I, I, I, I, I, P]
        for n in ("hevc_choose_luma_mode", "hevc_choose_luma_mode_10"):
            getattr(lib, n).argtypes = [I, P, P, I, I, I, I, I, P]
            getattr(lib, n).restype = I
        lib.hevc_chroma_qp_from_luma.argtypes = [I]
        lib.hevc_chroma_qp_from_luma.restype = I
        _ref = lib
    return _ref


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def pix(bd):
    return np.uint8 if bd == 8 else np.uint16


# ------------------------------------------------------------ transforms

@need_ref
@pytest.mark.parametrize("bd", [8, 10])
def test_tq4_and_dqit4_match_reference(bd):
    L = ref()
    rs = np.random.default_rng(40 + bd)
    vmax = (1 << bd) - 1
    tq = L.hevc_transform_quant_4x4 if bd == 8 else L.hevc_transform_quant_4x4_10
    dq = L.hevc_dequant_itransform_4x4 if bd == 8 else L.hevc_dequant_itransform_4x4_10
    qps = range(0, 52) if bd == 8 else range(0, 64)
    for qp in qps:
        for dst in (0, 1):
            for rnd in (2, 3, 4, 6):
                for k in range(6):
                    amp = [vmax, 40, 4, vmax][k % 4]
                    res = rs.integers(-amp, amp + 1, 16).astype(np.int16)
                    if k == 5:
                        res[:] = rs.integers(-vmax, vmax + 1)  # constant block
                    want = np.zeros(16, np.int16)
                    tq(p(res), qp, dst, rnd, p(want))
                    got = orc.hevc_tq4(res, qp, dst, rnd, bd)
                    assert np.array_equal(got, want), (bd, qp, dst, rnd, res)
                    back = np.zeros(16, np.int16)
                    dq(p(want), qp, dst, p(back))
                    assert np.array_equal(orc.hevc_dqit4(want, qp, dst, bd), back), (bd, qp, dst)
    # levels far outside what a residual produces still dequantise the same
    for qp in (0, 30, 51):
        lv = rs.integers(-32768, 32768, 16).astype(np.int16)
        back = np.zeros(16, np.int16)
        dq(p(lv), qp, 1, p(back))
        assert np.array_equal(orc.hevc_dqit4(lv, qp, 1, bd), back)


@need_ref
@pytest.mark.parametrize("bd", [8, 10])
def test_tq8_and_dqit8_match_reference(bd):
    L = ref()
    rs = np.random.default_rng(80 + bd)
    vmax = (1 << bd) - 1
    for qp in range(0, 64, 1 if bd == 10 else 2):
        for rnd in (2, 6):
            for amp in (vmax, 30, 3):
                res = rs.integers(-amp, amp + 1, 64).astype(np.int16)
                want = np.zeros(64, np.int16)
                L.hevc_transform_quant_8x8(p(res), qp, bd, rnd, p(want))
                assert np.array_equal(orc.hevc_tq8(res, qp, bd, rnd), want), (bd, qp, rnd, amp)
                back = np.zeros(64, np.int16)
                L.hevc_dequant_itransform_8x8(p(want), qp, bd, p(back))
                assert np.array_equal(orc.hevc_dqit8(want, qp, bd), back), (bd, qp)


@need_ref
def test_chroma_qp_matches_reference():
    for q in range(-3, 60):
        assert orc.lib.orc_hevc_chroma_qp(q) == ref().hevc_chroma_qp_from_luma(q)


def test_dc_only_round_trip_is_exact():
    """A constant residual v: forward (1, 8) shifts give coefficient 128*v at
    (0,0) and nothing else; dequant + inverse at a QP fine enough returns v."""
    for v in (-200, -17, -1, 1, 5, 99, 255):
        res = np.full(16, v, np.int16)
        lv = orc.hevc_tq4(res, 4, 0, 6)            # step 16*45 >> ... fine at QP 4
        assert np.count_nonzero(lv[1:]) == 0
        assert np.array_equal(orc.hevc_dqit4(lv, 4, 0), res)
        lv8 = orc.hevc_tq8(np.full(64, v, np.int16), 4, 8, 6)
        assert np.count_nonzero(lv8[1:]) == 0
        assert np.array_equal(orc.hevc_dqit8(lv8, 4, 8), np.full(64, v, np.int16))


def test_inverse_transform_is_matrix_product():
    """Dequantised coefficients through the inverse: (M^T C M) with the two
    rounding shifts of 8.6.4.2, for both 4x4 kinds and 8x8."""
    dct4 = np.array([[64, 64, 64, 64], [83, 36, -36, -83], [64, -64, -64, 64], [36, -83, 83, -36]])
    dst4 = np.array([[29, 55, 74, 84], [74, 74, 0, -74], [84, -29, -74, 55], [55, -84, 74, -29]])
    n8 = np.arange(8)
    # 8-point DCT-II from its definition: round(64*sqrt2*cos((2n+1)k*pi/16)) with the standard's tweaks
    dct8 = np.array([[64] * 8, [89, 75, 50, 18, -18, -50, -75, -89], [83, 36, -36, -83, -83, -36, 36, 83],
                     [75, -18, -89, -50, 50, 89, 18, -75], [64, -64, -64, 64, 64, -64, -64, 64],
                     [50, -89, 18, 75, -75, -18, 89, -50], [36, -83, 83, -36, -36, 83, -83, 36],
                     [18, -50, 75, -89, 89, -75, 50, -18]])
    ideal = np.round(64 * np.sqrt(2) * np.cos((2 * n8[None, :] + 1) * n8[:, None] * np.pi / 16))
    ideal[0] = 64
    assert np.abs(dct8 - ideal).max() <= 2  # same basis as the real DCT
    rs = np.random.default_rng(7)
    for M, n, kind in ((dct4, 4, 0), (dst4, 4, 1), (dct8, 8, 2)):
        for _ in range(50):
            # 8.6.3 scaling with the flat list, levelScale[q % 6] = {40,45,51,57,64,72}
            lv = rs.integers(-20, 21, n * n).astype(np.int16)
            q = 4
            ls = (40, 45, 51, 57, 64, 72)[q % 6]
            bds = 5 if n == 4 else 6
            d = np.clip((lv.astype(np.int64) * (ls << (q // 6 + 4)) + (1 << (bds - 1))) >> bds, -32768, 32767)
            D = d.reshape(n, n)
            g = np.clip((M.T @ D + 64) >> 7, -32768, 32767)
            want = np.clip((g @ M + 2048) >> 12, -32768, 32767).reshape(-1)
            got = orc.hevc_dqit4(lv, q, kind) if n == 4 else orc.hevc_dqit8(lv, q, 8)
            assert np.array_equal(got, want)


def test_quantiser_inverts_dequantiser():
    """With rounding to nearest, quantising a dequantised level gives the
    level back (the quantiser is the algebraic inverse of 8.6.3)."""
    rs = np.random.default_rng(9)
    for q in range(0, 52):
        lv = rs.integers(-300, 301, 64).astype(np.int16)
        d = np.zeros(64, np.int16)
        orc.lib.orc_hevc_dequant(p(lv), 64, q, 5, p(d))
        ok = np.abs(d.astype(int)) < 32767
        back = np.zeros(64, np.int16)
        orc.lib.orc_hevc_quant(p(d.astype(np.int32)), 64, 5, q, 6, p(back))
        assert np.array_equal(back[ok], lv[ok]), q


# ------------------------------------------------------------ MC

@need_ref
@pytest.mark.parametrize("bd", [8, 10])
def test_mc_uni_matches_reference(bd):
    L = ref()
    f = L.hevc_mc_uni_8 if bd == 8 else L.hevc_mc_uni_10
    rs = np.random.default_rng(3 + bd)
    plane = rs.integers(0, 1 << bd, (40, 56)).astype(pix(bd))
    for chroma in (0, 1):
        nfrac = 8 if chroma else 4
        for fy in range(nfrac):
            for fx in range(nfrac):
                for (x, y, w, h) in ((8, 8, 8, 8), (0, 0, 4, 4), (52, 36, 4, 4), (-9, 30, 8, 8), (20, -6, 16, 8)):
                    mvx = int(rs.integers(-6, 7)) * nfrac + fx
                    mvy = int(rs.integers(-6, 7)) * nfrac + fy
                    want = np.zeros((h, w), pix(bd))
                    f(p(plane), plane.shape[1], plane.shape[1], plane.shape[0], x, y, w, h, mvx, mvy, chroma,
                      p(want), w)
                    got = orc.hevc_mc_uni(plane, x, y, w, h, mvx, mvy, chroma, bd)
                    assert np.array_equal(got, want), (bd, chroma, fx, fy, x, y, w, h)


# ------------------------------------------------------------ intra

def _positions(w, h):
    out = []
    for y0 in range(0, h, 4):
        for x0 in range(0, w, 4):
            out.append((x0, y0))
    return out


@need_ref
@pytest.mark.parametrize("bd", [8, 10])
def test_predict4_matches_reference(bd):
    L = ref()
    f = L.hevc_predict_4x4 if bd == 8 else L.hevc_predict_4x4_10
    rs = np.random.default_rng(11 + bd)
    h, w = 48, 64
    plane = rs.integers(0, 1 << bd, (h, w)).astype(pix(bd))
    for is_luma in (1, 0):
        for y_min in (0, 16, 32):
            for (x0, y0) in _positions(w, h):
                for mode in (0, 1, 10, 26):
                    want = np.zeros(16, pix(bd))
                    f(p(plane), w, w, h, x0, y0, mode, is_luma, y_min, p(want))
                    got = orc.hevc_predict4(plane, x0, y0, mode, is_luma, y_min, bd)
                    assert np.array_equal(got, want.astype(np.int32)), (bd, is_luma, y_min, x0, y0, mode)


@need_ref
@pytest.mark.parametrize("bd", [8, 10])
def test_choose_mode_matches_reference(bd):
    L = ref()
    f = L.hevc_choose_luma_mode if bd == 8 else L.hevc_choose_luma_mode_10
    rs = np.random.default_rng(21 + bd)
    h, w = 48, 64
    for trial in range(3):
        base = rs.integers(0, 1 << bd, (h, w))
        recon = np.clip(base + rs.integers(-8, 9, (h, w)), 0, (1 << bd) - 1).astype(pix(bd))
        src = np.clip(base + rs.integers(-20, 21, (h, w)), 0, (1 << bd) - 1).astype(pix(bd))
        for (x0, y0) in _positions(w, h):
            want = np.zeros(16, pix(bd))
            m = f(16 * trial, p(src), p(recon), w, w, h, x0, y0, p(want))
            gm, gp = orc.hevc_choose_mode(src, recon, x0, y0, 16 * trial, bd)
            assert gm == m and np.array_equal(gp, want.astype(np.int32)), (bd, x0, y0)


# ------------------------------------------------------------ whole pictures

def _frame(rs, w, h, bd=8):
    """Smooth content with texture, so the mode decision and the quantiser
    both see varied input."""
    yy, xx = np.mgrid[0:h, 0:w]
    base = 128 + 60 * np.sin(xx / 7.0) * np.cos(yy / 5.0)
    y = np.clip(base + rs.normal(0, 6, (h, w)), 0, 255).astype(int)
    cb = np.clip(128 + 30 * np.sin(xx[:h // 2, :w // 2] / 4.0) + rs.normal(0, 3, (h // 2, w // 2)), 0, 255).astype(int)
    cr = np.clip(128 - 25 * np.cos(yy[:h // 2, :w // 2] / 3.0) + rs.normal(0, 3, (h // 2, w // 2)), 0, 255).astype(int)
    sc = 1 if bd == 8 else 4
    return [(a * sc).astype(pix(bd)) for a in (y, cb, cr)]


def _ref_encode_i(sy, scb, scr, qp, ns, rnd, bd):
    """encode_ctu's I branch with intra_trial, driven through the reference's
    compiled functions in its coding order."""
    L = ref()
    choose = L.hevc_choose_luma_mode if bd == 8 else L.hevc_choose_luma_mode_10
    pred4 = L.hevc_predict_4x4 if bd == 8 else L.hevc_predict_4x4_10
    tq = L.hevc_transform_quant_4x4 if bd == 8 else L.hevc_transform_quant_4x4_10
    dq = L.hevc_dequant_itransform_4x4 if bd == 8 else L.hevc_dequant_itransform_4x4_10
    qoff = 0 if bd == 8 else 12
    hc, wc = sy.shape
    rows, cols = hc // 16, wc // 16
    ry, rcb, rcr = np.zeros_like(sy), np.zeros_like(scb), np.zeros_like(scr)
    modes, coef = np.zeros(rows * cols * 16, np.uint8), np.zeros(rows * cols * 384, np.int16)
    vmax = (1 << bd) - 1
    cqp = L.hevc_chroma_qp_from_luma(qp) + qoff
    off = ((0, 0), (8, 0), (0, 8), (8, 8))
    for s in range(ns):
        r0, r1 = rows * s // ns, rows * (s + 1) // ns
        y_min = r0 * 16
        for ty in range(r0, r1):
            for tx in range(cols):
                t = ty * cols + tx
                for cu in range(4):
                    cx, cy = tx * 16 + off[cu][0], ty * 16 + off[cu][1]
                    for pu in range(4):
                        px, py = cx + off[pu][0] // 2, cy + off[pu][1] // 2
                        pred = np.zeros(16, pix(bd))
                        m = choose(y_min, p(sy), p(ry), wc, wc, hc, px, py, p(pred))
                        modes[t * 16 + cu * 4 + pu] = m
                        res = (sy[py:py + 4, px:px + 4].astype(np.int32).reshape(16) - pred).astype(np.int16)
                        lv = np.zeros(16, np.int16)
                        tq(p(res), qp + qoff, 1, rnd, p(lv))
                        coef[t * 384 + (cu * 4 + pu) * 16:][:16] = lv
                        rr = np.zeros(16, np.int16)
                        if lv.any():
                            dq(p(lv), qp + qoff, 1, p(rr))
                        ry[py:py + 4, px:px + 4] = np.clip(pred.astype(int) + rr, 0, vmax).reshape(4, 4)
                    ux, uy = cx // 2, cy // 2
                    for comp, (s_, r_) in enumerate(((scb, rcb), (scr, rcr))):
                        pred = np.zeros(16, pix(bd))
                        pred4(p(r_), wc // 2, wc // 2, hc // 2, ux, uy, 1, 0, y_min // 2, p(pred))
                        res = (s_[uy:uy + 4, ux:ux + 4].astype(np.int32).reshape(16) - pred).astype(np.int16)
                        lv = np.zeros(16, np.int16)
                        tq(p(res), cqp, 0, rnd, p(lv))
                        coef[t * 384 + 256 + comp * 64 + cu * 16:][:16] = lv
                        rr = np.zeros(16, np.int16)
                        if lv.any():
                            dq(p(lv), cqp, 0, p(rr))
                        r_[uy:uy + 4, ux:ux + 4] = np.clip(pred.astype(int) + rr, 0, vmax).reshape(4, 4)
    return dict(modes=modes, coef=coef, y=ry, cb=rcb, cr=rcr)


@need_ref
@pytest.mark.parametrize("bd,qp,ns", [(8, 22, 1), (8, 37, 2), (10, 27, 3)])
def test_encode_i_matches_reference_order(bd, qp, ns):
    rs = np.random.default_rng(qp)
    sy, scb, scr = _frame(rs, 64, 48, bd)
    want = _ref_encode_i(sy, scb, scr, qp, ns, orc.HEVC_ROUND_INTRA_I, bd)
    got = orc.hevc_encode_i(sy, scb, scr, qp, ns, orc.HEVC_ROUND_INTRA_I, bd)
    for k in ("modes", "coef", "y", "cb", "cr"):
        assert np.array_equal(got[k], want[k]), k
    assert len(set(got["modes"].tolist())) > 1  # the decision is exercised


def _spec_min_tb_addr_zs(x, y, width, ctb_log2):
    """6.5.2 (no tiles): MinTbAddrZs of the 4x4 block at sample (x, y):
    CtbAddrRs << (2 (CtbLog2 - 2)) plus, for every bit i of the block's
    coordinates inside the CTB, m*m for an x bit and 2*m*m for a y bit
    (m = 1 << i)."""
    pic_w_ctbs = -(-width // (1 << ctb_log2))
    bx, by = x >> 2, y >> 2
    tb_x, tb_y = (bx << 2) >> ctb_log2, (by << 2) >> ctb_log2
    addr = (pic_w_ctbs * tb_y + tb_x) << ((ctb_log2 - 2) * 2)
    for i in range(ctb_log2 - 2):
        m = 1 << i
        addr += (m * m if m & bx else 0) + (2 * m * m if m & by else 0)
    return addr


@pytest.mark.parametrize("w,h", [(48, 32), (112, 80), (64, 64)])
def test_zscan_rank_is_6_5_2(w, h):
    """The oracle's z-scan rank for CTU16 and CTU32 (luma, and chroma at half
    the CTU size) equals the standard's MinTbAddrZs at every 4x4 block,
    partial CTUs at the right and bottom edges included."""
    for ctu_log2 in (4, 5):
        for luma in (1, 0):
            cw, ch, lg = (w, h, ctu_log2) if luma else (w // 2, h // 2, ctu_log2 - 1)
            for y in range(0, ch, 4):
                for x in range(0, cw, 4):
                    assert orc.hevc_zrank(x, y, cw, luma, ctu_log2) == _spec_min_tb_addr_zs(x, y, cw, lg), \
                        (ctu_log2, luma, x, y)


def test_ctu32_slices_are_ctu32_rows():
    lib = orc.lib
    for rows16 in (1, 2, 5, 8, 9, 135):
        rows32 = (rows16 + 1) // 2
        for ns in (1, 2, 3, 4):
            for r16 in range(rows16):
                starts = [rows32 * s // min(ns, rows32) for s in range(min(ns, rows32))]
                r0 = max(b for b in starts if r16 // 2 >= b)
                assert lib.orc_hevc_slice_y_min_ctu(r16, rows16, ns, 5) == 32 * r0
                assert lib.orc_hevc_slice_y_min_ctu(r16, rows16, ns, 4) == lib.orc_hevc_slice_y_min(r16, rows16, ns)


@pytest.mark.parametrize("w,bd", [(32, 8), (16, 10)])
def test_ctu32_equals_ctu16_when_one_ctu32_wide(w, bd):
    """A picture one CTU32 wide (or half of one) is coded in the same order
    with the same neighbours either way, so CTU32 and CTU16 coding agree;
    a wider picture does not (the bottom-right quadrants lose their
    top-right neighbours)."""
    rs = np.random.default_rng(w)
    sy, scb, scr = _frame(rs, w, 96, bd)
    a = orc.hevc_encode_i(sy, scb, scr, 27, 1, orc.HEVC_ROUND_INTRA_I, bd, 4)
    b = orc.hevc_encode_i(sy, scb, scr, 27, 1, orc.HEVC_ROUND_INTRA_I, bd, 5)
    for k in ("modes", "coef", "y", "cb", "cr"):
        assert np.array_equal(a[k], b[k]), k
    sy, scb, scr = _frame(rs, 96, 64, bd)
    a = orc.hevc_encode_i(sy, scb, scr, 27, 1, orc.HEVC_ROUND_INTRA_I, bd, 4)
    b = orc.hevc_encode_i(sy, scb, scr, 27, 1, orc.HEVC_ROUND_INTRA_I, bd, 5)
    # the first CTU32's first three quadrants see the same neighbours
    assert np.array_equal(a["y"][:16, :32], b["y"][:16, :32]) and np.array_equal(a["y"][16:32, :16], b["y"][16:32, :16])
    assert not np.array_equal(a["y"], b["y"])


@need_ref
@pytest.mark.parametrize("bd,qp", [(8, 27), (10, 32)])
def test_encode_p_matches_reference_order(bd, qp):
    L = ref()
    mc = L.hevc_mc_uni_8 if bd == 8 else L.hevc_mc_uni_10
    tq = L.hevc_transform_quant_4x4 if bd == 8 else L.hevc_transform_quant_4x4_10
    dq = L.hevc_dequant_itransform_4x4 if bd == 8 else L.hevc_dequant_itransform_4x4_10
    rs = np.random.default_rng(5)
    sy, scb, scr = _frame(rs, 64, 48, bd)
    fy, fcb, fcr = _frame(np.random.default_rng(6), 64, 48, bd)
    hc, wc = sy.shape
    t = (wc // 16) * (hc // 16)
    mv = rs.integers(-40, 41, (t * 4, 2)).astype(np.int32)
    got = orc.hevc_encode_p(sy, scb, scr, fy, fcb, fcr, qp, mv, orc.HEVC_ROUND_INTER, bd)
    qoff = 0 if bd == 8 else 12
    cqp = L.hevc_chroma_qp_from_luma(qp) + qoff
    vmax = (1 << bd) - 1
    off = ((0, 0), (8, 0), (0, 8), (8, 8))
    for ti in range(t):
        for cu in range(4):
            cx, cy = (ti % (wc // 16)) * 16 + off[cu][0], (ti // (wc // 16)) * 16 + off[cu][1]
            mx, my = (int(v) for v in mv[ti * 4 + cu])
            pred = np.zeros(64, pix(bd))
            mc(p(fy), wc, wc, hc, cx, cy, 8, 8, mx, my, 0, p(pred), 8)
            res = (sy[cy:cy + 8, cx:cx + 8].astype(np.int32).reshape(64) - pred).astype(np.int16)
            lv = np.zeros(64, np.int16)
            L.hevc_transform_quant_8x8(p(res), qp + qoff, bd, orc.HEVC_ROUND_INTER, p(lv))
            assert np.array_equal(got["coef"][ti * 384 + cu * 64:][:64], lv)
            rr = np.zeros(64, np.int16)
            if lv.any():
                L.hevc_dequant_itransform_8x8(p(lv), qp + qoff, bd, p(rr))
            assert np.array_equal(got["y"][cy:cy + 8, cx:cx + 8].reshape(64), np.clip(pred.astype(int) + rr, 0, vmax))
            for comp, (s_, f_, r_) in enumerate(((scb, fcb, "cb"), (scr, fcr, "cr"))):
                cp = np.zeros(16, pix(bd))
                mc(p(f_), wc // 2, wc // 2, hc // 2, cx // 2, cy // 2, 4, 4, mx, my, 1, p(cp), 4)
                cres = (s_[cy // 2:cy // 2 + 4, cx // 2:cx // 2 + 4].astype(np.int32).reshape(16) - cp).astype(np.int16)
                clv = np.zeros(16, np.int16)
                tq(p(cres), cqp, 0, orc.HEVC_ROUND_INTER, p(clv))
                assert np.array_equal(got["coef"][ti * 384 + 256 + comp * 64 + cu * 16:][:16], clv)
                crr = np.zeros(16, np.int16)
                if clv.any():
                    dq(p(clv), cqp, 0, p(crr))
                assert np.array_equal(got[r_][cy // 2:cy // 2 + 4, cx // 2:cx // 2 + 4].reshape(16),
                                      np.clip(cp.astype(int) + crr, 0, vmax))


@need_ref
@pytest.mark.parametrize("bd,qp", [(8, 22), (10, 32)])
def test_encode_p_ctu32_large_cus_decode_to_recon(bd, qp):
    """CTU32 P coding: the CU choice follows the vectors, CU8 areas equal
    the CTU16 coding, and every CU16 / CU32 reconstructs as the reference
    decoder would from its levels: the reference's MC (hevc_mc_uni) plus
    its inverse transform (hevcd_transform) of the flat-list dequantised
    levels, gathered from the per-area layout."""
    L = ref()
    L.hevcd_transform.argtypes = [P, I, C.c_bool, I]
    mc = L.hevc_mc_uni_8 if bd == 8 else L.hevc_mc_uni_10
    rs = np.random.default_rng(qp)
    w, h = 112, 80                          # 7 x 5 areas: partial CTU32s at the right and bottom
    sy, scb, scr = _frame(rs, w, h, bd)
    fy, fcb, fcr = _frame(np.random.default_rng(qp + 1), w, h, bd)
    wtu, htu = w // 16, h // 16
    t = wtu * htu
    mv = rs.integers(-40, 41, (t * 4, 2)).astype(np.int32)
    for a in (0, 1, wtu, wtu + 1):          # CTU32 (0, 0): one vector -> CU32
        mv[a * 4:a * 4 + 4] = (13, -6)
    mv[2 * 4:2 * 4 + 4] = (5, 7)            # CTU32 (1, 0): one CU16 (area 2), the rest CU8
    mv[6 * 4:6 * 4 + 4] = (-9, 3)           # area 6: right edge, CTU32 (3, 0) partial -> CU16
    mv[(4 * wtu + 2) * 4:(4 * wtu + 2) * 4 + 4] = (1, 1)   # bottom row (partial CTU32) -> CU16
    want_lg = {0: 5, 1: 5, wtu: 5, wtu + 1: 5, 2: 4, 6: 4, 4 * wtu + 2: 4}
    got = orc.hevc_encode_p(sy, scb, scr, fy, fcb, fcr, qp, mv, orc.HEVC_ROUND_INTER, bd, 5)
    base = orc.hevc_encode_p(sy, scb, scr, fy, fcb, fcr, qp, mv, orc.HEVC_ROUND_INTER, bd, 4)
    for a in range(t):
        assert (got["culog2"][a * 4:a * 4 + 4] == want_lg.get(a, 3)).all(), a
        if a not in want_lg:
            ax, ay = (a % wtu) * 16, (a // wtu) * 16
            assert np.array_equal(got["coef"][a * 384:(a + 1) * 384], base["coef"][a * 384:(a + 1) * 384])
            assert np.array_equal(got["y"][ay:ay + 16, ax:ax + 16], base["y"][ay:ay + 16, ax:ax + 16])
    qoff = 0 if bd == 8 else 12
    cqp = L.hevc_chroma_qp_from_luma(qp) + qoff
    vmax = (1 << bd) - 1
    nonzero = 0
    for (x0, y0, log2n, areas) in ((0, 0, 5, (0, 1, wtu, wtu + 1)), (32, 0, 4, (2,)), (96, 0, 4, (6,)),
                                   (32, 64, 4, (4 * wtu + 2,))):
        mx, my = (int(v) for v in mv[areas[0] * 4])
        for comp, (plane, ref_p, key) in enumerate(((sy, fy, "y"), (scb, fcb, "cb"), (scr, fcr, "cr"))):
            lg = log2n if comp == 0 else log2n - 1
            n, chroma = 1 << lg, comp > 0
            bx, by = (x0, y0) if not chroma else (x0 // 2, y0 // 2)
            pw, ph = (w, h) if not chroma else (w // 2, h // 2)
            pred = np.zeros(n * n, pix(bd))
            mc(p(np.ascontiguousarray(ref_p)), pw, pw, ph, bx, by, n, n, mx, my, int(chroma), p(pred), n)
            per, slot = (256, 0) if comp == 0 else (64, 256 + (comp - 1) * 64)
            per = per if log2n == 5 else n * n
            lv = np.concatenate([got["coef"][a * 384 + slot:][:per] for a in areas]).astype(np.int16)
            d = np.zeros(n * n, np.int16)
            orc.lib.orc_hevc_dequant(p(lv), n * n, cqp if chroma else qp + qoff, bd + lg - 5, p(d))
            L.hevcd_transform(p(d), lg, False, bd)
            rec = np.clip(pred.astype(int) + d, 0, vmax).reshape(n, n)
            assert np.array_equal(got[key][by:by + n, bx:bx + n], rec), (x0, y0, key)
            nonzero += int(lv.any())
    assert nonzero >= 4  # levels are exercised


@need_ref
@pytest.mark.parametrize("bd", [8, 10])
def test_refine8_rings_brute_force(bd):
    """The HEVC 8x8 refinement by brute force: the integer start, then the
    8 half-sample candidates (2,0) (-2,0) (0,2) (0,-2) (2,2) (2,-2) (-2,2)
    (-2,-2) around it, then the same offsets halved around the half winner,
    each taken only when its SAD against the reference's own uni-prediction
    (hevc_mc_uni_8 / _10) is strictly lower.  Two-level and flat content
    make ties frequent."""
    L = ref()
    f = L.hevc_mc_uni_8 if bd == 8 else L.hevc_mc_uni_10
    rs = np.random.default_rng(40 + bd)
    h, w = 40, 48
    yy, xx = np.mgrid[0:h, 0:w]
    sc = 1 << (bd - 8)
    planes = [(100 * sc + rs.integers(0, 2, (h, w))).astype(pix(bd)),
              np.clip(128 * sc + 50 * sc * np.sin(xx / 2.7) * np.cos(yy / 3.1), 0, 255 * sc).astype(pix(bd)),
              np.full((h, w), 77 * sc, pix(bd))]
    ring = [(2, 0), (-2, 0), (0, 2), (0, -2), (2, 2), (2, -2), (-2, 2), (-2, -2)]

    def sad(cur, refp, x0, y0, mx, my):
        pr = np.zeros((8, 8), pix(bd))
        f(p(refp), w, w, h, x0, y0, 8, 8, mx, my, 0, p(pr), 8)
        return int(np.abs(pr.astype(int) - cur[y0:y0 + 8, x0:x0 + 8]).sum())

    # p(x, y) = g(x + y): a half-sample move in x predicts exactly what one in
    # y does, so (2,0) and (0,2) tie and the ring order decides
    diag = np.clip(128 * sc + 60 * sc * np.sin((xx + yy) / 2.3), 0, 255 * sc).astype(pix(bd))
    half = np.zeros((h, w), pix(bd))
    f(p(diag), w, w, h, 0, 0, w, h, 2, 0, 0, p(half), w)
    moved = 0
    pairs = [(np.ascontiguousarray(np.roll(r, (1, -1), axis=(0, 1))), r) for r in planes] + [(half, diag)]
    for cur, refp in pairs:
        for (x0, y0) in ((0, 0), (8, 16), (24, 8), (40, 32)):
            for (dx, dy) in ((0, 0), (-1, 0), (1, -1)):
                cx, cy = 4 * dx, 4 * dy
                best = sad(cur, refp, x0, y0, cx, cy)
                for div in (1, 2):
                    bx = by = 0
                    for ox, oy in ring:
                        s_ = sad(cur, refp, x0, y0, cx + ox // div, cy + oy // div)
                        if s_ < best:
                            best, bx, by = s_, ox // div, oy // div
                    cx, cy = cx + bx, cy + by
                got = orc.hevc_refine8(cur, refp, x0, y0, 4 * dx, 4 * dy, bd)
                assert got == (cx, cy, best), (bd, x0, y0, dx, dy)
                moved += (cx, cy) != (4 * dx, 4 * dy)
    assert moved > 0


def test_refine8_never_worse_and_consistent():
    rs = np.random.default_rng(12)
    h, w = 48, 64
    ref_ = np.clip(128 + 50 * np.sin(np.arange(w)[None, :] / 3.0 + np.arange(h)[:, None] / 4.0), 0, 255).astype(np.uint8)
    cur = np.roll(ref_, (1, 2), axis=(0, 1))
    for (x0, y0) in ((8, 8), (24, 16), (40, 32)):
        for (dx, dy) in ((0, 0), (-2, -1), (1, 1)):
            mx, my, sad = orc.hevc_refine8(cur, ref_, x0, y0, dx * 4, dy * 4)
            start = np.abs(cur[y0:y0 + 8, x0:x0 + 8].astype(int) -
                           orc.hevc_mc_uni(ref_, x0, y0, 8, 8, dx * 4, dy * 4, 0).astype(int)).sum()
            assert sad <= start
            again = np.abs(cur[y0:y0 + 8, x0:x0 + 8].astype(int) -
                           orc.hevc_mc_uni(ref_, x0, y0, 8, 8, mx, my, 0).astype(int)).sum()
            assert again == sad
            assert abs(mx - dx * 4) <= 3 and abs(my - dy * 4) <= 3


# ------------------------------------------------------------ 16x16 / 32x32

@need_ref
def test_dct_matrices_match_reference_table():
    """The N-point matrices generated from the standard's 32-point magnitudes
    equal the reference decoder's hevcd_dct rows k*32/N (and the 4/8-point
    ones the encoder uses)."""
    table = np.frombuffer((C.c_int8 * 1024).in_dll(ref(), "hevcd_dct"), np.int8).reshape(32, 32).astype(int)
    for n in (4, 8, 16, 32):
        m = orc.hevc_dct_matrix(n)
        assert np.array_equal(m, table[::32 // n, :n]), n
    assert np.array_equal(orc.hevc_dct_matrix(4), [[64, 64, 64, 64], [83, 36, -36, -83], [64, -64, -64, 64],
                                                   [36, -83, 83, -36]])
    # and they are the DCT-II basis: within 2 of 64*sqrt(2)*cos(pi*(2n+1)k/(2N))
    for n in (16, 32):
        k, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        ideal = np.round(64 * np.sqrt(2) * np.cos(np.pi * (2 * x + 1) * k / (2 * n)))
        ideal[0] = 64
        assert np.abs(orc.hevc_dct_matrix(n) - ideal).max() <= 2


@need_ref
@pytest.mark.parametrize("bd", [8, 10])
@pytest.mark.parametrize("log2n", [4, 5])
def test_inverse_nxn_matches_reference_decoder(bd, log2n):
    L = ref()
    L.hevcd_transform.argtypes = [P, I, C.c_bool, I]
    n2 = 1 << (2 * log2n)
    rs = np.random.default_rng(log2n * 10 + bd)
    for qp in (4, 22, 27, 37, 51, 60):
        for dens in (0.05, 0.3, 1.0):
            lv = (rs.integers(-60, 61, n2) * (rs.random(n2) < dens)).astype(np.int16)
            lv[rs.integers(0, n2)] = 0
            d = np.zeros(n2, np.int16)
            orc.lib.orc_hevc_dequant(p(lv), n2, qp, bd + log2n - 5, p(d))
            want = d.copy()
            L.hevcd_transform(p(want), log2n, False, bd)
            assert np.array_equal(orc.hevc_dqitn(lv, log2n, qp, bd), want), (qp, dens)


@pytest.mark.parametrize("log2n", [2, 3, 4, 5])
def test_forward_nxn_round_trips(log2n):
    """DC-only residuals come back exactly; textured residuals within a
    quantisation step's error at a fine QP; the 4x4 and 8x8 forms equal the
    dedicated ones the reference's encoder uses."""
    n2 = 1 << (2 * log2n)
    rs = np.random.default_rng(log2n)
    for v in (-200, -3, 1, 77, 255):
        res = np.full(n2, v, np.int16)
        lv = orc.hevc_tqn(res, log2n, 4, 8, 6)
        assert np.count_nonzero(lv[1:]) == 0
        assert np.array_equal(orc.hevc_dqitn(lv, log2n, 4, 8), res)
    # the integer bases are near-orthogonal, not exact: an impulse through the
    # 32-point pair comes back 2% high, so the bound grows with the size
    tol = {2: 3, 3: 3, 4: 3, 5: 6}[log2n]
    for _ in range(10):
        res = rs.integers(-255, 256, n2).astype(np.int16)
        back = orc.hevc_dqitn(orc.hevc_tqn(res, log2n, 0, 8, 6), log2n, 0, 8)
        assert np.abs(back.astype(int) - res).max() <= tol
        if log2n == 3:
            assert np.array_equal(orc.hevc_tqn(res, 3, 27, 8, 2), orc.hevc_tq8(res, 27, 8, 2))
            assert np.array_equal(orc.hevc_dqitn(orc.hevc_tq8(res, 27, 8, 2), 3, 27, 8),
                                  orc.hevc_dqit8(orc.hevc_tq8(res, 27, 8, 2), 27, 8))
        if log2n == 2:
            r4 = np.clip(res, -255, 255)
            assert np.array_equal(orc.hevc_tqn(r4, 2, 27, 8, 4), orc.hevc_tq4(r4, 27, 0, 4, 8))


# ------------------------------------------------------------ deblocking

@need_ref
def test_deblock_tables_match_reference():
    beta = np.frombuffer((C.c_uint8 * 52).in_dll(ref(), "hevcd_beta"), np.uint8)
    tc = np.frombuffer((C.c_uint8 * 54).in_dll(ref(), "hevcd_tc"), np.uint8)
    assert [orc.lib.orc_hevc_beta(q) for q in range(52)] == beta.tolist()
    assert [orc.lib.orc_hevc_tc(q) for q in range(54)] == tc.tolist()


def _blocky_picture(rs, h, w, step, noise):
    yy, xx = np.mgrid[0:h, 0:w]
    off = rs.integers(-step, step + 1, (h // 8 + 1, w // 8 + 1))
    v = 90 + (xx * 50) // w + (yy * 40) // h + off[yy // 8, xx // 8] + rs.integers(-noise, noise + 1, (h, w))
    return np.clip(v, 0, 255).astype(np.uint8)


@pytest.mark.skipif(not os.path.exists(orc.REF_HEVC_DEBLOCK), reason="oracle/_ref not built (reference tree absent)")
def test_deblock_matches_reference_decoder():
    """The oracle's picture deblocking equals the reference decoder's own
    loop filter (hevcd_filter_stage, hevc_filter.c:507, run by the harness
    oracle/ref_hevc_deblock.c) on seeded blocky pictures: 8 and 10 bits, I and
    P, every QP band, coded-luma flags and vectors that give bS 0/1/2, CTU-row
    slices."""
    lib = C.CDLL(orc.REF_HEVC_DEBLOCK, mode=os.RTLD_LAZY)
    lib.ref_hevc_deblock.argtypes = [P, P, I, I, I, I, P, P, I, I]
    lib.ref_hevc_deblock.restype = I
    rs = np.random.default_rng(11)
    cases = 0
    changed = 0
    for (w, h, bd) in ((64, 48, 8), (96, 64, 8), (160, 96, 8), (96, 64, 10), (128, 80, 10)):
        for qp in (17, 22, 27, 30, 34, 37, 41, 45, 51):
            for intra in (0, 1):
                for ns in (1, 2, 3):
                    step = int(rs.integers(2, 20))
                    y = _blocky_picture(rs, h, w, step, int(rs.integers(0, 3)))
                    uv = _blocky_picture(rs, h // 2, w, step // 2 + 1, 1)
                    if bd == 10:   # 10-bit pictures: 4x the 8-bit one plus a dither
                        y = y.astype(np.uint16) * 4 + rs.integers(0, 4, y.shape).astype(np.uint16)
                        uv = uv.astype(np.uint16) * 4 + rs.integers(0, 4, uv.shape).astype(np.uint16)
                    ncu = (w // 8) * (h // 8)
                    cbf = (rs.random(ncu) < 0.4).astype(np.uint8)
                    mv4 = np.zeros((ncu, 4), np.int32)
                    mv4[:, :2] = rs.integers(-8, 9, (ncu, 2))
                    mv4[rs.random(ncu) < 0.3, :2] = 0
                    wy, wuv = orc.hevc_deblock(y, uv, qp, intra, cbf, mv4, ns, bd)
                    ry, ruv = y.copy(), uv.copy()
                    assert lib.ref_hevc_deblock(ry.ctypes.data, ruv.ctypes.data, w, h, qp, intra,
                                                cbf.ctypes.data, mv4.ctypes.data, ns, bd) == 0
                    assert np.array_equal(wy, ry), (w, h, bd, qp, intra, ns)
                    assert np.array_equal(wuv, ruv), (w, h, bd, qp, intra, ns)
                    cases += 1
                    changed += int(not np.array_equal(wy, y)) + int(not np.array_equal(wuv, uv))
    assert cases == 270 and changed > 250


@pytest.mark.skipif(not os.path.exists(orc.REF_HEVC_DEBLOCK), reason="oracle/_ref not built (reference tree absent)")
def test_deblock_ctu32_large_cus_matches_reference_decoder():
    """CTU32 pictures through the reference decoder's loop filter: CTB 32
    (slices as CTU32-row ranges) and, for P pictures, CU16 / CU32 whose inner
    8x8 cells carry no edge bits; the oracle's CU-size map must skip exactly
    those edges.  CU sizes drawn per CTU32 (partial CTU32s at the edges keep
    CU16 at most), cbf per transform block."""
    lib = C.CDLL(orc.REF_HEVC_DEBLOCK, mode=os.RTLD_LAZY)
    lib.ref_hevc_deblock_ctu.argtypes = [P, P, I, I, I, I, P, P, I, I, I, P]
    lib.ref_hevc_deblock_ctu.restype = I
    rs = np.random.default_rng(12)
    cases = 0
    for (w, h, bd) in ((96, 64, 8), (112, 80, 8), (160, 96, 10)):
        wtu, htu = w // 16, h // 16
        for qp in (22, 30, 37, 45):
            for intra in (0, 1):
                for ns in (1, 2, 3):
                    step = int(rs.integers(2, 20))
                    y = _blocky_picture(rs, h, w, step, int(rs.integers(0, 3)))
                    uv = _blocky_picture(rs, h // 2, w, step // 2 + 1, 1)
                    if bd == 10:
                        y = y.astype(np.uint16) * 4 + rs.integers(0, 4, y.shape).astype(np.uint16)
                        uv = uv.astype(np.uint16) * 4 + rs.integers(0, 4, uv.shape).astype(np.uint16)
                    ncu = wtu * htu * 4
                    culog2 = np.full(ncu, 3, np.uint8)
                    if not intra:
                        for Y in range((htu + 1) // 2):
                            for X in range((wtu + 1) // 2):
                                areas = [(2 * Y + (a >> 1)) * wtu + 2 * X + (a & 1) for a in range(4)
                                         if 2 * X + (a & 1) < wtu and 2 * Y + (a >> 1) < htu]
                                kind = int(rs.integers(0, 3))
                                if kind == 2 and len(areas) == 4:
                                    for t in areas:
                                        culog2[t * 4:t * 4 + 4] = 5
                                elif kind >= 1:
                                    for t in areas:
                                        if rs.random() < 0.6:
                                            culog2[t * 4:t * 4 + 4] = 4
                    mv4 = np.zeros((ncu, 4), np.int32)
                    mv4[:, :2] = rs.integers(-8, 9, (ncu, 2))
                    for cu in range(ncu):   # one vector per CU
                        L = int(culog2[cu])
                        if L >= 4:
                            t = cu // 4
                            t0 = t if L == 4 else ((t // wtu) & ~1) * wtu + ((t % wtu) & ~1)
                            mv4[cu] = mv4[t0 * 4]
                    cbf = (rs.random(ncu) < 0.4).astype(np.uint8)
                    for cu in range(ncu):
                        L = int(culog2[cu])
                        if L >= 4:
                            t = cu // 4
                            t0 = t if L == 4 else ((t // wtu) & ~1) * wtu + ((t % wtu) & ~1)
                            cbf[cu] = cbf[t0 * 4]
                    wy, wuv = orc.hevc_deblock(y, uv, qp, intra, cbf, mv4, ns, bd, 5, None if intra else culog2)
                    ry, ruv = y.copy(), uv.copy()
                    assert lib.ref_hevc_deblock_ctu(ry.ctypes.data, ruv.ctypes.data, w, h, qp, intra,
                                                    cbf.ctypes.data, mv4.ctypes.data, ns, bd, 5,
                                                    None if intra else culog2.ctypes.data) == 0
                    assert np.array_equal(wy, ry), (w, h, bd, qp, intra, ns)
                    assert np.array_equal(wuv, ruv), (w, h, bd, qp, intra, ns)
                    cases += 1
    assert cases == 72


def _two_columns(left, right, w=32, h=16):
    y = np.full((h, w), right, np.uint8)
    y[:, :8] = left
    uv = np.full((h // 2, w), 128, np.uint8)
    return y, uv


def test_deblock_line_kats():
    """Hand-worked 8.7.2.5 cases on the vertical edge x = 8 of a 32x16
    picture whose other edges are flat (so nothing else moves)."""
    ncu = 2 * 4
    mv = np.zeros((ncu, 4), np.int32)
    # normal filter, bS 1 (coded luma), QP 37: beta 36, tc 4; p = 60 flat, q = 70 flat:
    # delta = (9*10 - 3*10 + 8) >> 4 = 4; p1/q1 move by (+-4 >> 1) = +-2 (|| <= tc/2)
    y, uv = _two_columns(60, 70)
    cbf = np.ones(ncu, np.uint8)
    oy, _ = orc.hevc_deblock(y, uv, 37, 0, cbf, mv)
    assert oy[:, 4:12].tolist() == [[60, 60, 62, 64, 66, 68, 70, 70]] * 16
    # strong filter, bS 2 (intra), QP 40: beta 42, tc 7; p = 50, q = 56:
    # p0' = (50+100+100+112+56+4)>>3 = 52, p1' = 208>>2 = 52, p2' = 410>>3 = 51,
    # q0' = 434>>3 = 54, q1' = 220>>2 = 55, q2' = 446>>3 = 55
    y, uv = _two_columns(50, 56)
    oy, ouv = orc.hevc_deblock(y, uv, 40, 1, np.zeros(ncu, np.uint8), mv)
    assert oy[:, 4:12].tolist() == [[50, 51, 52, 52, 54, 55, 55, 56]] * 16
    # chroma, bS 2, QP 30: QpC 29, tc = tc'[31] = 3; p = 40, q = 60 on the chroma
    # edge at chroma x = 8 (luma 16): delta = clip(+-3, (80 - 20 + 4) >> 3) = 3
    y = np.full((16, 32), 90, np.uint8)
    uv = np.full((8, 32), 60, np.uint8)
    uv[:, :16] = 40                     # chroma columns 0..7 (both components) = 40
    _, ouv = orc.hevc_deblock(y, uv, 30, 1, np.zeros(ncu, np.uint8), mv)
    assert ouv[:, 12:20].tolist() == [[40, 40, 43, 43, 57, 57, 60, 60]] * 8


def test_deblock_invariants():
    rs = np.random.default_rng(3)
    h, w = 48, 64
    ncu = (w // 8) * (h // 8)
    flat = np.full((h, w), 99, np.uint8), np.full((h // 2, w), 120, np.uint8)
    oy, ouv = orc.hevc_deblock(*flat, 30, 1, np.ones(ncu, np.uint8), np.zeros((ncu, 4), np.int32))
    assert (oy == 99).all() and (ouv == 120).all()
    # blocky but smooth inside the blocks: a ramp plus a random offset per 8x8
    # (luma) / 4x4 (chroma) block, so the edges carry steps the filter acts on
    yy, xx = np.mgrid[0:h, 0:w]
    y = (100 + xx // 2 + yy // 3 + np.kron(rs.integers(-9, 10, (h // 8, w // 8)), np.ones((8, 8), int)))
    y = y.astype(np.uint8)
    uv = (128 + np.kron(rs.integers(-6, 7, (h // 8, w // 8)), np.ones((4, 8), int))).astype(np.uint8)
    oy, ouv = orc.hevc_deblock(y, uv, 30, 0, np.zeros(ncu, np.uint8), np.zeros((ncu, 4), np.int32))
    assert np.array_equal(oy, y) and np.array_equal(ouv, uv)  # bS 0 everywhere
    # Symmetry.  The strong luma filter is symmetric in p and q, so a
    # mirrored picture gives the mirrored result.  The normal luma and the
    # chroma filters round with "+c >> s", which is not antisymmetric; they
    # are symmetric under mirror plus sample negation (v -> 255 - v).  Each
    # picture below forces one kind of filter on every edge.
    def mirror_lr(a, b):
        bb = b.reshape(b.shape[0], -1, 2)[:, ::-1, :].reshape(b.shape)
        return np.ascontiguousarray(a[:, ::-1]), np.ascontiguousarray(bb)

    def check(y, uv, qp, intra, cbf, negate):
        mv = np.zeros((ncu, 4), np.int32)
        oy, ouv = orc.hevc_deblock(y, uv, qp, intra, cbf, mv)
        assert not np.array_equal(oy, y)
        f = (lambda a: 255 - a) if negate else (lambda a: a)
        for flip in ("lr", "ud"):
            if flip == "lr":
                my, muv = mirror_lr(f(y), f(uv))
                ey, euv = mirror_lr(f(oy), f(ouv))
            else:
                my, muv = np.ascontiguousarray(f(y)[::-1]), np.ascontiguousarray(f(uv)[::-1])
                ey, euv = f(oy)[::-1], f(ouv)[::-1]
            gy, guv = orc.hevc_deblock(my, muv, qp, intra, cbf, mv)
            assert np.array_equal(gy, ey), flip
            if negate:
                assert np.array_equal(guv, euv), flip
        return oy, ouv

    # strong only: gentle ramp, steps of at most 4 between blocks, intra QP 40
    small = np.kron(rs.integers(-2, 3, (h // 8, w // 8)), np.ones((8, 8), int))
    ys = (100 + xx // 4 + yy // 4 + small).astype(np.uint8)
    check(ys, uv, 40, 1, np.zeros(ncu, np.uint8), negate=False)
    # normal only: checkerboard of +-15 per block (every step 30 >= the strong
    # limit (5*tc+1)/2 = 10 at tc 4), coded luma (bS 1), QP 37; chroma edges
    # need bS 2, so the chroma check is an intra picture with the same steps
    board = 15 * (2 * ((np.arange(h // 8)[:, None] + np.arange(w // 8)[None, :]) % 2) - 1)
    yn = (120 + xx // 4 + np.kron(board, np.ones((8, 8), int))).astype(np.uint8)
    check(yn, uv, 37, 0, np.ones(ncu, np.uint8), negate=True)
    cb = (128 + np.kron(board // 3, np.ones((4, 8), int))).astype(np.uint8)
    _, ouv = check(yn, cb, 37, 1, np.zeros(ncu, np.uint8), negate=True)
    assert not np.array_equal(ouv, cb)
