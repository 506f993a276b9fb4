"""ctypes binding of the CPU oracle (oracle/liborc.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
this module; the product path never does.  Arrays follow the coded-size
layout documented in orc.h.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liborc.so")
REF_H264 = os.path.join(_HERE, "_ref", "libref_h264.so")

MV = np.dtype([("mvx", "<i4"), ("mvy", "<i4"), ("sad", "<u4"), ("pad", "<u4")])
IMV = np.dtype([("dx", "<i4"), ("dy", "<i4"), ("cost", "<u4"), ("sad", "<u4")])

if not os.path.exists(LIB):
    raise RuntimeError(f"{LIB} missing: run `make -C {_HERE}`")
lib = C.CDLL(LIB)
_P = C.c_void_p
_I = C.c_int


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


_sig("orc_chroma_qp", _I, [_I])
_sig("orc_fwd4x4", None, [_P, _P])
_sig("orc_inv4x4", None, [_P, _P])
_sig("orc_quant_ac", _I, [_I, _I, _I, _I])
_sig("orc_dequant_ac", _I, [_I, _I, _I])
_sig("orc_quant_dc_luma", _I, [_I, _I])
_sig("orc_quant_dc_chroma", _I, [_I, _I])
_sig("orc_dequant_dc_luma", _I, [_I, _I])
_sig("orc_dequant_dc_chroma", _I, [_I, _I])
_sig("orc_luma_dc_hadamard_fwd", None, [_P, _P])
_sig("orc_luma_dc_hadamard_inv", None, [_P, _P])
_sig("orc_chroma_dc_hadamard", None, [_P, _P])
_sig("orc_intra16_luma_dc_levels", None, [_P, _I, _P])
_sig("orc_luma_qpel", _I, [_P, _I, _I, _I, _I, _I])
_sig("orc_chroma_8th", _I, [_P, _I, _I, _I, _I, _I, _I])
_sig("orc_pred_luma_block", None, [_P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P, _I])
_sig("orc_me_fullsearch", None, [_P, _P, _I, _I, _I, _I, _I, _P, _P])
_sig("orc_me_fullsearch_mb", None, [_P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P])
_sig("orc_me_refine", None, [_P, _P, _I, _I, _I, _P, _P, _P, _P])
_sig("orc_me_refine_mb", None, [_P, _P, _I, _I, _I, _I, _I, _P, _P, _P, _P])
_sig("orc_h264_encode_p", None, [_P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P])
_sig("orc_h264_encode_i", None, [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P])
_sig("orc_slice_of", _I, [_I, _I, _I])
_sig("orc_h264_deblock", None, [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _I, _I])
_sig("orc_h264_filter_line_luma", None, [_P, _I, _I, _I])
_sig("orc_h264_filter_line_chroma", None, [_P, _I, _I, _I])
_sig("orc_h264_alpha", _I, [_I])
_sig("orc_h264_beta", _I, [_I])
_sig("orc_h264_tc0", _I, [_I, _I])
_sig("orc_h264_encode_frame", None, [_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I,
                                     _P, _P, _P, _P, _P, _P, _P, _P])


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def fwd4x4(blk):
    i, o = c32(blk).reshape(16), np.zeros(16, np.int32)
    lib.orc_fwd4x4(p(i), p(o))
    return o


def inv4x4(blk):
    i, o = c32(blk).reshape(16), np.zeros(16, np.int32)
    lib.orc_inv4x4(p(i), p(o))
    return o


def luma_qpel(ref, xq, yq):
    ref = np.ascontiguousarray(ref, np.uint8)
    return lib.orc_luma_qpel(p(ref), ref.strides[0], ref.shape[1], ref.shape[0], xq, yq)


def me_fullsearch(cur, ref, rng, lambda8):
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    hc, wc = cur.shape
    nmb = (wc // 16) * (hc // 16)
    i16, i8 = np.zeros(nmb, IMV), np.zeros(nmb * 4, IMV)
    lib.orc_me_fullsearch(p(cur), p(ref), cur.strides[0], wc, hc, rng, lambda8, p(i16), p(i8))
    return i16, i8


_sig("orc_square_me", None, [_P, _I, _P, _I, _I, _I, _I, _P])


def square_me(cur, ref, width, height, search_radius):
    """The reference's Tier-2 square search over MB-padded planes; returns MV [nmb]."""
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    nmb = ((width + 15) // 16) * ((height + 15) // 16)
    out = np.zeros(nmb, MV)
    lib.orc_square_me(p(cur), cur.strides[0], p(ref), ref.strides[0], width, height, search_radius, p(out))
    return out


_sig("orc_diamond_me", None, [_P, _P, _I, _I, _I, _I, _P])


def diamond_me(cur, ref, lam=5):
    """The reference's diamond search with integer SADs over coded-size planes; MV [nmb]."""
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    hc, wc = cur.shape
    out = np.zeros((wc // 16) * (hc // 16), MV)
    lib.orc_diamond_me(p(cur), p(ref), wc, wc, hc, lam, p(out))
    return out


def me_fullsearch_mb(cur, ref, rng, lambda8, mbx, mby):
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    hc, wc = cur.shape
    i16, i8 = np.zeros(1, IMV), np.zeros(4, IMV)
    lib.orc_me_fullsearch_mb(p(cur), p(ref), cur.strides[0], wc, hc, rng, lambda8, mbx, mby, p(i16), p(i8))
    return i16[0], i8


def me_refine(cur, ref, i16, i8):
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    hc, wc = cur.shape
    nmb = (wc // 16) * (hc // 16)
    m16, m8 = np.zeros(nmb, MV), np.zeros(nmb * 4, MV)
    lib.orc_me_refine(p(cur), p(ref), cur.strides[0], wc, hc, p(np.ascontiguousarray(i16)),
                      p(np.ascontiguousarray(i8)), p(m16), p(m8))
    return m16, m8


def me_refine_mb(cur, ref, mbx, mby, i16, i8):
    """me_refine of one MB: i16 one IMV, i8 its four 8x8 IMVs -> (MV, MV[4])."""
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    hc, wc = cur.shape
    a = np.array([i16], IMV)
    b = np.ascontiguousarray(i8, IMV)
    m16, m8 = np.zeros(1, MV), np.zeros(4, MV)
    lib.orc_me_refine_mb(p(cur), p(ref), cur.strides[0], wc, hc, mbx, mby, p(a), p(b), p(m16), p(m8))
    return m16[0], m8


def _outs(wc, hc):
    nmb = (wc // 16) * (hc // 16)
    return dict(levels=np.zeros(nmb * 384, np.int16), dc=np.zeros(nmb * 24, np.int32),
                nz=np.zeros(nmb * 24, np.uint32), modes=np.zeros(nmb, np.uint32),
                y=np.zeros((hc, wc), np.uint8), uv=np.zeros((hc // 2, wc), np.uint8))


def encode_p(cy, cuv, ry, ruv, qp, mvs):
    hc, wc = cy.shape
    o = _outs(wc, hc)
    args = [np.ascontiguousarray(a, np.uint8) for a in (cy, cuv, ry, ruv)]
    mvs = np.ascontiguousarray(mvs, MV)
    lib.orc_h264_encode_p(p(args[0]), p(args[1]), p(args[2]), p(args[3]), args[0].strides[0],
                          args[1].strides[0], wc, hc, qp, p(mvs), p(o["levels"]), p(o["dc"]), p(o["nz"]),
                          p(o["modes"]), p(o["y"]), p(o["uv"]))
    return o


def encode_i(cy, cuv, qp, num_slices):
    hc, wc = cy.shape
    o = _outs(wc, hc)
    cy = np.ascontiguousarray(cy, np.uint8)
    cuv = np.ascontiguousarray(cuv, np.uint8)
    lib.orc_h264_encode_i(p(cy), p(cuv), cy.strides[0], cuv.strides[0], wc, hc, qp, num_slices,
                          p(o["levels"]), p(o["dc"]), p(o["nz"]), p(o["modes"]), p(o["y"]), p(o["uv"]))
    return o


def deblock(y, uv, qp, is_intra, nz, mvs, num_slices, idc):
    y = np.array(y, np.uint8, copy=True)
    uv = np.array(uv, np.uint8, copy=True)
    hc, wc = y.shape
    nz = np.ascontiguousarray(nz, np.uint32)
    mvs = np.ascontiguousarray(mvs, MV)
    lib.orc_h264_deblock(p(y), p(uv), y.strides[0], uv.strides[0], wc, hc, qp, int(is_intra), p(nz), p(mvs),
                         num_slices, idc)
    return y, uv


def encode_frame(cy, cuv, ry, ruv, qp, is_intra, num_slices=1, rng=16, lambda8=1, idc=0):
    """All stages of one frame in the pipeline's order; returns a dict."""
    hc, wc = cy.shape
    o = _outs(wc, hc)
    nmb = (wc // 16) * (hc // 16)
    o["mvs"] = np.zeros(nmb, MV)
    o["mv8"] = np.zeros(nmb * 4, MV)
    a = [np.ascontiguousarray(x, np.uint8) for x in (cy, cuv, ry, ruv)]
    lib.orc_h264_encode_frame(p(a[0]), p(a[1]), p(a[2]), p(a[3]), a[0].strides[0], a[1].strides[0], wc, hc, qp,
                              int(is_intra), num_slices, rng, lambda8, idc, p(o["mvs"]), p(o["mv8"]),
                              p(o["levels"]), p(o["dc"]), p(o["nz"]), p(o["modes"]), p(o["y"]), p(o["uv"]))
    return o


# ------------------------------------------------------------ H.264 CAVLC
REF_CAVLC = os.path.join(_HERE, "_ref", "libref_cavlc.so")
_L64 = C.c_int64
_sig("orc_cavlc_block", _I, [_P, _I, _I, _P, _I, _P])
_sig("orc_cavlc_chroma_dc", _I, [_P, _P, _I, _P])
_sig("orc_exp_golomb", _I, [_I, _I, _P, _I, _P])
_sig("orc_h264_mb_header", _I, [_I, _I, _I, _I, _I, _P, _I, _P])
_sig("orc_h264_mv_pred", None, [_P, _I, _I, _I, _P, _P])
_sig("orc_h264_skip_mv", None, [_P, _I, _I, _I, _P, _P])
_sig("orc_h264_cavlc_slices", _L64, [_P, _P, _P, _P, _P, _P, _L64, _P, _P])

# slice configuration, in the order of bc250_h264_slice_cfg_t
SLICE_CFG = ("width_in_mbs", "height_in_mbs", "qp", "is_idr", "num_slices", "pps_id", "pic_init_qp_minus26",
             "frame_num", "log2_max_frame_num_minus4", "idr_pic_id", "poc", "log2_max_poc_lsb_minus4",
             "deblock_idc", "relax_chroma_skip")


def slice_cfg(wmb, hmb, qp, is_idr, num_slices=1, pps_id=0, pic_init_qp_minus26=0, frame_num=0,
              log2_max_frame_num_minus4=0, idr_pic_id=0, poc=0, log2_max_poc_lsb_minus4=0, deblock_idc=0,
              relax_chroma_skip=0):
    """The slice configuration as the int32 array both writers take."""
    v = dict(locals())
    v["width_in_mbs"], v["height_in_mbs"] = wmb, hmb
    return np.array([int(v[k]) for k in SLICE_CFG], np.int32)


def _bits(buf, n):
    return np.unpackbits(np.frombuffer(bytes(buf), np.uint8))[:n]


def cavlc_block(coef, ac, nc):
    """One 4x4 residual block (raster coefficients, ac = 1 skips the DC
    position) -> (bit array, TotalCoeff)."""
    c = np.ascontiguousarray(coef, np.int16).reshape(16)
    buf, nb = np.zeros(256, np.uint8), C.c_int(0)
    tc = lib.orc_cavlc_block(p(c), int(ac), int(nc), p(buf), 256, C.byref(nb))
    return _bits(buf, nb.value), tc


def cavlc_chroma_dc(c4):
    c = c32(c4).reshape(4)
    buf, nb = np.zeros(64, np.uint8), C.c_int(0)
    tc = lib.orc_cavlc_chroma_dc(p(c), p(buf), 64, C.byref(nb))
    return _bits(buf, nb.value), tc


def exp_golomb(value, signed):
    buf, nb = np.zeros(16, np.uint8), C.c_int(0)
    lib.orc_exp_golomb(int(value), int(signed), p(buf), 16, C.byref(nb))
    return _bits(buf, nb.value)


def mb_header(kind, a, b, c, d):
    """kind 0: I16x16 header (pred mode, chroma mode, cbp chroma, cbp luma),
    1: P16x16 header (mvd x, mvd y, cbp, qp delta), 2: mb_skip_run a."""
    buf, nb = np.zeros(32, np.uint8), C.c_int(0)
    lib.orc_h264_mb_header(kind, a, b, c, d, p(buf), 32, C.byref(nb))
    return _bits(buf, nb.value)


def h264_mv_pred(mvs, wmb, start, mb, skip=False):
    m = np.ascontiguousarray(mvs, MV)
    x, y = C.c_int(0), C.c_int(0)
    (lib.orc_h264_skip_mv if skip else lib.orc_h264_mv_pred)(p(m), wmb, start, mb, C.byref(x), C.byref(y))
    return x.value, y.value


def h264_cavlc_slices(levels, dc, modes, mvs, cfg):
    """The frame's slice RBSPs (list of bytes) written by the CAVLC oracle.
    levels [nmb, 24, 16] int16, dc [nmb, 24] int32, modes [nmb] uint32,
    mvs [nmb] MV (ignored for IDR frames), cfg from slice_cfg()."""
    cfg = np.ascontiguousarray(cfg, np.int32)
    nmb, ns = int(cfg[0]) * int(cfg[1]), int(cfg[4])
    lv = np.ascontiguousarray(levels, np.int16).reshape(nmb, 24, 16)
    d = c32(dc).reshape(nmb, 24)
    md = np.ascontiguousarray(modes, np.uint32).reshape(nmb)
    mv = np.ascontiguousarray(mvs if mvs is not None else np.zeros(nmb, MV), MV).reshape(nmb)
    cap = nmb * 2560 + 4096 * ns
    out = np.zeros(cap, np.uint8)
    off, ln = np.zeros(ns, np.int64), np.zeros(ns, np.int64)
    r = lib.orc_h264_cavlc_slices(p(lv), p(d), p(md), p(mv), p(cfg), p(out), cap, p(off), p(ln))
    if r < 0:
        raise RuntimeError("CAVLC oracle: output capacity exceeded")
    return [out[o:o + n].tobytes() for o, n in zip(off, ln)]


# ------------------------------------------------------------------- HEVC
REF_HEVC = os.path.join(_HERE, "_ref", "libref_hevc.so")
REF_HEVC_DEBLOCK = os.path.join(_HERE, "_ref", "libref_hevc_deblock.so")

_sig("orc_hevc_chroma_qp", _I, [_I])
_sig("orc_hevc_quant", None, [_P, _I, _I, _I, _I, _P])
_sig("orc_hevc_dequant", None, [_P, _I, _I, _I, _P])
_sig("orc_hevc_tq4", None, [_P, _I, _I, _I, _I, _P])
_sig("orc_hevc_dqit4", None, [_P, _I, _I, _I, _P])
_sig("orc_hevc_tq8", None, [_P, _I, _I, _I, _P])
_sig("orc_hevc_dqit8", None, [_P, _I, _I, _P])
_sig("orc_hevc_mc_uni", None, [_P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P, _I])
_sig("orc_hevc_predict4", None, [_P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P])
_sig("orc_hevc_choose_mode", _I, [_P, _P, _I, _I, _I, _I, _I, _I, _I, _P])
_sig("orc_hevc_slice_y_min", _I, [_I, _I, _I])
_sig("orc_hevc_encode_i", None, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _P, _P])
_sig("orc_hevc_encode_i_ctu", None, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P])
_sig("orc_hevc_slice_y_min_ctu", _I, [_I, _I, _I, _I])
_sig("orc_hevc_zrank", C.c_longlong, [_I, _I, _I, _I, _I])
_sig("orc_hevc_encode_p", None, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _P])
_sig("orc_hevc_encode_p_ctu", None, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _I, _P, _P])
_sig("orc_hevc_refine8", C.c_uint32, [_P, _P, _I, _I, _I, _I, _I, C.POINTER(_I), C.POINTER(_I)])

HEVC_ROUND_INTER, HEVC_ROUND_INTRA_P, HEVC_ROUND_INTRA_I, HEVC_ROUND_NEAREST = 2, 3, 4, 6


def _pix(bd):
    return np.uint8 if bd == 8 else np.uint16


def hevc_tq4(res, qp, use_dst, round_q12, bd=8):
    r, o = np.ascontiguousarray(res, np.int16).reshape(16), np.zeros(16, np.int16)
    lib.orc_hevc_tq4(p(r), qp, use_dst, round_q12, bd, p(o))
    return o


def hevc_dqit4(coef, qp, use_dst, bd=8):
    c, o = np.ascontiguousarray(coef, np.int16).reshape(16), np.zeros(16, np.int16)
    lib.orc_hevc_dqit4(p(c), qp, use_dst, bd, p(o))
    return o


def hevc_tq8(res, qp, bd, round_q12):
    r, o = np.ascontiguousarray(res, np.int16).reshape(64), np.zeros(64, np.int16)
    lib.orc_hevc_tq8(p(r), qp, bd, round_q12, p(o))
    return o


def hevc_dqit8(coef, qp, bd):
    c, o = np.ascontiguousarray(coef, np.int16).reshape(64), np.zeros(64, np.int16)
    lib.orc_hevc_dqit8(p(c), qp, bd, p(o))
    return o


def hevc_mc_uni(ref, x, y, w, h, mvx, mvy, chroma, bd=8):
    ref = np.ascontiguousarray(ref, _pix(bd))
    out = np.zeros((h, w), _pix(bd))
    lib.orc_hevc_mc_uni(p(ref), ref.shape[1], ref.shape[1], ref.shape[0], x, y, w, h, mvx, mvy, chroma, bd,
                        p(out), w)
    return out


def hevc_predict4(plane, x0, y0, mode, is_luma, y_min, bd=8):
    plane = np.ascontiguousarray(plane, _pix(bd))
    o = np.zeros(16, np.int32)
    lib.orc_hevc_predict4(p(plane), bd, plane.shape[1], plane.shape[1], plane.shape[0], x0, y0, mode, is_luma,
                          y_min, p(o))
    return o


def hevc_choose_mode(src, recon, x0, y0, y_min, bd=8):
    src = np.ascontiguousarray(src, _pix(bd))
    recon = np.ascontiguousarray(recon, _pix(bd))
    o = np.zeros(16, np.int32)
    m = lib.orc_hevc_choose_mode(p(src), p(recon), bd, src.shape[1], src.shape[1], src.shape[0], x0, y0, y_min,
                                 p(o))
    return m, o


def split_uv(uv):
    """Interleaved NV12 chroma -> separate (Cb, Cr) planes."""
    return np.ascontiguousarray(uv[:, 0::2]), np.ascontiguousarray(uv[:, 1::2])


def hevc_encode_i(sy, scb, scr, qp, num_slices=1, round_q12=HEVC_ROUND_INTRA_I, bd=8, ctu_log2=4):
    """ctu_log2 4: the reference's CTU16 order (orc_hevc_encode_i); 5: CTU32
    (16x16 quadrants in z-order, CTU32-row slices)."""
    hc, wc = sy.shape
    t = (wc // 16) * (hc // 16)
    sy, scb, scr = (np.ascontiguousarray(a, _pix(bd)) for a in (sy, scb, scr))
    ry, rcb, rcr = np.zeros_like(sy), np.zeros_like(scb), np.zeros_like(scr)
    modes, coef = np.zeros(t * 16, np.uint8), np.zeros(t * 384, np.int16)
    if ctu_log2 == 4:
        lib.orc_hevc_encode_i(p(sy), p(scb), p(scr), p(ry), p(rcb), p(rcr), bd, wc, hc, qp, num_slices, round_q12,
                              p(modes), p(coef))
    else:
        lib.orc_hevc_encode_i_ctu(p(sy), p(scb), p(scr), p(ry), p(rcb), p(rcr), bd, wc, hc, qp, num_slices,
                                  round_q12, ctu_log2, p(modes), p(coef))
    return dict(modes=modes, coef=coef, y=ry, cb=rcb, cr=rcr)


def hevc_zrank(x, y, width, is_luma, ctu_log2):
    return lib.orc_hevc_zrank(x, y, width, int(is_luma), ctu_log2)


def hevc_encode_p(sy, scb, scr, fy, fcb, fcr, qp, mv, round_q12=HEVC_ROUND_INTER, bd=8, ctu_log2=4):
    """mv: int32 [nctu*4][2] quarter-sample vectors, CUs in z-order per CTU.
    ctu_log2 5: CU16 / CU32 where the vectors agree (orc_hevc_encode_p_ctu);
    culog2 (uint8 [nctu*4]) = log2 of the CU holding each CU8."""
    hc, wc = sy.shape
    t = (wc // 16) * (hc // 16)
    a = [np.ascontiguousarray(x, _pix(bd)) for x in (sy, scb, scr, fy, fcb, fcr)]
    ry, rcb, rcr = np.zeros_like(a[0]), np.zeros_like(a[1]), np.zeros_like(a[2])
    coef = np.zeros(t * 384, np.int16)
    culog2 = np.zeros(t * 4, np.uint8)
    mv = np.ascontiguousarray(mv, np.int32).reshape(t * 4, 2)
    lib.orc_hevc_encode_p_ctu(p(a[0]), p(a[1]), p(a[2]), p(a[3]), p(a[4]), p(a[5]), p(ry), p(rcb), p(rcr), bd, wc,
                              hc, qp, round_q12, p(mv), ctu_log2, p(coef), p(culog2))
    return dict(coef=coef, y=ry, cb=rcb, cr=rcr, culog2=culog2)


def hevc_refine8(cur, ref, x0, y0, mvx, mvy, bd=8):
    cur = np.ascontiguousarray(cur, _pix(bd))
    ref = np.ascontiguousarray(ref, _pix(bd))
    mx, my = C.c_int(mvx), C.c_int(mvy)
    sad = lib.orc_hevc_refine8(p(cur), p(ref), bd, cur.shape[1], cur.shape[0], x0, y0, C.byref(mx), C.byref(my))
    return mx.value, my.value, sad


def hevc_encode_frame(cy, cuv, ry, ruv, qp, is_intra, num_slices=1, rng=16, lambda8=1, deblock=True, bd=8,
                      ctu_log2=4):
    """One picture of the HEVC pipeline in its stage order: I -> CTU intra
    coding (rounding 4); P -> +-rng full search on 8x8 blocks (at 10 bits on
    the top 8 bits of each sample, v >> 2), quarter-sample refine with the
    HEVC interpolation at the full bit depth, 8x8 inter CU coding (rounding
    2); then, with deblock, the 8.7.2 filter (bS from intra / coded luma /
    vector difference).  Planes are coded-size NV12 (cuv / ruv interleaved;
    uint8 at bd 8, else uint16 samples); returns coef, modes (uint8
    [nctu*16]), mv8 (MV [nctu*4]) and the recon as NV12 (y, uv).  ctu_log2 5
    codes CTU32s: the I-picture order and availability and the slice rows
    change; outputs stay indexed by 16x16 area."""
    hc, wc = cy.shape
    t = (wc // 16) * (hc // 16)
    pix = _pix(bd)
    cb, cr = split_uv(cuv)
    mv8 = np.zeros(t * 4, MV)
    if is_intra:
        o = hevc_encode_i(cy, cb, cr, qp, num_slices, HEVC_ROUND_INTRA_I, bd, ctu_log2)
        modes = o["modes"]
    else:
        s8 = (lambda a: np.asarray(a, np.uint8)) if bd == 8 else (lambda a: (np.asarray(a) >> (bd - 8)).astype(np.uint8))
        _, i8 = me_fullsearch(s8(cy), s8(ry), rng, lambda8)
        mv = np.zeros((t * 4, 2), np.int32)
        for cu in range(t * 4):
            ctu, z = cu // 4, cu % 4
            x0, y0 = (ctu % (wc // 16)) * 16 + (z & 1) * 8, (ctu // (wc // 16)) * 16 + (z >> 1) * 8
            mx, my, sad = hevc_refine8(cy, ry, x0, y0, int(i8[cu]["dx"]) * 4, int(i8[cu]["dy"]) * 4, bd)
            mv[cu] = (mx, my)
            mv8[cu] = (mx, my, sad, 0)
        return hevc_code_p(cy, cuv, ry, ruv, qp, mv8, num_slices, deblock, bd, ctu_log2)
    uv = np.empty((hc // 2, wc), pix)
    uv[:, 0::2], uv[:, 1::2] = o["cb"], o["cr"]
    y = o["y"]
    if deblock:
        cus = o["culog2"] if not is_intra and ctu_log2 == 5 else None
        cbf = hevc_cu_cbf(o["coef"], cus, wc) if not is_intra else np.zeros(t * 4, np.uint8)
        mv4 = np.zeros((t * 4, 4), np.int32)
        mv4[:, 0], mv4[:, 1] = mv8["mvx"], mv8["mvy"]
        y, uv = hevc_deblock(y, uv, qp, is_intra, cbf, mv4, num_slices, bd, ctu_log2, cus)
    return dict(coef=o["coef"], modes=modes, mv8=mv8, y=y, uv=uv)


def hevc_code_p(cy, cuv, ry, ruv, qp, mv8, num_slices=1, deblock=True, bd=8, ctu_log2=4):
    """The P picture of hevc_encode_frame after its search, from given CU
    vectors mv8 (MV [nctu*4], quarter samples): CU coding (rounding 2) and,
    with deblock, the 8.7.2 filter; returns the same dict."""
    hc, wc = cy.shape
    t = (wc // 16) * (hc // 16)
    pix = _pix(bd)
    mv8 = np.ascontiguousarray(mv8, MV)
    mv = np.stack([mv8["mvx"], mv8["mvy"]], axis=1).astype(np.int32)
    cb, cr = split_uv(cuv)
    fb, fr = split_uv(ruv)
    o = hevc_encode_p(cy, cb, cr, ry, fb, fr, qp, mv, HEVC_ROUND_INTER, bd, ctu_log2)
    modes = np.zeros(t * 16, np.uint8)
    if ctu_log2 == 5:
        modes.reshape(t, 16)[:, :4] = o["culog2"].reshape(t, 4)
    uv = np.empty((hc // 2, wc), pix)
    uv[:, 0::2], uv[:, 1::2] = o["cb"], o["cr"]
    y = o["y"]
    if deblock:
        cus = o["culog2"] if ctu_log2 == 5 else None
        cbf = hevc_cu_cbf(o["coef"], cus, wc)
        mv4 = np.zeros((t * 4, 4), np.int32)
        mv4[:, :2] = mv
        y, uv = hevc_deblock(y, uv, qp, False, cbf, mv4, num_slices, bd, ctu_log2, cus)
    return dict(coef=o["coef"], modes=modes, mv8=mv8, y=y, uv=uv)


def hevc_cu_cbf(coef, culog2=None, wc=None):
    """Per-CU8 coded-luma flag of a P picture's levels (CU z of CTU t holds
    coef[384 t + 64 z : +64]); with culog2 (and the coded width wc) the flag
    of the transform block holding the CU8 (16x16: the area's 256 levels,
    32x32: the four areas' luma levels)."""
    a = np.asarray(coef, np.int16).reshape(-1, 384)[:, :256]
    cbf = (a.reshape(-1, 64) != 0).any(axis=1).astype(np.uint8)
    if culog2 is None:
        return cbf
    wtu = wc // 16
    area = (a != 0).any(axis=1)
    for cu in range(cbf.size):
        t, L = cu // 4, int(culog2[cu])
        if L == 4:
            cbf[cu] = area[t]
        elif L == 5:
            tx, ty = (t % wtu) & ~1, (t // wtu) & ~1
            cbf[cu] = area[ty * wtu + tx] | area[ty * wtu + tx + 1] | area[(ty + 1) * wtu + tx] | \
                area[(ty + 1) * wtu + tx + 1]
    return cbf

_sig("orc_hevc_dct_matrix", None, [_I, _P])
_sig("orc_hevc_tqn", None, [_P, _I, _I, _I, _I, _P])
_sig("orc_hevc_dqitn", None, [_P, _I, _I, _I, _P])


def hevc_dct_matrix(n):
    m = np.zeros(n * n, np.int32)
    lib.orc_hevc_dct_matrix(n, p(m))
    return m.reshape(n, n)


def hevc_tqn(res, log2n, qp, bd, round_q12):
    n2 = 1 << (2 * log2n)
    r, o = np.ascontiguousarray(res, np.int16).reshape(n2), np.zeros(n2, np.int16)
    lib.orc_hevc_tqn(p(r), log2n, qp, bd, round_q12, p(o))
    return o


def hevc_dqitn(coef, log2n, qp, bd):
    n2 = 1 << (2 * log2n)
    c, o = np.ascontiguousarray(coef, np.int16).reshape(n2), np.zeros(n2, np.int16)
    lib.orc_hevc_dqitn(p(c), log2n, qp, bd, p(o))
    return o

_sig("orc_h264_pred4x4", None, [_P, _P, _I, _I, _I, _I, _P])
_sig("orc_h264_pred8x8", None, [_P, _P, _I, _I, _I, _I, _I, _I, _P])


def h264_pred4x4(top8, left4, corner, mode, at, al):
    t, l = np.ascontiguousarray(top8, np.uint8), np.ascontiguousarray(left4, np.uint8)
    o = np.zeros(16, np.uint8)
    lib.orc_h264_pred4x4(p(t), p(l), int(corner), mode, int(at), int(al), p(o))
    return o


def h264_pred8x8(top16, left8, corner, mode, at, al, ac, atr):
    t, l = np.ascontiguousarray(top16, np.uint8), np.ascontiguousarray(left8, np.uint8)
    o = np.zeros(64, np.uint8)
    lib.orc_h264_pred8x8(p(t), p(l), int(corner), mode, int(at), int(al), int(ac), int(atr), p(o))
    return o


# modes and the neighbours each needs (a decoder never uses a mode without them)
H264_I4_NEEDS = {0: "t", 1: "l", 2: "", 3: "t", 4: "tlc", 5: "tlc", 6: "tlc", 7: "t", 8: "l"}

_sig("orc_h264_class8", _I, [_I, _I])
_sig("orc_h264_v8", _I, [_I, _I])
_sig("orc_h264_tq8x8", None, [_P, _I, _I, _P])
_sig("orc_h264_dqit8x8", None, [_P, _I, _P])
_sig("orc_h264_dequant8x8", None, [_P, _I, _P])


def h264_tq8x8(res, qp, round_q12):
    r, o = np.ascontiguousarray(res, np.int16).reshape(64), np.zeros(64, np.int16)
    lib.orc_h264_tq8x8(p(r), qp, round_q12, p(o))
    return o


def h264_dqit8x8(lv, qp):
    c, o = np.ascontiguousarray(lv, np.int16).reshape(64), np.zeros(64, np.int16)
    lib.orc_h264_dqit8x8(p(c), qp, p(o))
    return o

_sig("orc_hevc_beta", _I, [_I])
_sig("orc_hevc_tc", _I, [_I])
_sig("orc_hevc_deblock", None, [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _I])


_sig("orc_hevc_deblock_bd", None, [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _I, _I])
_sig("orc_hevc_deblock_ctu", None, [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _I, _I, _I, _P])


def hevc_deblock(y, uv, qp, is_intra, cbf, mv4, num_slices=1, bd=8, ctu_log2=4, culog2=None):
    """Returns deblocked copies of the NV12 planes (uint8 at bd 8, else
    uint16 holding the sample in the low bits); cbf uint8 [ncu], mv4 int32
    [ncu][4]."""
    t = np.uint8 if bd == 8 else np.uint16
    y = np.ascontiguousarray(y, t).copy()
    uv = np.ascontiguousarray(uv, t).copy()
    hc, wc = y.shape
    cbf = np.ascontiguousarray(cbf, np.uint8)
    mv4 = np.ascontiguousarray(mv4, np.int32).reshape(-1, 4)
    cu = None if culog2 is None else np.ascontiguousarray(culog2, np.uint8)
    lib.orc_hevc_deblock_ctu(p(y), p(uv), wc, wc, wc, hc, qp, int(is_intra), p(cbf), p(mv4), num_slices,
                             ctu_log2, bd, None if cu is None else p(cu))
    return y, uv


# ---- the CPU baseline encoder (orc_baseline.c, liborc_mt.so): the oracle's
# stages with an AVX2 search over OpenMP threads; measurement, not a checker
LIB_MT = os.path.join(_HERE, "liborc_mt.so")
_mt = None


def _mtlib():
    global _mt
    if _mt is None:
        _mt = C.CDLL(LIB_MT)
        _mt.orc_baseline_has_avx2.restype = C.c_int
        _mt.orc_baseline_fullsearch.argtypes = [_P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P]
        _mt.orc_baseline_fullsearch.restype = None
        _mt.orc_baseline_hevc_seq.argtypes = [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P]
        _mt.orc_baseline_hevc_seq.restype = _I
    return _mt


def baseline_fullsearch(cur, ref, rng, lambda8, nthreads=1, simd=1):
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    hc, wc = cur.shape
    nmb = (wc // 16) * (hc // 16)
    i16, i8 = np.zeros(nmb, IMV), np.zeros(nmb * 4, IMV)
    _mtlib().orc_baseline_fullsearch(p(cur), p(ref), wc, wc, hc, rng, lambda8, nthreads, simd, p(i16), p(i8))
    return i16, i8


def baseline_hevc_seq(ys, uvs, qp, rng, lambda8, num_slices, nthreads, simd=1):
    """ys [n, hc, wc], uvs [n, hc/2, wc] NV12 pictures, IPPP; returns
    (per-picture ms, last reconstruction y, uv)."""
    ys = np.ascontiguousarray(ys, np.uint8)
    uvs = np.ascontiguousarray(uvs, np.uint8)
    n, hc, wc = ys.shape
    ms = np.zeros(n, np.float64)
    oy, ouv = np.zeros((hc, wc), np.uint8), np.zeros((hc // 2, wc), np.uint8)
    rc = _mtlib().orc_baseline_hevc_seq(p(ys), p(uvs), n, wc, hc, qp, rng, lambda8, num_slices, nthreads, simd,
                                        p(ms), p(oy), p(ouv))
    if rc != 0:
        raise MemoryError("baseline buffers")
    return ms, oy, ouv


def baseline_has_avx2():
    return bool(_mtlib().orc_baseline_has_avx2())
