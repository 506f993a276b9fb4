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
