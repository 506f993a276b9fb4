"""Decodes a frame parsed by h264_cavlc_parse.SliceParser with the reference
decoder's own prediction, motion compensation, DC transforms and inverse
transforms (oracle/_ref/libref_h264.so, built from src/h264_pred.c,
h264_mc.c, h264_recon.c) -- test helper, not a test module.  decode_frame
stops before the loop filter; loop_filter then applies clause 8.7
(tests/h264_loopfilter.py) with the parsed slices' settings."""
import ctypes as C

import numpy as np

import orc
from test_oracle_h264 import _decode_mb_residual, _dequant, p, ref


def _chroma_from_levels(lib, dst_uv, mbx, mby, levels, cdc, qp):
    """Chroma residual of one MB from its transmitted DC levels and AC
    levels (the DC-level form of test_oracle_h264._decode_chroma)."""
    qpc = orc.lib.orc_chroma_qp(qp)
    dqc = _dequant(lib, qpc)
    lv = levels.reshape(24, 16)
    for comp in range(2):
        cl = np.ascontiguousarray(cdc[comp], np.int16)
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


def decode_frame(ps, qp, ref_y=None, ref_uv=None):
    """Coded-size luma / NV12 chroma of the frame in parser `ps`; P frames
    predict from ref_y / ref_uv (coded size)."""
    lib = ref()
    wmb, hmb = ps.wmb, ps.hmb
    wc, hc = wmb * 16, hmb * 16
    dy = np.zeros((hc, wc), np.uint8)
    duv = np.zeros((hc // 2, wc), np.uint8)
    buf = np.zeros(64 * 64, np.uint8)
    for m in range(wmb * hmb):
        mbx, mby = m % wmb, m // wmb
        y0, x0 = mby * 16, mbx * 16
        lv = ps.levels[m].astype(np.int16).reshape(-1)
        if ps.kind[m] == "I":
            ta = mby > 0 and ps.slice_of[m - wmb] == ps.slice_of[m]
            la = mbx > 0 and ps.slice_of[m - 1] == ps.slice_of[m]
            top = dy[y0 - 1, x0:x0 + 16].copy() if mby > 0 else np.full(16, 128, np.uint8)
            left = dy[y0:y0 + 16, x0 - 1].copy() if mbx > 0 else np.full(16, 128, np.uint8)
            corner = int(dy[y0 - 1, x0 - 1]) if (mbx > 0 and mby > 0) else 128
            pred = np.zeros((16, 16), np.uint8)
            lib.h264d_pred16x16(p(pred), 16, int(ps.pm[m]), p(top), p(left), corner, ta, la)
            dy[y0:y0 + 16, x0:x0 + 16] = pred
            _decode_mb_residual(lib, dy, duv, mbx, mby, lv, qp, True, ps.ldc[m])
            for comp in range(2):
                t = duv[mby * 8 - 1, x0 + comp:x0 + 16:2].copy() if ta else np.full(8, 128, np.uint8)
                l = duv[mby * 8:mby * 8 + 8, x0 - 2 + comp].copy() if la else np.full(8, 128, np.uint8)
                cp = np.zeros((8, 8), np.uint8)
                lib.h264d_pred_chroma(p(cp), 8, int(ps.cm[m]), p(np.ascontiguousarray(t)),
                                      p(np.ascontiguousarray(l)), 128, ta, la)
                duv[mby * 8:mby * 8 + 8, x0 + comp:x0 + 16:2] = cp
        else:
            mx, my = int(ps.mv[m][0]), int(ps.mv[m][1])
            ss = C.c_int()
            src = lib.h264d_mc_fetch_luma(p(buf), C.byref(ss), p(ref_y), wc, wc, hc, x0 + (mx >> 2), y0 + (my >> 2),
                                          16, 16)
            pred = np.zeros((16, 16), np.uint8)
            lib.h264d_mc_luma(p(pred), 16, src, ss.value, 16, 16, mx & 3, my & 3)
            dy[y0:y0 + 16, x0:x0 + 16] = pred
            for comp in range(2):
                plane = np.ascontiguousarray(ref_uv[:, comp::2])
                src = lib.h264d_mc_fetch_chroma(p(buf), C.byref(ss), p(plane), wc // 2, wc // 2, hc // 2,
                                                mbx * 8 + (mx >> 3), mby * 8 + (my >> 3), 8, 8)
                cp = np.zeros((8, 8), np.uint8)
                lib.h264d_mc_chroma(p(cp), 8, src, ss.value, 8, 8, mx & 7, my & 7)
                duv[mby * 8:mby * 8 + 8, x0 + comp:x0 + 16:2] = cp
            _decode_mb_residual(lib, dy, duv, mbx, mby, lv, qp, False, None)
        _chroma_from_levels(lib, duv, mbx, mby, lv, ps.cdc[m], qp)
    return dy, duv


def loop_filter(ps, qp, dy, duv):
    """Clause 8.7 over a decoded frame in place (dy coded-size luma, duv NV12
    chroma): per MB intra or not, QP_Y = qp, the non-zero 4x4 luma blocks
    from the parsed TotalCoeff (blkIdx order -> raster), one vector per
    inter MB, one reference picture; each slice's
    disable_deblocking_filter_idc and offsets (slice_alpha_c0_offset_div2,
    slice_beta_offset_div2, doubled) from its header."""
    import h264_loopfilter as lf
    n = ps.wmb * ps.hmb
    nnz = []
    for m in range(n):
        r = [0] * 16
        for k in range(16):
            x, y = (k & 1) | (k & 4) >> 1, (k & 2) >> 1 | (k & 8) >> 2
            r[y * 4 + x] = int(ps.tcl[m][k])
        nnz.append(r)
    intra = [ps.kind[m] == "I" for m in range(n)]
    mv = [[(0, 0) if intra[m] else (int(ps.mv[m][0]), int(ps.mv[m][1]))] * 16 for m in range(n)]
    hd = ps.headers
    pic = lf.Picture(ps.wmb, ps.hmb, intra=intra, qp=[qp] * n, nnz=nnz, mv=mv, ref=[0] * n,
                     slice_of=[int(v) for v in ps.slice_of], idc=[h["deblock_idc"] for h in hd],
                     off_a=[2 * h.get("alpha", 0) for h in hd], off_b=[2 * h.get("beta", 0) for h in hd])
    cb, cr = np.ascontiguousarray(duv[:, 0::2]), np.ascontiguousarray(duv[:, 1::2])
    lf.deblock(pic, dy, cb, cr)
    duv[:, 0::2], duv[:, 1::2] = cb, cr
    return dy, duv
