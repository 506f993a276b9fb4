"""A literal reading of Rec. ITU-T H.264 clause 8.7 (deblocking filter
process) for frame pictures without MBAFF, 4:2:0, 8-bit, 4x4 transforms:
test helper, not a test module, and independent of the oracle's C filter
(oracle/orc_h264.c) and of the GPU kernel it checks.  It follows the
clause's structure sample by sample: 8.7 (macroblock order, edge flags,
luma vertical then horizontal edges, then chroma), 8.7.1 (the edge's
samples), 8.7.2.1 (bS), 8.7.2.2 (indexA / indexB, alpha, beta,
filterSamplesFlag), 8.7.2.3 (bS < 4) and 8.7.2.4 (bS == 4).

A picture's macroblocks are described by per-MB arrays: intra (bool), qp
(QP_Y), nnz[16] (non-zero coefficient levels per 4x4 luma block, raster
order in the MB), mv[16] ((x, y) quarter-sample vector per 4x4 block) and
ref (reference picture id, ignored for intra), slice (slice number) and
per-slice disable_deblocking_filter_idc, FilterOffsetA and FilterOffsetB."""
import numpy as np

# Table 8-16: alpha' and beta' by indexA / indexB
ALPHA = [0] * 16 + [4, 4, 5, 6, 7, 8, 9, 10, 12, 13, 15, 17, 20, 22, 25, 28, 32, 36, 40, 45, 50, 56, 63, 71, 80, 90,
                    101, 113, 127, 144, 162, 182, 203, 226, 255, 255]
BETA = [0] * 16 + [2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15,
                   15, 16, 16, 17, 17, 18, 18]
# Table 8-17: tC0' by indexA and bS = 1, 2, 3
TC0 = [(0, 0, 0)] * 17 + [(0, 0, 1), (0, 0, 1), (0, 0, 1), (0, 0, 1), (0, 1, 1), (0, 1, 1), (1, 1, 1), (1, 1, 1),
                          (1, 1, 1), (1, 1, 1), (1, 1, 2), (1, 1, 2), (1, 1, 2), (1, 1, 2), (1, 2, 3), (1, 2, 3),
                          (2, 2, 3), (2, 2, 4), (2, 3, 4), (2, 3, 4), (3, 3, 5), (3, 4, 6), (3, 4, 6), (4, 5, 7),
                          (4, 5, 8), (4, 6, 9), (5, 7, 10), (6, 8, 11), (6, 8, 13), (7, 10, 14), (8, 11, 16),
                          (9, 12, 18), (10, 13, 20), (11, 15, 23), (13, 17, 25)]
# Table 8-15: QPc by qPI (qPI < 30: QPc = qPI)
_QPC_HI = [29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39]


def qpc_of(qp, offset=0):
    qpi = min(max(qp + offset, 0), 51)   # Clip3(-QpBdOffsetC, 51, QP_Y + chroma_qp_index_offset), 8-bit
    return qpi if qpi < 30 else _QPC_HI[qpi - 30]


def clip3(lo, hi, v):
    return lo if v < lo else hi if v > hi else v


class Picture:
    def __init__(self, mbw, mbh, intra, qp, nnz, mv, ref, slice_of, idc, off_a, off_b, chroma_qp_offset=0):
        self.mbw, self.mbh = mbw, mbh
        self.intra, self.qp, self.nnz, self.mv, self.ref = intra, qp, nnz, mv, ref
        self.slice_of, self.idc, self.off_a, self.off_b = slice_of, idc, off_a, off_b
        self.cqp = chroma_qp_offset


def _bs(pic, mp, bp, mq, bq, mb_edge):
    """8.7.2.1 for p0 in 4x4 block bp of MB mp and q0 in block bq of mq
    (frame macroblocks, mixedModeEdgeFlag 0)."""
    if mb_edge and (pic.intra[mp] or pic.intra[mq]):
        return 4
    if pic.intra[mp] or pic.intra[mq]:
        return 3
    if pic.nnz[mp][bp] or pic.nnz[mq][bq]:
        return 2
    # different reference pictures or numbers of vectors (one list here)
    if pic.ref[mp] != pic.ref[mq]:
        return 1
    (px, py), (qx, qy) = pic.mv[mp][bp], pic.mv[mq][bq]
    if abs(px - qx) >= 4 or abs(py - qy) >= 4:
        return 1
    return 0


def _filter(s, bs, chroma, qp_p, qp_q, off_a, off_b):
    """8.7.2.2 - 8.7.2.4 on one set of samples s = [p3, p2, p1, p0, q0,
    q1, q2, q3] (chroma: p1, p0, q0, q1 in positions 2..5); returns the
    filtered list."""
    qpav = (qp_p + qp_q + 1) >> 1
    ia, ib = clip3(0, 51, qpav + off_a), clip3(0, 51, qpav + off_b)
    alpha, beta = ALPHA[ia], BETA[ib]
    p3, p2, p1, p0, q0, q1, q2, q3 = s
    if not (bs != 0 and abs(p0 - q0) < alpha and abs(p1 - p0) < beta and abs(q1 - q0) < beta):
        return s
    o = list(s)
    if bs < 4:
        tc0 = TC0[ia][bs - 1]
        ap, aq = abs(p2 - p0), abs(q2 - q0)
        tc = tc0 + 1 if chroma else tc0 + (1 if ap < beta else 0) + (1 if aq < beta else 0)
        delta = clip3(-tc, tc, (((q0 - p0) << 2) + (p1 - q1) + 4) >> 3)
        o[3] = clip3(0, 255, p0 + delta)
        o[4] = clip3(0, 255, q0 - delta)
        if not chroma and ap < beta:
            o[2] = p1 + clip3(-tc0, tc0, (p2 + ((p0 + q0 + 1) >> 1) - (p1 << 1)) >> 1)
        if not chroma and aq < beta:
            o[5] = q1 + clip3(-tc0, tc0, (q2 + ((p0 + q0 + 1) >> 1) - (q1 << 1)) >> 1)
    else:
        ap, aq = abs(p2 - p0), abs(q2 - q0)
        if not chroma and ap < beta and abs(p0 - q0) < ((alpha >> 2) + 2):
            o[3] = (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3
            o[2] = (p2 + p1 + p0 + q0 + 2) >> 2
            o[1] = (2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3
        else:
            o[3] = (2 * p1 + p0 + q1 + 2) >> 2
        if not chroma and aq < beta and abs(p0 - q0) < ((alpha >> 2) + 2):
            o[4] = (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3
            o[5] = (p0 + q0 + q1 + q2 + 2) >> 2
            o[6] = (2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3
        else:
            o[4] = (2 * q1 + q0 + p1 + 2) >> 2
    return o


def _edge(plane, pic, m, vertical, e, chroma, mb_edge):
    """8.7.1: the samples across edge e (luma: at 4e; chroma: at 4e in the
    8x8 block) of MB m, k along the edge."""
    mbx, mby = m % pic.mbw, m // pic.mbw
    n = 8 if chroma else 16
    bw = 8 if chroma else 16
    x0, y0 = mbx * bw, mby * bw
    mp = (m - 1 if vertical else m - pic.mbw) if mb_edge else m
    sl = pic.slice_of[m]
    for k in range(n):
        lk = 2 * k if chroma else k          # the luma sample row/column of chroma sample k
        le = 2 * (4 * e) if chroma else 4 * e  # the luma coordinate of the edge
        if vertical:
            bq = (lk // 4) * 4 + le // 4
            bp = (lk // 4) * 4 + (3 if mb_edge else le // 4 - 1)
        else:
            bq = (le // 4) * 4 + lk // 4
            bp = (3 if mb_edge else le // 4 - 1) * 4 + lk // 4
        bs = _bs(pic, mp, bp, m, bq, mb_edge)
        if chroma:
            qp_p, qp_q = qpc_of(pic.qp[mp], pic.cqp), qpc_of(pic.qp[m], pic.cqp)
        else:
            qp_p, qp_q = pic.qp[mp], pic.qp[m]
        if vertical:
            yy, xs = y0 + k, [x0 + 4 * e + d for d in range(-4, 4)]
            s = [int(plane[yy, x]) for x in xs] if not chroma else \
                [0, 0] + [int(plane[yy, x]) for x in xs[2:6]] + [0, 0]
            o = _filter(s, bs, chroma, qp_p, qp_q, pic.off_a[sl], pic.off_b[sl])
            for i, x in enumerate(xs):
                if not chroma or 2 <= i < 6:
                    plane[yy, x] = o[i]
        else:
            xx, ys = x0 + k, [y0 + 4 * e + d for d in range(-4, 4)]
            s = [int(plane[y, xx]) for y in ys] if not chroma else \
                [0, 0] + [int(plane[y, xx]) for y in ys[2:6]] + [0, 0]
            o = _filter(s, bs, chroma, qp_p, qp_q, pic.off_a[sl], pic.off_b[sl])
            for i, y in enumerate(ys):
                if not chroma or 2 <= i < 6:
                    plane[y, xx] = o[i]


def deblock(pic, y, cb, cr):
    """8.7 over the picture, in place: y (16 mbh x 16 mbw), cb / cr
    (8 mbh x 8 mbw) uint8 planes."""
    y16, cb16, cr16 = y.astype(np.int32), cb.astype(np.int32), cr.astype(np.int32)
    for m in range(pic.mbw * pic.mbh):
        mbx, mby = m % pic.mbw, m // pic.mbw
        idc = pic.idc[pic.slice_of[m]]
        if idc == 1:
            continue
        # filterLeftMbEdgeFlag / filterTopMbEdgeFlag: the neighbour exists
        # and, with idc 2, lies in the same slice
        left = mbx > 0 and not (idc == 2 and pic.slice_of[m - 1] != pic.slice_of[m])
        top = mby > 0 and not (idc == 2 and pic.slice_of[m - pic.mbw] != pic.slice_of[m])
        for vertical, first in ((True, left), (False, top)):
            for e in range(4):
                if e == 0 and not first:
                    continue
                _edge(y16, pic, m, vertical, e, False, e == 0)
        for plane in (cb16, cr16):
            for vertical, first in ((True, left), (False, top)):
                for e in range(2):               # chroma edges at 0 and 4 (transform_size_8x8_flag 0)
                    if e == 0 and not first:
                        continue
                    _edge(plane, pic, m, vertical, e, True, e == 0)
    y[:] = y16
    cb[:] = cb16
    cr[:] = cr16
