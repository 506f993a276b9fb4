"""Checks a sequence the GPU pipeline encoded against the oracle, frame by
frame, the way the sequence was coded: each P frame is checked on the
oracle's own reconstruction of the frame before it, so one wrong frame
fails there and not only later.

At full size the oracle's +-64 integer full search takes ~14 ms per MB, so
the search is checked on a seeded sample of MBs (always including the four
corners; the sample of the sub-pel refinement is the same), and everything
after it -- the P coding, deblocking, CAVLC slices -- on the whole frame
with the GPU's own vectors.  A wrong vector outside the sample passes here;
the sample is drawn afresh for every frame, and the small-frame tests check
the search on every MB.

Used by the GPU tests and by bench.py's CPU-baseline leg (the checker, never
the thing measured)."""
import numpy as np

import orc


class Mismatch(AssertionError):
    pass


def _same(what, frame, got, want):
    got, want = np.asarray(got), np.asarray(want)
    if got.shape != want.shape:
        raise Mismatch(f"frame {frame}: {what} shape {got.shape} != oracle {want.shape}")
    if got.dtype.names:
        diff = np.zeros(got.shape, bool)
        for k in got.dtype.names:
            diff |= got[k] != want[k]
    else:
        diff = got != want
    if diff.any():
        idx = np.flatnonzero(diff)
        raise Mismatch(f"frame {frame}: {what} differs from the oracle at {idx.size} of {diff.size} entries "
                       f"(first at flat index {idx[0]}: {got.reshape(-1)[idx[0]]} vs {want.reshape(-1)[idx[0]]})")


def _sample_mbs(nmb, mbw, n, seed):
    """A seeded sample of n MBs plus the four corners; none when n is 0 (the
    caller checks the search itself)."""
    if n == 0:
        return []
    mbh = nmb // mbw
    corners = [0, mbw - 1, (mbh - 1) * mbw, nmb - 1]
    rs = np.random.default_rng(seed)
    pick = rs.choice(nmb, size=min(n, nmb), replace=False) if n < nmb else np.arange(nmb)
    return sorted(set(int(m) for m in pick) | set(corners))


def frame_types(n, gop):
    return [i == 0 or (gop > 0 and i % gop == 0) for i in range(n)]


def check_h264(frames, outs, qp, num_slices, rng, lambda8, gop=0, idc=0, sample=48, seed=0, slice_hdr=None,
               last_idr_pic_id=-1):
    """frames: coded-size (y, uv) of every frame; outs: the GPU's per-frame
    outputs (Encoder.encode_sequence(collect=True)).  Checks modes, vectors,
    levels, DC terms, nz masks and, when present, the CAVLC slices
    (frame_num / POC / idr_pic_id counted the way the context counts them;
    last_idr_pic_id is that of the context's previous IDR, -1 for none).
    Returns the oracle's last reconstruction (y, uv)."""
    hdr = dict(slice_hdr or {})
    ry = ruv = None
    fn = poc = 0
    idr = last_idr_pic_id
    for i, ((cy, cuv), got, intra) in enumerate(zip(frames, outs, frame_types(len(frames), gop))):
        if isinstance(got, Exception):
            raise got
        if got.get("is_intra", intra) != intra:
            raise Mismatch(f"frame {i}: coded as intra={got['is_intra']}, expected {intra}")
        hc, wc = cy.shape
        mbw, nmb = wc // 16, (wc // 16) * (hc // 16)
        if intra:
            e = orc.encode_i(cy, cuv, qp, num_slices)
            mvs = np.zeros(nmb, orc.MV)
            _same("16x16 vectors (zero in I frames)", i, got["mv"], mvs)
            _same("8x8 vectors (zero in I frames)", i, got["mv8"], np.zeros(nmb * 4, orc.MV))
            fn = poc = 0
            idr = (idr + 1) & 0xffff
        else:
            for m in _sample_mbs(nmb, mbw, sample, seed * 1000003 + i):
                i16, i8 = orc.me_fullsearch_mb(cy, ry, rng, lambda8, m % mbw, m // mbw)
                m16, m8 = orc.me_refine_mb(cy, ry, m % mbw, m // mbw, i16, i8)
                _same(f"16x16 vector of MB {m}", i, got["mv"][m], m16)
                _same(f"8x8 vectors of MB {m}", i, got["mv8"][4 * m:4 * m + 4], m8)
            mvs = got["mv"]
            e = orc.encode_p(cy, cuv, ry, ruv, qp, mvs)
        _same("modes", i, got["pred_mode"], e["modes"])
        _same("levels", i, got["quant"], e["levels"])
        _same("dc", i, got["dc"], e["dc"])
        _same("nz", i, got["nz"], e["nz"])
        if "slices" in got:
            fb, pb = hdr.get("log2_max_frame_num_minus4", 0) + 4, hdr.get("log2_max_poc_lsb_minus4", 0) + 4
            cfg = orc.slice_cfg(mbw, nmb // mbw, qp, intra, num_slices, frame_num=fn % (1 << fb),
                                poc=poc % (1 << pb), idr_pic_id=max(idr, 0), deblock_idc=idc, **hdr)
            want = orc.h264_cavlc_slices(e["levels"], e["dc"], e["modes"], mvs, cfg)
            if got["slices"] != want:
                raise Mismatch(f"frame {i}: CAVLC slices differ from the oracle writer's")
        if idc != 1:
            ry, ruv = orc.deblock(e["y"], e["uv"], qp, intra, e["nz"], mvs, num_slices, idc)
        else:
            ry, ruv = e["y"], e["uv"]
        fn, poc = fn + 1, poc + 2
    return ry, ruv


def check_hevc(frames, outs, qp, num_slices, rng, lambda8, gop=0, deblock=True, bd=8, ctu_log2=4, sample=48,
               seed=0):
    """The HEVC pipeline's frames: coefficients, PU modes (I) / CU sizes
    (P at CTU32) and CU vectors.  Returns the oracle's last reconstruction."""
    ry = ruv = None
    to8 = (lambda a: np.asarray(a, np.uint8)) if bd == 8 else (lambda a: (np.asarray(a) >> (bd - 8)).astype(np.uint8))
    for i, ((cy, cuv), got, intra) in enumerate(zip(frames, outs, frame_types(len(frames), gop))):
        if isinstance(got, Exception):
            raise got
        if got.get("is_intra", intra) != intra:
            raise Mismatch(f"frame {i}: coded as intra={got['is_intra']}, expected {intra}")
        hc, wc = cy.shape
        wt, nt = wc // 16, (wc // 16) * (hc // 16)
        if intra:
            o = orc.hevc_encode_frame(cy, cuv, cy, cuv, qp, True, num_slices, rng, lambda8, deblock, bd, ctu_log2)
            _same("CU vectors (zero in I pictures)", i, got["mv8"], np.zeros(nt * 4, orc.MV))
        else:
            c8, r8 = to8(cy), to8(ry)
            for t in _sample_mbs(nt, wt, sample, seed * 1000003 + i):
                _, i8 = orc.me_fullsearch_mb(c8, r8, rng, lambda8, t % wt, t // wt)
                for z in range(4):
                    x0, y0 = (t % wt) * 16 + (z & 1) * 8, (t // wt) * 16 + (z >> 1) * 8
                    mx, my, sad = orc.hevc_refine8(cy, ry, x0, y0, int(i8[z]["dx"]) * 4, int(i8[z]["dy"]) * 4, bd)
                    g = got["mv8"][4 * t + z]
                    if (int(g["mvx"]), int(g["mvy"]), int(g["sad"])) != (mx, my, sad):
                        raise Mismatch(f"frame {i}: vector of CU {z} of CTU {t} is "
                                       f"({g['mvx']}, {g['mvy']}, sad {g['sad']}), oracle ({mx}, {my}, sad {sad})")
            o = orc.hevc_code_p(cy, cuv, ry, ruv, qp, got["mv8"], num_slices, deblock, bd, ctu_log2)
        _same("coefficients", i, got["quant"], o["coef"])
        _same("modes", i, got["hevc_mode"], o["modes"])
        ry, ruv = o["y"], o["uv"]
    return ry, ruv
