"""The frame-by-frame sequence checker (tests/seqcheck.py) on the CPU: it
accepts a sequence the oracle itself encoded and rejects one with a single
planted mistake (a level, a vector, a slice byte), in the frame where the
mistake is."""
import numpy as np
import pytest

import orc
import seqcheck


def _frames(pkg, w, h, n, seed):
    g = pkg.FrameGen(seed, w, h, 2)
    try:
        return [pkg.pad_coded(*g.frame(i)) for i in range(n)]
    finally:
        g.close()


def _h264_outs(frames, qp, ns, rng, gop):
    outs, ref = [], None
    fn = poc = 0
    idr = -1
    for i, (cy, cuv) in enumerate(frames):
        intra = i == 0 or (gop and i % gop == 0)
        ry, ruv = (cy, cuv) if ref is None else ref
        e = orc.encode_frame(cy, cuv, ry, ruv, qp, intra, ns, rng, 1, 0)
        if intra:
            fn = poc = 0
            idr += 1
        wmb, hmb = cy.shape[1] // 16, cy.shape[0] // 16
        cfg = orc.slice_cfg(wmb, hmb, qp, intra, ns, frame_num=fn, poc=poc, idr_pic_id=idr)
        outs.append(dict(quant=e["levels"], dc=e["dc"], pred_mode=e["modes"], mv=e["mvs"], nz=e["nz"],
                         mv8=e["mv8"], is_intra=intra,
                         slices=orc.h264_cavlc_slices(e["levels"], e["dc"], e["modes"], e["mvs"], cfg)))
        ref = (e["y"], e["uv"])
        fn, poc = fn + 1, poc + 2
    return outs, ref


def test_h264_checker_accepts_the_oracle_and_finds_planted_mistakes(pkg):
    frames = _frames(pkg, 96, 64, 5, 11)
    outs, last = _h264_outs(frames, 30, 2, 16, 3)
    y, uv = seqcheck.check_h264(frames, outs, 30, 2, 16, 1, gop=3, sample=1000)
    assert np.array_equal(y, last[0]) and np.array_equal(uv, last[1])
    bad = [dict(o) for o in outs]
    bad[2]["quant"] = bad[2]["quant"].copy()
    bad[2]["quant"][100] += 1
    with pytest.raises(seqcheck.Mismatch, match="frame 2: levels"):
        seqcheck.check_h264(frames, bad, 30, 2, 16, 1, gop=3, sample=1000)
    bad = [dict(o) for o in outs]
    bad[1]["mv"] = bad[1]["mv"].copy()
    bad[1]["mv"][5]["mvx"] += 1
    with pytest.raises(seqcheck.Mismatch, match="frame 1: 16x16 vector of MB 5"):
        seqcheck.check_h264(frames, bad, 30, 2, 16, 1, gop=3, sample=1000)
    bad = [dict(o) for o in outs]
    bad[4]["slices"] = [bad[4]["slices"][0][:-1] + b"\x01"] + bad[4]["slices"][1:]
    with pytest.raises(seqcheck.Mismatch, match="frame 4: CAVLC"):
        seqcheck.check_h264(frames, bad, 30, 2, 16, 1, gop=3, sample=1000)


def test_hevc_checker_accepts_the_oracle_and_finds_planted_mistakes(pkg):
    frames = _frames(pkg, 80, 48, 3, 12)
    outs, ref = [], None
    for i, (cy, cuv) in enumerate(frames):
        ry, ruv = (cy, cuv) if ref is None else ref
        o = orc.hevc_encode_frame(cy, cuv, ry, ruv, 32, i == 0, 1, 16, 1, True)
        outs.append(dict(quant=o["coef"], hevc_mode=o["modes"], mv8=o["mv8"], is_intra=i == 0))
        ref = (o["y"], o["uv"])
    y, uv = seqcheck.check_hevc(frames, outs, 32, 1, 16, 1, sample=1000)
    assert np.array_equal(y, ref[0]) and np.array_equal(uv, ref[1])
    bad = [dict(o) for o in outs]
    bad[2]["mv8"] = bad[2]["mv8"].copy()
    bad[2]["mv8"][7]["mvy"] += 1
    with pytest.raises(seqcheck.Mismatch, match="frame 2: vector of CU 3 of CTU 1"):
        seqcheck.check_hevc(frames, bad, 32, 1, 16, 1, sample=1000)
