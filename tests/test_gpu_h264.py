"""GPU parity of the H.264 compute pipeline against the CPU oracle.

Every comparison is bit-exact (integer/byte work).  Stage tests call the
launchers of include/bc250_stages.h on device buffers; pipeline tests go
through the gpu_compute_* ABI exactly as the reference's host encoder does
(begin_picture / dispatch_encode_ext / end_picture / sync_slot / staging).
"""
import ctypes as C
import sys

import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def P(t):
    return C.c_void_p(t.data_ptr())


def frames(pkg, seed, w, h, n, nrect):
    g = pkg.FrameGen(seed, w, h, nrect)
    out = [pkg.pad_coded(*g.frame(i)) for i in range(n)]
    g.close()
    return out


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------------ stages

@pytest.mark.parametrize("w,h,rng,lam", [(128, 96, 16, 1), (320, 240, 32, 2), (320, 176, 64, 1), (64, 48, 64, 0)])
def test_me_fullsearch_stage(pkg, w, h, rng, lam):
    torch = _torch()
    say(f"me_fullsearch {w}x{h} R={rng} lambda8={lam}")
    (cy, _), (ry, _) = frames(pkg, 7, w, h, 2, 3)
    # make the second picture the current one so motion is non-trivial
    cur, ref = ry, cy
    hc, wc = cur.shape
    nmb = (wc // 16) * (hc // 16)
    dcur, dref = dev(torch, cur), dev(torch, ref)
    o16 = torch.zeros(nmb * 4, dtype=torch.int32, device="cuda")
    o8 = torch.zeros(nmb * 16, dtype=torch.int32, device="cuda")
    rc = pkg.hip.bc250_launch_me_fullsearch(C.c_void_p(dcur.data_ptr()), wc, C.c_void_p(dref.data_ptr()), wc, wc,
                                            hc, rng, lam, C.c_void_p(o16.data_ptr()), C.c_void_p(o8.data_ptr()),
                                            None)
    assert rc == 0
    torch.cuda.synchronize()
    e16, e8 = orc.me_fullsearch(cur, ref, rng, lam)
    g16 = host(o16).view(orc.IMV)
    g8 = host(o8).view(orc.IMV)
    np.testing.assert_array_equal(g16, e16)
    np.testing.assert_array_equal(g8, e8)


def test_me_refine_stage(pkg):
    torch = _torch()
    w, h, rng = 192, 128, 16
    say(f"me_refine {w}x{h}")
    (cy, _), (ry, _) = frames(pkg, 11, w, h, 2, 4)
    cur, ref = ry, cy
    hc, wc = cur.shape
    i16, i8 = orc.me_fullsearch(cur, ref, rng, 1)
    e16, e8 = orc.me_refine(cur, ref, i16, i8)
    nmb = len(i16)
    dcur, dref = dev(torch, cur), dev(torch, ref)
    di16, di8 = dev(torch, i16.view(np.int32)), dev(torch, i8.view(np.int32))
    m16 = torch.zeros(nmb * 4, dtype=torch.int32, device="cuda")
    m8 = torch.zeros(nmb * 16, dtype=torch.int32, device="cuda")
    rc = pkg.hip.bc250_launch_me_refine(C.c_void_p(dcur.data_ptr()), wc, C.c_void_p(dref.data_ptr()), wc, wc, hc,
                                        C.c_void_p(di16.data_ptr()), C.c_void_p(di8.data_ptr()),
                                        C.c_void_p(m16.data_ptr()), C.c_void_p(m8.data_ptr()), 1, None)
    assert rc == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(m16).view(orc.MV), e16)
    np.testing.assert_array_equal(host(m8).view(orc.MV), e8)


def _random_mvs(rs, nmb, span):
    mv = np.zeros(nmb, orc.MV)
    mv["mvx"] = rs.integers(-span, span + 1, nmb)
    mv["mvy"] = rs.integers(-span, span + 1, nmb)
    return mv


@pytest.mark.parametrize("qp", [0, 17, 28, 40, 51])
def test_p_encode_stage(pkg, qp):
    torch = _torch()
    w, h = 176, 144
    say(f"p_encode {w}x{h} qp={qp}")
    (cy, cuv), (ry, ruv) = frames(pkg, 3, w, h, 2, 3)
    hc, wc = cy.shape
    nmb = (wc // 16) * (hc // 16)
    mvs = _random_mvs(np.random.default_rng(qp), nmb, 90)  # crosses the frame edges
    exp = orc.encode_p(cy, cuv, ry, ruv, qp, mvs)
    d = {k: dev(torch, v) for k, v in dict(cy=cy, cuv=cuv, ry=ry, ruv=ruv, mv=mvs.view(np.int32)).items()}
    oy = torch.zeros((hc, wc), dtype=torch.uint8, device="cuda")
    ouv = torch.zeros((hc // 2, wc), dtype=torch.uint8, device="cuda")
    lv = torch.zeros(nmb * 384, dtype=torch.int16, device="cuda")
    dc = torch.zeros(nmb * 24, dtype=torch.int32, device="cuda")
    nz = torch.zeros(nmb * 24, dtype=torch.int32, device="cuda")
    md = torch.zeros(nmb, dtype=torch.int32, device="cuda")
    P = lambda t: C.c_void_p(t.data_ptr())
    rc = pkg.hip.bc250_launch_p_encode(P(d["cy"]), P(d["cuv"]), wc, wc, P(d["ry"]), P(d["ruv"]), wc, wc, P(oy),
                                       P(ouv), wc, wc, wc, hc, qp, P(d["mv"]), P(lv), P(dc), P(nz), P(md), None)
    assert rc == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(lv), exp["levels"])
    np.testing.assert_array_equal(host(dc), exp["dc"])
    np.testing.assert_array_equal(host(nz).view(np.uint32), exp["nz"])
    np.testing.assert_array_equal(host(md).view(np.uint32), exp["modes"])
    np.testing.assert_array_equal(host(oy), exp["y"])
    np.testing.assert_array_equal(host(ouv), exp["uv"])


@pytest.mark.parametrize("qp,ns", [(10, 1), (28, 1), (28, 3), (45, 2)])
def test_i_encode_stage(pkg, qp, ns):
    torch = _torch()
    w, h = 320, 240  # 20 x 15 MBs: slice starts off column 0 for ns=3
    say(f"i_encode {w}x{h} qp={qp} slices={ns}")
    (cy, cuv), = frames(pkg, 5, w, h, 1, 6)
    hc, wc = cy.shape
    nmb = (wc // 16) * (hc // 16)
    exp = orc.encode_i(cy, cuv, qp, ns)
    dcy, dcuv = dev(torch, cy), dev(torch, cuv)
    oy = torch.zeros((hc, wc), dtype=torch.uint8, device="cuda")
    ouv = torch.zeros((hc // 2, wc), dtype=torch.uint8, device="cuda")
    lv = torch.zeros(nmb * 384, dtype=torch.int16, device="cuda")
    dc = torch.zeros(nmb * 24, dtype=torch.int32, device="cuda")
    nz = torch.zeros(nmb * 24, dtype=torch.int32, device="cuda")
    md = torch.zeros(nmb, dtype=torch.int32, device="cuda")
    nwork = pkg.hip.bc250_intra_work_bytes(wc, hc)
    sync = torch.zeros((nwork + 3) // 4, dtype=torch.int32, device="cuda")
    P = lambda t: C.c_void_p(t.data_ptr())
    rc = pkg.hip.bc250_launch_i_encode(P(dcy), P(dcuv), wc, wc, P(oy), P(ouv), wc, wc, wc, hc, qp, ns, P(lv), P(dc),
                                       P(nz), P(md), P(sync), 1 << 22, None)
    assert rc == 0
    torch.cuda.synchronize()
    assert host(sync)[1] == 0, "wavefront wait gave up"
    np.testing.assert_array_equal(host(md).view(np.uint32), exp["modes"])
    np.testing.assert_array_equal(host(lv), exp["levels"])
    np.testing.assert_array_equal(host(dc), exp["dc"])
    np.testing.assert_array_equal(host(nz).view(np.uint32), exp["nz"])
    np.testing.assert_array_equal(host(oy), exp["y"])
    np.testing.assert_array_equal(host(ouv), exp["uv"])


@pytest.mark.parametrize("intra,qp,ns,idc", [(1, 30, 1, 0), (0, 30, 1, 0), (0, 45, 3, 2), (1, 51, 2, 2),
                                             (0, 20, 1, 0)])
def test_deblock_stage(pkg, intra, qp, ns, idc):
    torch = _torch()
    w, h = 320, 240
    say(f"deblock {w}x{h} intra={intra} qp={qp} slices={ns} idc={idc}")
    rs = np.random.default_rng(qp * 10 + intra)
    (y, uv), = frames(pkg, 9, w, h, 1, 8)
    # blocky input: add per-4x4 offsets so the filters have edges to act on
    off = rs.integers(-12, 13, (y.shape[0] // 4, y.shape[1] // 4))
    y = np.clip(y.astype(int) + np.kron(off, np.ones((4, 4), int)), 0, 255).astype(np.uint8)
    coff = rs.integers(-6, 7, (uv.shape[0] // 4, uv.shape[1] // 8))
    uv = np.clip(uv.astype(int) + np.kron(coff, np.ones((4, 8), int)), 0, 255).astype(np.uint8)
    hc, wc = y.shape
    nmb = (wc // 16) * (hc // 16)
    nz = np.where(rs.random(nmb * 24) < 0.3, rs.integers(1, 1 << 16, nmb * 24), 0).astype(np.uint32)
    mvs = _random_mvs(rs, nmb, 6)
    ey, euv = orc.deblock(y, uv, qp, intra, nz, mvs, ns, idc)
    dy, duv = dev(torch, y), dev(torch, uv)
    dnz, dmv = dev(torch, nz.view(np.int32)), dev(torch, mvs.view(np.int32))
    nwork = pkg.hip.bc250_deblock_work_bytes(wc, hc)
    sync = torch.zeros((nwork + 3) // 4, dtype=torch.int32, device="cuda")
    P = lambda t: C.c_void_p(t.data_ptr())
    rc = pkg.hip.bc250_launch_deblock(P(dy), P(duv), wc, wc, wc, hc, qp, intra, P(dnz), P(dmv), ns, idc, P(sync),
                                      1 << 22, None)
    assert rc == 0
    torch.cuda.synchronize()
    assert host(sync)[1] == 0, "wavefront wait gave up"
    # boundary strengths left in the workspace (after the flags and hand-off records)
    mbw, mbh = wc // 16, hc // 16
    off = 256 + nmb * 24 * 8
    words = host(sync).view(np.uint8)[off:off + nmb * 16].view(np.uint32).reshape(nmb, 4)
    np.testing.assert_array_equal(words, expected_bs_words(mbw, mbh, intra, nz, mvs, ns, idc))
    assert (ey != y).any()  # the case exercises the filter
    np.testing.assert_array_equal(host(dy), ey)
    np.testing.assert_array_equal(host(duv), euv)


def expected_bs_words(mbw, mbh, intra, nz, mvs, ns, idc):
    """H.264 8.7.2.1 strengths, 4 bits per (edge, segment): words vertical
    edges 0-1, vertical 2-3, horizontal 0-1, horizontal 2-3; 0 where an edge
    is not filtered."""
    nmb = mbw * mbh
    out = np.zeros((nmb, 4), np.uint32)
    sl = [orc.lib.orc_slice_of(m, nmb, ns) for m in range(nmb)]
    for m in range(nmb):
        x, y = m % mbw, m // mbw
        for d in range(2):
            w = [0, 0]
            for e in range(4):
                for i in range(4):
                    if e == 0:
                        nb = m - 1 if d == 0 else m - mbw
                        exists = (x > 0) if d == 0 else (y > 0)
                        if exists and idc == 2 and sl[nb] != sl[m]:
                            exists = False
                        if not exists:
                            continue
                    if intra:
                        b = 4 if e == 0 else 3
                    else:
                        pm = m if e else (m - 1 if d == 0 else m - mbw)
                        bq = i * 4 + e if d == 0 else e * 4 + i
                        bp = (i * 4 + e - 1 if e else i * 4 + 3) if d == 0 else ((e - 1) * 4 + i if e else 12 + i)
                        if nz[pm * 24 + bp] or nz[m * 24 + bq]:
                            b = 2
                        elif abs(int(mvs[pm]["mvx"]) - int(mvs[m]["mvx"])) >= 4 or \
                                abs(int(mvs[pm]["mvy"]) - int(mvs[m]["mvy"])) >= 4:
                            b = 1
                        else:
                            b = 0
                    w[e >> 1] |= b << (4 * ((e & 1) * 4 + i))
            out[m, 2 * d:2 * d + 2] = w
    return out


# ------------------------------------------------------------------ pipeline

def _run_sequence(pkg, w, h, n, qp, ns, rng, lam, seed, nrect, idc=0, gop=0):
    say(f"pipeline {w}x{h} frames={n} qp={qp} slices={ns} R={rng}")
    fr = frames(pkg, seed, w, h, n, nrect)
    g = pkg.FrameGen(seed, w, h, nrect)
    enc = pkg.Encoder(w, h, search_range=rng, lambda8=lam, deblock_idc=idc)
    ref_y = ref_uv = None
    try:
        for i in range(n):
            vy, vuv = g.frame(i)
            im = enc.new_image()
            enc.upload(im, vy, vuv)
            intra = i == 0 or (gop and i % gop == 0)
            slot = enc.encode(im, qp, intra, ns)
            got = enc.results(slot)
            cy, cuv = fr[i]
            ry, ruv = (cy, cuv) if ref_y is None else (ref_y, ref_uv)
            exp = orc.encode_frame(cy, cuv, ry, ruv, qp, intra, ns, rng, lam, idc)
            np.testing.assert_array_equal(got["mv"], exp["mvs"], err_msg=f"frame {i} mvs")
            np.testing.assert_array_equal(got["mv8"], exp["mv8"], err_msg=f"frame {i} mv8")
            np.testing.assert_array_equal(got["pred_mode"], exp["modes"], err_msg=f"frame {i} modes")
            np.testing.assert_array_equal(got["quant"], exp["levels"], err_msg=f"frame {i} levels")
            np.testing.assert_array_equal(got["dc"], exp["dc"], err_msg=f"frame {i} dc")
            np.testing.assert_array_equal(got["nz"], exp["nz"], err_msg=f"frame {i} nz")
            gy, guv = enc.recon()
            np.testing.assert_array_equal(gy, exp["y"][:h, :w], err_msg=f"frame {i} recon y")
            np.testing.assert_array_equal(guv, exp["uv"][:(h + 1) // 2, :2 * ((w + 1) // 2)],
                                          err_msg=f"frame {i} recon uv")
            my, muv = enc.map_recon()  # ctx.recon_image, as the CPU ME route reads it
            np.testing.assert_array_equal(my, exp["y"], err_msg=f"frame {i} mapped recon y")
            np.testing.assert_array_equal(muv, exp["uv"], err_msg=f"frame {i} mapped recon uv")
            ref_y, ref_uv = exp["y"], exp["uv"]
            say(f"  frame {i} ok ({'I' if intra else 'P'})")
    finally:
        enc.close()
        g.close()


def test_pipeline_small_ipppp(pkg):
    _torch()
    _run_sequence(pkg, 128, 96, 5, 28, 1, 16, 1, seed=21, nrect=3)


def test_pipeline_odd_size_slices(pkg):
    _torch()
    # 200x120 is not 16-aligned: exercises the coded-size padding
    _run_sequence(pkg, 200, 120, 4, 36, 3, 16, 2, seed=22, nrect=2, idc=2, gop=3)


def test_pipeline_640x480_cqp32_4slices(pkg):
    """The 640x480 workload: CQP 32, 4 slices, +-32 full search (first frames)."""
    _torch()
    _run_sequence(pkg, 640, 480, 3, 32, 4, 32, 1, seed=2024, nrect=6)


# ------------------------------------------------------------------ full size

def test_pipeline_640x368_r64_chain(pkg):
    """+-64 over an I + 4 P chain, 4 slices: every stage against the oracle."""
    _torch()
    _run_sequence(pkg, 640, 368, 5, 27, 4, 64, 1, seed=64, nrect=6)


def test_pipeline_1080p_ipppp_chain(pkg):
    """1920x1080, I then 5 P frames, 4 slices, +-16: every stage of every
    frame against the oracle's own chain (each P frame predicts from the
    previous reconstruction)."""
    _torch()
    _run_sequence(pkg, 1920, 1080, 6, 27, 4, 16, 1, seed=1080, nrect=12)


def test_1080p_stages_full_frame(pkg):
    """1920x1080 at +-64: I frame, then a P frame; every stage output is
    compared with the oracle at full size except the integer search, which
    is checked by recomputing each winner's SAD and cost and by the oracle's
    exhaustive search on a sample of macroblocks."""
    torch = _torch()
    w, h, rng, lam, qp = 1920, 1080, 64, 1, 27
    say("1080p full frame")
    g = pkg.FrameGen(1080, w, h, 12)
    enc = pkg.Encoder(w, h, search_range=rng, lambda8=lam, deblock_idc=0)
    try:
        v0, v1 = g.frame(0), g.frame(1)
        c0, c1 = pkg.pad_coded(*v0), pkg.pad_coded(*v1)
        im0, im1 = enc.new_image(), enc.new_image()
        enc.upload(im0, *v0)
        enc.upload(im1, *v1)
        s0 = enc.encode(im0, qp, True, 1)
        r0 = enc.results(s0)
        e0 = orc.encode_i(c0[0], c0[1], qp, 1)
        np.testing.assert_array_equal(r0["pred_mode"], e0["modes"])
        np.testing.assert_array_equal(r0["quant"], e0["levels"])
        ey0, euv0 = orc.deblock(e0["y"], e0["uv"], qp, 1, e0["nz"], np.zeros(len(e0["modes"]), orc.MV), 1, 0)
        gy0, guv0 = enc.recon()
        np.testing.assert_array_equal(gy0, ey0[:h, :w])
        say("  I frame ok")
        s1 = enc.encode(im1, qp, False, 1)
        r1 = enc.results(s1)
        mv = r1["mv"]
        cur, ref = c1[0], ey0
        hc, wc = cur.shape
        mbw = wc // 16
        # sub-pel refine and everything after it at full size, from the GPU's own MVs
        rs = np.random.default_rng(0)
        sample = rs.choice(len(mv), 24, replace=False)
        i16 = np.zeros(len(mv), orc.IMV)
        i8 = np.zeros(len(mv) * 4, orc.IMV)
        for m in sample:
            a, b = orc.me_fullsearch_mb(cur, ref, rng, lam, m % mbw, m // mbw)
            i16[m], i8[4 * m:4 * m + 4] = a, b
        e16, e8 = orc.me_refine(cur, ref, i16, i8)
        np.testing.assert_array_equal(mv[sample], e16[sample])
        np.testing.assert_array_equal(r1["mv8"].reshape(-1, 4)[sample], e8.reshape(-1, 4)[sample])
        ep = orc.encode_p(cur, c1[1], ey0, euv0, qp, mv)
        np.testing.assert_array_equal(r1["quant"], ep["levels"])
        np.testing.assert_array_equal(r1["nz"], ep["nz"])
        np.testing.assert_array_equal(r1["dc"], ep["dc"])
        ey1, euv1 = orc.deblock(ep["y"], ep["uv"], qp, 0, ep["nz"], mv, 1, 0)
        gy1, guv1 = enc.recon()
        np.testing.assert_array_equal(gy1, ey1[:h, :w])
        np.testing.assert_array_equal(guv1, euv1[:h // 2, :w])
        say("  P frame ok")
    finally:
        enc.close()
        g.close()


def test_1080p_integer_search_costs(pkg):
    """Every MB's integer winner at +-64 carries the SAD it claims, and no
    candidate of a sampled MB beats it (exhaustive numpy search)."""
    torch = _torch()
    w, h, rng, lam = 1920, 1080, 64, 1
    say("1080p integer search properties")
    g = pkg.FrameGen(77, w, h, 12)
    (ry, _), (cy, _) = pkg.pad_coded(*g.frame(0)), pkg.pad_coded(*g.frame(1))
    g.close()
    hc, wc = cy.shape
    nmb = (wc // 16) * (hc // 16)
    dcur, dref = dev(torch, cy), dev(torch, ry)
    o16 = torch.zeros(nmb * 4, dtype=torch.int32, device="cuda")
    o8 = torch.zeros(nmb * 16, dtype=torch.int32, device="cuda")
    assert pkg.hip.bc250_launch_me_fullsearch(C.c_void_p(dcur.data_ptr()), wc, C.c_void_p(dref.data_ptr()), wc, wc,
                                              hc, rng, lam, C.c_void_p(o16.data_ptr()),
                                              C.c_void_p(o8.data_ptr()), None) == 0
    torch.cuda.synchronize()
    g16 = host(o16).view(orc.IMV)
    pad = np.pad(ry.astype(np.int32), rng, mode="edge")
    mbw = wc // 16
    blocks = cy.astype(np.int32).reshape(hc // 16, 16, mbw, 16).transpose(0, 2, 1, 3).reshape(nmb, 16, 16)
    # SAD at the reported MV, all MBs
    ys = (np.arange(nmb) // mbw) * 16 + rng + g16["dy"]
    xs = (np.arange(nmb) % mbw) * 16 + rng + g16["dx"]
    idx_y = ys[:, None, None] + np.arange(16)[None, :, None]
    idx_x = xs[:, None, None] + np.arange(16)[None, None, :]
    sad = np.abs(blocks - pad[idx_y, idx_x]).sum(axis=(1, 2))
    np.testing.assert_array_equal(g16["sad"], sad)
    l1 = np.abs(g16["dx"]) + np.abs(g16["dy"])
    np.testing.assert_array_equal(g16["cost"], sad + 4 * lam * l1)
    assert (np.abs(g16["dx"]) <= rng).all() and (np.abs(g16["dy"]) <= rng).all()
    # exhaustive search of sampled MBs
    for m in np.random.default_rng(1).choice(nmb, 6, replace=False):
        y0, x0 = (m // mbw) * 16, (m % mbw) * 16
        win = pad[y0:y0 + 16 + 2 * rng, x0:x0 + 16 + 2 * rng]
        sw = np.lib.stride_tricks.sliding_window_view(win, (16, 16))  # [dy][dx][16][16]
        costs = np.abs(sw - blocks[m]).sum(axis=(2, 3))
        d = np.arange(-rng, rng + 1)
        costs = costs + 4 * lam * (np.abs(d)[:, None] + np.abs(d)[None, :])
        k = np.argmin(costs)  # first in raster order on ties
        assert (g16["dy"][m], g16["dx"][m]) == (k // (2 * rng + 1) - rng, k % (2 * rng + 1) - rng)


# ------------------------------------------------------------------ intra 4x4 / 8x8

@pytest.mark.parametrize("size", [4, 8])
def test_intra_nxn_pred_batch(pkg, size):
    torch = _torch()
    say(f"intra {size}x{size} prediction batch")
    rs = np.random.default_rng(size)
    reps = []
    for mode, needs in orc.H264_I4_NEEDS.items():
        for at in (0, 1):
            for al in (0, 1):
                for ac in (0, 1):
                    for atr in (0, 1):
                        if ("t" in needs and not at) or ("l" in needs and not al) or ("c" in needs and not ac):
                            continue
                        reps += [(mode, at, al, ac, atr)] * 12
    n = len(reps)
    stride = 16 if size == 4 else 32
    nbr = rs.integers(0, 256, (n, stride)).astype(np.uint8)
    ctl = np.array([m | (at << 4) | (al << 5) | (ac << 6) | (atr << 7) for (m, at, al, ac, atr) in reps], np.uint8)
    if size == 4:
        for i, (m, at, al, ac, atr) in enumerate(reps):
            if not atr:
                nbr[i, 9:13] = nbr[i, 8]  # missing top-right run = p[3,-1] (the caller's substitution at 4x4)
    dn, dc = dev(torch, nbr), dev(torch, ctl)
    dout = torch.zeros(n * size * size, dtype=torch.uint8, device="cuda")
    f = pkg.hip.bc250_h264_pred4x4_batch if size == 4 else pkg.hip.bc250_h264_pred8x8_batch
    assert f(C.c_void_p(dn.data_ptr()), C.c_void_p(dc.data_ptr()), n, C.c_void_p(dout.data_ptr()), None) == 0
    torch.cuda.synchronize()
    got = host(dout).reshape(n, size * size)
    for i, (m, at, al, ac, atr) in enumerate(reps):
        if size == 4:
            want = orc.h264_pred4x4(nbr[i, 5:13], nbr[i, 0:4], nbr[i, 4], m, at, al)
        else:
            want = orc.h264_pred8x8(nbr[i, 9:25], nbr[i, 0:8], nbr[i, 8], m, at, al, ac, atr)
        assert np.array_equal(got[i], want), (i, m, at, al, ac, atr)


@pytest.mark.parametrize("qp", [0, 12, 27, 36, 51])
def test_t8x8_batch(pkg, qp):
    torch = _torch()
    say(f"8x8 transform batch qp={qp}")
    rs = np.random.default_rng(qp + 3)
    n = 257
    res = rs.integers(-255, 256, (n, 64)).astype(np.int16)
    res[::2] //= 13
    for rnd in (2, 4):
        dres, dlv = dev(torch, res), torch.zeros((n, 64), dtype=torch.int16, device="cuda")
        P = lambda t: C.c_void_p(t.data_ptr())
        assert pkg.hip.bc250_h264_tq8x8_batch(P(dres), n, qp, rnd, P(dlv), None) == 0
        dback = torch.zeros_like(dlv)
        assert pkg.hip.bc250_h264_dqit8x8_batch(P(dlv), n, qp, P(dback), None) == 0
        torch.cuda.synchronize()
        lv, back = host(dlv), host(dback)
        for b in range(0, n, 3):
            want = orc.h264_tq8x8(res[b], qp, rnd)
            assert np.array_equal(lv[b], want), (b, rnd)
            assert np.array_equal(back[b], orc.h264_dqit8x8(want, qp)), b


# ------------------------------------------------- the reference's Tier-2 search

@pytest.mark.parametrize("w,h,radius", [(128, 96, 2), (200, 120, 8), (320, 240, 16), (64, 64, 8)])
def test_square_me_stage(pkg, w, h, radius):
    torch = _torch()
    say(f"square me {w}x{h} R={radius}")
    fr = frames(pkg, 60 + radius, w, h, 2, 4)
    (ry, _), (cy, _) = fr
    want = orc.square_me(cy, ry, w, h, radius)
    hc, wc = cy.shape
    dcy, dry = dev(torch, cy), dev(torch, ry)
    out = torch.zeros(want.shape[0] * 4, dtype=torch.int32, device="cuda")
    assert pkg.hip.bc250_launch_square_me(P(dcy), wc, P(dry), wc, w, h, radius, P(out), None) == 0
    torch.cuda.synchronize()
    got = host(out).view(orc.MV)
    np.testing.assert_array_equal(got, want)
    assert w < 128 or (want["mvx"] != 0).sum() > 3
    # a ramp moved by half a sample: zero-vector SAD 768, every half-sample
    # branch of the search taken
    xx = np.arange(wc)[None, :].repeat(hc, 0)
    r2 = np.clip(20 + 6 * (xx % 32), 0, 255).astype(np.uint8)
    c2 = ((r2.astype(int) + np.roll(r2, -1, 1) + 1) >> 1).astype(np.uint8)
    want = orc.square_me(c2, r2, w, h, radius)
    assert ((want["mvx"] % 4) != 0).sum() > 0
    dcy, dry = dev(torch, c2), dev(torch, r2)
    assert pkg.hip.bc250_launch_square_me(P(dcy), wc, P(dry), wc, w, h, radius, P(out), None) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(out).view(orc.MV), want)


def test_square_me_reference_interface(pkg):
    """cpu_simd_me_search_frame through the reference-named C entry points
    (bc250_cpu_me.h), on the reference test's frame: one textured MB moved by
    (+2, +2) is found at (8, 8) quarter samples; a flat +10 offset costs 2560."""
    _torch()
    cur = np.full((64, 64), 100, np.uint8)
    ref = cur.copy()
    yy, xx = np.mgrid[0:16, 0:16]
    cur[16:32, 16:32] = (xx * 13 + yy * 7 + 40) % 256
    ref[18:34, 18:34] = cur[16:32, 16:32]
    cfg = pkg.CpuSimdMeConfig()
    pkg.hip.cpu_simd_me_config_init.argtypes = [C.c_void_p, C.c_uint, C.c_uint]
    pkg.hip.cpu_simd_me_config_init(C.addressof(cfg), 64, 64)
    assert cfg.width_in_mbs == 4 and cfg.core_ids[0] == -1 and cfg.search_radius == 0
    cfg.search_radius = 8
    mvs = np.zeros(16, orc.MV)
    assert pkg.hip.cpu_simd_me_search_frame(cur.ctypes.data, 64, ref.ctypes.data, 64, 64, 64, mvs.ctypes.data,
                                            C.addressof(cfg)) == 0
    assert (mvs[5]["mvx"], mvs[5]["mvy"]) == (8, 8) and (mvs[0]["mvx"], mvs[0]["mvy"]) == (0, 0)
    np.testing.assert_array_equal(mvs, orc.square_me(cur, ref, 64, 64, 8))
    a = np.full((64, 64), 128, np.uint8)
    b = np.full((64, 64), 138, np.uint8)
    assert pkg.hip.cpu_simd_sad_16x16(a.ctypes.data, 64, b.ctypes.data, 64) == 2560


def test_pipeline_square_me(pkg):
    """The H.264 pipeline with the Tier-2 search (square_radius 2, the radius
    the reference's encoder runs with) in place of the full search."""
    _torch()
    w, h, qp, ns, R = 160, 96, 30, 2, 2
    fr = frames(pkg, 23, w, h, 4, 3)
    g = pkg.FrameGen(23, w, h, 3)
    enc = pkg.Encoder(w, h, deblock_idc=0, square_radius=R)
    ref = None
    try:
        for i in range(4):
            im = enc.new_image()
            enc.upload(im, *g.frame(i))
            got = enc.results(enc.encode(im, qp, i == 0, ns))
            cy, cuv = fr[i]
            if i == 0:
                e = orc.encode_i(cy, cuv, qp, ns)
                mvs = np.zeros((w // 16) * (h // 16), orc.MV)
            else:
                mvs = orc.square_me(cy, ref[0], w, h, R)
                e = orc.encode_p(cy, cuv, ref[0], ref[1], qp, mvs)
            y, uv = orc.deblock(e["y"], e["uv"], qp, i == 0, e["nz"], mvs, ns, 0)
            np.testing.assert_array_equal(got["mv"], mvs, err_msg=f"frame {i} mvs")
            np.testing.assert_array_equal(got["quant"], e["levels"], err_msg=f"frame {i} levels")
            my, muv = enc.map_recon()
            np.testing.assert_array_equal(my, y, err_msg=f"frame {i} recon")
            np.testing.assert_array_equal(muv, uv, err_msg=f"frame {i} recon uv")
            ref = (y, uv)
            say(f"  frame {i} ok")
    finally:
        enc.close()
        g.close()


# ------------------------------------------------- the reference's diamond search

@pytest.mark.parametrize("w,h,lam", [(128, 96, 5), (320, 176, 8), (200, 120, 5)])
def test_diamond_me_stage(pkg, w, h, lam):
    torch = _torch()
    say(f"diamond me {w}x{h} lambda={lam}")
    (ry, _), (cy, _) = frames(pkg, 70 + lam, w, h, 2, 5)
    hc, wc = cy.shape
    want = orc.diamond_me(cy, ry, lam)
    dcy, dry = dev(torch, cy), dev(torch, ry)
    out = torch.zeros(want.shape[0] * 4, dtype=torch.int32, device="cuda")
    assert pkg.hip.bc250_launch_diamond_me(P(dcy), wc, P(dry), wc, wc, hc, lam, P(out), None) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(out).view(orc.MV), want)
    assert ((want["mvx"] % 4) != 0).sum() > 0 and (want["mvx"] != 0).sum() > 3


@pytest.mark.parametrize("me_mode", [0, 1])
def test_pipeline_diamond_me(pkg, me_mode):
    """The H.264 pipeline with the reference's diamond search in place of the
    full search; me_mode 1 is the reference's fast mode (lambda 8)."""
    _torch()
    w, h, qp, ns = 160, 112, 28, 1
    fr = frames(pkg, 24, w, h, 3, 4)
    g = pkg.FrameGen(24, w, h, 4)
    enc = pkg.Encoder(w, h, deblock_idc=0, diamond=1)
    ref = None
    try:
        for i in range(3):
            im = enc.new_image()
            enc.upload(im, *g.frame(i))
            got = enc.results(enc.encode(im, qp, i == 0, ns, me_mode=me_mode))
            cy, cuv = fr[i]
            if i == 0:
                e = orc.encode_i(cy, cuv, qp, ns)
                mvs = np.zeros((w // 16) * (h // 16), orc.MV)
            else:
                mvs = orc.diamond_me(cy, ref[0], 8 if me_mode == 1 else 5)
                e = orc.encode_p(cy, cuv, ref[0], ref[1], qp, mvs)
            y, uv = orc.deblock(e["y"], e["uv"], qp, i == 0, e["nz"], mvs, ns, 0)
            np.testing.assert_array_equal(got["mv"], mvs, err_msg=f"frame {i} mvs")
            np.testing.assert_array_equal(got["quant"], e["levels"], err_msg=f"frame {i} levels")
            my, muv = enc.map_recon()
            np.testing.assert_array_equal(my, y, err_msg=f"frame {i} recon")
            ref = (y, uv)
            say(f"  frame {i} ok")
    finally:
        enc.close()
        g.close()
