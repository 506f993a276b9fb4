"""GPU parity of the HEVC stages (include/bc250_hevc.h) against the CPU
oracle (oracle/orc_hevc.c, itself pinned to the reference's compiled HEVC
functions).  Every comparison is bit-exact."""
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


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def P(t):
    return C.c_void_p(t.data_ptr())


def host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("bd", [8, 10])
def test_tq4_dqit4_batch(pkg, bd):
    torch = _torch()
    rs = np.random.default_rng(100 + bd)
    vmax = (1 << bd) - 1
    n = 999
    res = rs.integers(-vmax, vmax + 1, (n, 16)).astype(np.int16)
    res[::3] //= 37  # small residuals too
    res[5] = 0
    for qp in ([0, 13, 22, 27, 37, 51] if bd == 8 else [0, 12, 34, 39, 49, 63]):
        for dst in (0, 1):
            for rnd in (2, 4, 6):
                say(f"tq4 bd={bd} qp={qp} dst={dst} rnd={rnd}")
                dres, dout = dev(torch, res), torch.zeros((n, 16), dtype=torch.int16, device="cuda")
                assert pkg.hip.bc250_hevc_tq4_batch(P(dres), n, qp, dst, rnd, bd, P(dout), None) == 0
                dback = torch.zeros_like(dout)
                assert pkg.hip.bc250_hevc_dqit4_batch(P(dout), n, qp, dst, bd, P(dback), None) == 0
                torch.cuda.synchronize()
                lv, back = host(dout), host(dback)
                for b in range(0, n, 7):
                    want = orc.hevc_tq4(res[b], qp, dst, rnd, bd)
                    assert np.array_equal(lv[b], want), (b, qp)
                    assert np.array_equal(back[b], orc.hevc_dqit4(want, qp, dst, bd)), (b, qp)


@pytest.mark.parametrize("bd", [8, 10])
def test_tq8_dqit8_batch(pkg, bd):
    torch = _torch()
    rs = np.random.default_rng(200 + bd)
    vmax = (1 << bd) - 1
    n = 301
    res = rs.integers(-vmax, vmax + 1, (n, 64)).astype(np.int16)
    res[1::2] //= 29
    for qp in (0, 22, 27, 37, 51, 63):
        say(f"tq8 bd={bd} qp={qp}")
        dres, dout = dev(torch, res), torch.zeros((n, 64), dtype=torch.int16, device="cuda")
        assert pkg.hip.bc250_hevc_tq8_batch(P(dres), n, qp, bd, 2, P(dout), None) == 0
        dback = torch.zeros_like(dout)
        assert pkg.hip.bc250_hevc_dqit8_batch(P(dout), n, qp, bd, P(dback), None) == 0
        torch.cuda.synchronize()
        lv, back = host(dout), host(dback)
        for b in range(0, n, 5):
            want = orc.hevc_tq8(res[b], qp, bd, 2)
            assert np.array_equal(lv[b], want), (b, qp)
            assert np.array_equal(back[b], orc.hevc_dqit8(want, qp, bd)), (b, qp)


@pytest.mark.parametrize("bd", [8, 10])
def test_mc_uni_batch(pkg, bd):
    torch = _torch()
    rs = np.random.default_rng(300 + bd)
    pix = np.uint8 if bd == 8 else np.uint16
    plane = rs.integers(0, 1 << bd, (72, 88)).astype(pix)
    dplane = dev(torch, plane.view(np.int8 if bd == 8 else np.int16))
    for chroma in (0, 1):
        nfrac = 8 if chroma else 4
        for (w, h) in ((4, 4), (8, 8), (16, 16), (8, 4)):
            say(f"mc bd={bd} chroma={chroma} {w}x{h}")
            n = 200
            blk = np.stack([rs.integers(-12, 88, n), rs.integers(-12, 72, n),
                            rs.integers(-10 * nfrac, 10 * nfrac, n), rs.integers(-10 * nfrac, 10 * nfrac, n)],
                           1).astype(np.int32)
            dblk = dev(torch, blk)
            dout = torch.zeros(n * w * h * (1 if bd == 8 else 2), dtype=torch.uint8, device="cuda")
            assert pkg.hip.bc250_hevc_mc_uni_batch(P(dplane), 88, 88, 72, P(dblk), n, w, h, chroma, bd, P(dout),
                                                   None) == 0
            torch.cuda.synchronize()
            got = host(dout).view(pix).reshape(n, h, w)
            for i in range(n):
                want = orc.hevc_mc_uni(plane, int(blk[i, 0]), int(blk[i, 1]), w, h, int(blk[i, 2]), int(blk[i, 3]),
                                       chroma, bd)
                assert np.array_equal(got[i], want), (i, blk[i])


def _nv12(y, cb, cr):
    uv = np.empty((cb.shape[0], cb.shape[1] * 2), np.uint8)
    uv[:, 0::2], uv[:, 1::2] = cb, cr
    return y, uv


def _frames(pkg, seed, w, h, n, nrect=3):
    g = pkg.FrameGen(seed, w, h, nrect)
    out = [pkg.pad_coded(*g.frame(i)) for i in range(n)]
    g.close()
    return out


def _ten(a, rs):
    """10-bit samples from 8-bit ones as the 4K P010 workload makes them: 4 v +
    uniform [0, 3]."""
    return np.asarray(a, np.uint16) * 4 + rs.integers(0, 4, np.shape(a)).astype(np.uint16)


def _p010(torch, a):
    """Device copy of a plane of 10-bit samples in P010 form (sample << 6)."""
    return dev(torch, (np.asarray(a, np.uint16) << 6).view(np.int16))


def _unp010(t):
    v = host(t).view(np.uint16)
    assert not (v & 63).any()
    return v >> 6


@pytest.mark.parametrize("w,h,rng,bd", [(128, 96, 16, 8), (320, 176, 32, 8), (128, 96, 16, 10), (256, 160, 32, 10)])
def test_refine8_stage(pkg, w, h, rng, bd):
    torch = _torch()
    say(f"hevc refine8 {w}x{h} R={rng} bd={bd}")
    (cy, cuv), (ry, ruv) = _frames(pkg, 77, w, h, 2)
    cy, ry = ry, cy  # search frame 0 in frame 1's reconstruction stand-in
    hc, wc = cy.shape
    _, i8 = orc.me_fullsearch(cy, ry, rng, 1)
    if bd == 10:
        rs = np.random.default_rng(5)
        cy, ry = _ten(cy, rs), _ten(ry, rs)
        dcy, dry, pitch = _p010(torch, cy), _p010(torch, ry), wc * 2
    else:
        dcy, dry, pitch = dev(torch, cy), dev(torch, ry), wc
    di8 = dev(torch, i8.view(np.int32))
    dout = torch.zeros(i8.shape[0] * 4, dtype=torch.int32, device="cuda")
    assert pkg.hip.bc250_hevc_launch_refine8(P(dcy), pitch, P(dry), pitch, wc, hc, P(di8), P(dout), bd, None) == 0
    torch.cuda.synchronize()
    got = host(dout).reshape(-1, 4)
    for cu in range(i8.shape[0]):
        ctu, z = cu // 4, cu % 4
        x0 = (ctu % (wc // 16)) * 16 + (z & 1) * 8
        y0 = (ctu // (wc // 16)) * 16 + (z >> 1) * 8
        mx, my, sad = orc.hevc_refine8(cy, ry, x0, y0, int(i8[cu]["dx"]) * 4, int(i8[cu]["dy"]) * 4, bd)
        assert (got[cu, 0], got[cu, 1], got[cu, 2]) == (mx, my, sad), cu


@pytest.mark.parametrize("qp,bd", [(0, 8), (22, 8), (32, 8), (45, 8), (51, 8), (0, 10), (27, 10), (51, 10)])
def test_p_encode_stage(pkg, qp, bd):
    torch = _torch()
    say(f"hevc P coding qp={qp} bd={bd}")
    w, h = 192, 112
    (cy, cuv), (ry, ruv) = _frames(pkg, 91 + qp, w, h, 2, 4)
    hc, wc = cy.shape
    t = (wc // 16) * (hc // 16)
    rs = np.random.default_rng(qp)
    mv = rs.integers(-70, 71, (t * 4, 2)).astype(np.int32)
    mv[::5] = 0
    if bd == 10:
        cy, cuv, ry, ruv = (_ten(a, rs) for a in (cy, cuv, ry, ruv))
    cb, cr = orc.split_uv(cuv)
    fb, fr = orc.split_uv(ruv)
    want = orc.hevc_encode_p(cy, cb, cr, ry, fb, fr, qp, mv, orc.HEVC_ROUND_INTER, bd)
    mv4 = np.zeros((t * 4, 4), np.int32)
    mv4[:, :2] = mv
    if bd == 10:
        d = [_p010(torch, a) for a in (cy, cuv, ry, ruv)] + [dev(torch, mv4)]
        pitch, back = wc * 2, _unp010
    else:
        d = [dev(torch, a) for a in (cy, cuv, ry, ruv, mv4)]
        pitch, back = wc, host
    oy, ouv = torch.zeros_like(d[0]), torch.zeros_like(d[1])
    coef = torch.zeros(t * 384, dtype=torch.int16, device="cuda")
    assert pkg.hip.bc250_hevc_launch_p_encode(P(d[0]), P(d[1]), pitch, pitch, P(d[2]), P(d[3]), pitch, pitch, P(oy),
                                              P(ouv), pitch, pitch, wc, hc, qp, orc.HEVC_ROUND_INTER, P(d[4]), P(coef),
                                              bd, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(host(coef), want["coef"])
    assert np.array_equal(back(oy), want["y"])
    gb, gr = orc.split_uv(back(ouv))
    assert np.array_equal(gb, want["cb"]) and np.array_equal(gr, want["cr"])
    assert (want["coef"] != 0).any()


@pytest.mark.parametrize("w,h,qp,bd,ctu", [(128, 96, 27, 8, 4), (112, 80, 37, 8, 5), (144, 96, 32, 10, 4),
                                            (112, 80, 22, 10, 5)])
def test_refine_p_encode_fused_equals_stages(pkg, w, h, qp, bd, ctu):
    """The fused refinement + CU coding equals refine8 followed by
    p_encode_ctu on the same inputs: vectors, levels, reconstruction, modes,
    CU sizes and coded-luma flags."""
    torch = _torch()
    say(f"hevc refine+P fused {w}x{h} qp={qp} bd={bd} ctu_log2={ctu}")
    (cy, cuv), (ry, ruv) = _frames(pkg, 301 + qp, w, h, 2, 4)
    hc, wc = cy.shape
    t = (wc // 16) * (hc // 16)
    _, i8 = orc.me_fullsearch(ry, cy, 16, 1)
    if bd == 10:
        rs = np.random.default_rng(qp)
        cy, cuv, ry, ruv = (_ten(a, rs) for a in (cy, cuv, ry, ruv))
        d = [_p010(torch, a) for a in (cy, cuv, ry, ruv)]
        pitch = wc * 2
    else:
        d = [dev(torch, a) for a in (cy, cuv, ry, ruv)]
        pitch = wc
    di8 = dev(torch, i8.view(np.int32))
    outs = []
    for fused in (False, True):
        oy, ouv = torch.zeros_like(d[0]), torch.zeros_like(d[1])
        mv8 = torch.zeros(t * 16, dtype=torch.int32, device="cuda")
        coef = torch.zeros(t * 384, dtype=torch.int16, device="cuda")
        cul = torch.zeros(t * 4, dtype=torch.uint8, device="cuda")
        modes = torch.full((t * 16,), 9, dtype=torch.uint8, device="cuda")
        cbf = torch.full((t * 4,), 7, dtype=torch.uint8, device="cuda")
        cbfp = None if ctu == 5 else P(cbf)
        if fused:
            assert pkg.hip.bc250_hevc_launch_refine_p_encode(P(d[0]), P(d[1]), pitch, pitch, P(d[2]), P(d[3]), pitch,
                                                             pitch, P(oy), P(ouv), pitch, pitch, wc, hc, qp,
                                                             orc.HEVC_ROUND_INTER, P(di8), P(mv8), P(coef), ctu,
                                                             P(cul), P(modes), cbfp, bd, None) == 0
        else:
            assert pkg.hip.bc250_hevc_launch_refine8(P(d[0]), pitch, P(d[2]), pitch, wc, hc, P(di8), P(mv8), bd,
                                                     None) == 0
            assert pkg.hip.bc250_hevc_launch_p_encode_ctu(P(d[0]), P(d[1]), pitch, pitch, P(d[2]), P(d[3]), pitch,
                                                          pitch, P(oy), P(ouv), pitch, pitch, wc, hc, qp,
                                                          orc.HEVC_ROUND_INTER, P(mv8), P(coef), ctu, P(cul),
                                                          P(modes), cbfp, bd, None) == 0
        torch.cuda.synchronize()
        outs.append([host(x) for x in (mv8, coef, oy, ouv, cul, modes, cbf)])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    assert not outs[1][5].reshape(t, 16)[:, 4:].any()
    if ctu == 4:
        want = orc.hevc_cu_cbf(outs[1][1])
        assert np.array_equal(outs[1][6], want) and want.any()


@pytest.mark.parametrize("w,h,qp,bd", [(112, 80, 22, 8), (256, 160, 37, 8), (112, 80, 27, 10), (192, 128, 4, 10)])
def test_p_encode_ctu32_stage(pkg, w, h, qp, bd):
    """CTU32 P coding: CU16 / CU32 where the CU8 vectors agree (whole CTU32s,
    single areas, areas of partial CTU32s at the edges), CU8 elsewhere; the
    per-area level layout, the CU-size map, the modes copy and the
    per-transform-block cbf."""
    torch = _torch()
    say(f"hevc P ctu32 {w}x{h} qp={qp} bd={bd}")
    (cy, cuv), (ry, ruv) = _frames(pkg, 191 + qp, w, h, 2, 4)
    hc, wc = cy.shape
    wtu, htu = wc // 16, hc // 16
    t = wtu * htu
    rs = np.random.default_rng(qp + w)
    mv = rs.integers(-70, 71, (t * 4, 2)).astype(np.int32)
    for Y in range((htu + 1) // 2):
        for X in range((wtu + 1) // 2):
            areas = [(2 * Y + (a >> 1)) * wtu + 2 * X + (a & 1) for a in range(4)
                     if 2 * X + (a & 1) < wtu and 2 * Y + (a >> 1) < htu]
            kind = int(rs.integers(0, 3))
            v = rs.integers(-70, 71, 2)
            for a in areas:
                if kind == 2 or (kind == 1 and rs.random() < 0.5):
                    mv[a * 4:a * 4 + 4] = v if kind == 2 else rs.integers(-70, 71, 2)
    if bd == 10:
        cy, cuv, ry, ruv = (_ten(a, rs) for a in (cy, cuv, ry, ruv))
    cb, cr = orc.split_uv(cuv)
    fb, fr = orc.split_uv(ruv)
    want = orc.hevc_encode_p(cy, cb, cr, ry, fb, fr, qp, mv, orc.HEVC_ROUND_INTER, bd, 5)
    assert (want["culog2"] == 5).any() and (want["culog2"] == 4).any() and (want["culog2"] == 3).any()
    mv4 = np.zeros((t * 4, 4), np.int32)
    mv4[:, :2] = mv
    if bd == 10:
        d = [_p010(torch, a) for a in (cy, cuv, ry, ruv)] + [dev(torch, mv4)]
        pitch, back = wc * 2, _unp010
    else:
        d = [dev(torch, a) for a in (cy, cuv, ry, ruv, mv4)]
        pitch, back = wc, host
    oy, ouv = torch.zeros_like(d[0]), torch.zeros_like(d[1])
    coef = torch.zeros(t * 384, dtype=torch.int16, device="cuda")
    culog2 = torch.zeros(t * 4, dtype=torch.uint8, device="cuda")
    modes = torch.zeros(t * 16, dtype=torch.uint8, device="cuda")
    assert pkg.hip.bc250_hevc_launch_p_encode_ctu(P(d[0]), P(d[1]), pitch, pitch, P(d[2]), P(d[3]), pitch, pitch,
                                                  P(oy), P(ouv), pitch, pitch, wc, hc, qp, orc.HEVC_ROUND_INTER,
                                                  P(d[4]), P(coef), 5, P(culog2), P(modes), None, bd, None) == 0
    cbf = torch.zeros(t * 4, dtype=torch.uint8, device="cuda")
    assert pkg.hip.bc250_hevc_cu_cbf_ctu(P(coef), wc, hc, P(culog2), P(cbf), None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(host(culog2), want["culog2"])
    assert np.array_equal(host(modes).reshape(t, 16)[:, :4].reshape(-1), want["culog2"])
    assert not host(modes).reshape(t, 16)[:, 4:].any()
    assert np.array_equal(host(coef), want["coef"])
    assert np.array_equal(back(oy), want["y"])
    gb, gr = orc.split_uv(back(ouv))
    assert np.array_equal(gb, want["cb"]) and np.array_equal(gr, want["cr"])
    assert np.array_equal(host(cbf), orc.hevc_cu_cbf(want["coef"], want["culog2"], wc))
    # deblocking with the CU-size map: no filtering inside a CU16 / CU32
    uv = np.empty((hc // 2, wc), want["y"].dtype)
    uv[:, 0::2], uv[:, 1::2] = want["cb"], want["cr"]
    wcbf = orc.hevc_cu_cbf(want["coef"], want["culog2"], wc)
    wy, wuv = orc.hevc_deblock(want["y"], uv, qp, 0, wcbf, mv4, 2, bd, 5, want["culog2"])
    if bd == 10:
        dy, duv = _p010(torch, want["y"]), _p010(torch, uv)
    else:
        dy, duv = dev(torch, want["y"]), dev(torch, uv)
    assert pkg.hip.bc250_hevc_launch_deblock_ctu(P(dy), P(duv), pitch, pitch, wc, hc, qp, 0, P(cbf), P(d[4]),
                                                 P(culog2), 2, 5, bd, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(back(dy), wy) and np.array_equal(back(duv), wuv)


@pytest.mark.parametrize("w,h,qp,ns,bd", [(64, 48, 22, 1, 8), (128, 96, 37, 2, 8), (200, 120, 27, 3, 8),
                                           (320, 240, 32, 4, 8), (64, 48, 22, 1, 10), (128, 96, 45, 2, 10),
                                           (208, 128, 27, 3, 10), (320, 240, 0, 4, 10)])
def test_i_encode_stage(pkg, w, h, qp, ns, bd):
    _i_encode_case(pkg, w, h, qp, ns, bd, 4)


@pytest.mark.parametrize("w,h,qp,ns,bd", [(112, 80, 27, 1, 8), (128, 96, 37, 2, 8), (320, 240, 32, 4, 8),
                                           (48, 16, 27, 1, 8), (208, 144, 22, 3, 10), (64, 48, 45, 1, 10)])
def test_i_encode_ctu32_stage(pkg, w, h, qp, ns, bd):
    """CTU32 order and availability: partial CTU32s at the right and bottom
    edges (112x80 is 7x5 16x16 areas), a single 16-row picture (no lower
    row), CTU32-row slices, 10 bit."""
    _i_encode_case(pkg, w, h, qp, ns, bd, 5)


def _i_encode_case(pkg, w, h, qp, ns, bd, ctu_log2):
    torch = _torch()
    say(f"hevc I {w}x{h} qp={qp} slices={ns} bd={bd} ctu_log2={ctu_log2}")
    (cy, cuv), = _frames(pkg, 5 + qp, w, h, 1, 3)
    hc, wc = cy.shape
    t = (wc // 16) * (hc // 16)
    if bd == 10:
        rs = np.random.default_rng(qp + 1)
        cy, cuv = _ten(cy, rs), _ten(cuv, rs)
    cb, cr = orc.split_uv(cuv)
    want = orc.hevc_encode_i(cy, cb, cr, qp, ns, orc.HEVC_ROUND_INTRA_I, bd, ctu_log2)
    if bd == 10:
        d, pitch, back = [_p010(torch, a) for a in (cy, cuv)], wc * 2, _unp010
    else:
        d, pitch, back = [dev(torch, a) for a in (cy, cuv)], wc, host
    oy, ouv = torch.zeros_like(d[0]), torch.zeros_like(d[1])
    modes = torch.zeros(t * 16, dtype=torch.uint8, device="cuda")
    coef = torch.zeros(t * 384, dtype=torch.int16, device="cuda")
    nwork = pkg.hip.bc250_hevc_intra_work_bytes(wc, hc)
    sync = torch.zeros((nwork + 3) // 4, dtype=torch.int32, device="cuda")
    if ctu_log2 == 4:
        assert pkg.hip.bc250_hevc_launch_i_encode(P(d[0]), P(d[1]), pitch, pitch, P(oy), P(ouv), pitch, pitch, wc, hc,
                                                  qp, ns, orc.HEVC_ROUND_INTRA_I, P(modes), P(coef), P(sync), 1 << 22,
                                                  bd, None) == 0
    else:
        assert pkg.hip.bc250_hevc_launch_i_encode_ctu(P(d[0]), P(d[1]), pitch, pitch, P(oy), P(ouv), pitch, pitch,
                                                      wc, hc, qp, ns, orc.HEVC_ROUND_INTRA_I, P(modes), P(coef),
                                                      P(sync), 1 << 22, ctu_log2, bd, None) == 0
    torch.cuda.synchronize()
    assert host(sync)[1] == 0, "wavefront wait gave up"
    assert np.array_equal(host(modes), want["modes"])
    assert np.array_equal(host(coef), want["coef"])
    assert np.array_equal(back(oy), want["y"])
    gb, gr = orc.split_uv(back(ouv))
    assert np.array_equal(gb, want["cb"]) and np.array_equal(gr, want["cr"])


def _hevc_sequence(pkg, w, h, n, qp, ns, rng, lam, seed, nrect, gop=0, deblock=True, bd=8, ctu_log2=4):
    say(f"hevc pipeline {w}x{h} frames={n} qp={qp} slices={ns} R={rng} deblock={deblock} bd={bd} ctu={ctu_log2}")
    g = pkg.FrameGen(seed, w, h, nrect)
    enc = pkg.Encoder(w, h, search_range=rng, lambda8=lam, deblock_idc=0 if deblock else 1,
                      codec=pkg.GPU_CODEC_HEVC, bit_depth=bd, ctu_log2=ctu_log2)
    rs = np.random.default_rng(seed)
    ref = None
    try:
        for i in range(n):
            vy, vuv = g.frame(i)
            if bd == 10:
                vy, vuv = _ten(vy, rs), _ten(vuv, rs)
            im = enc.new_image()
            enc.upload(im, vy, vuv)
            intra = i == 0 or (gop and i % gop == 0)
            slot = enc.encode(im, qp, intra, ns)
            got = enc.results(slot)
            cy, cuv = pkg.pad_coded(vy, vuv)
            ry, ruv = (cy, cuv) if ref is None else ref
            exp = orc.hevc_encode_frame(cy, cuv, ry, ruv, qp, intra, ns, rng, lam, deblock, bd, ctu_log2)
            assert np.array_equal(got["mv8"], exp["mv8"]), f"frame {i} mv8"
            assert np.array_equal(got["hevc_mode"], exp["modes"]), f"frame {i} modes"
            assert np.array_equal(got["quant"], exp["coef"]), f"frame {i} coef"
            my, muv = enc.map_recon()
            assert np.array_equal(my, exp["y"]) and np.array_equal(muv, exp["uv"]), f"frame {i} recon"
            assert enc.error_flags() == 0
            ref = (exp["y"], exp["uv"])
            say(f"  frame {i} ok ({'I' if intra else 'P'})")
    finally:
        enc.close()
        g.close()


def test_hevc_pipeline_small_ipppp(pkg):
    _torch()
    _hevc_sequence(pkg, 128, 96, 5, 27, 1, 16, 1, seed=31, nrect=3)


def test_hevc_pipeline_no_deblock(pkg):
    """deblock_idc 1: the reference encoder's setting (filter disabled in the PPS)."""
    _torch()
    _hevc_sequence(pkg, 128, 96, 4, 32, 2, 16, 1, seed=33, nrect=3, deblock=False)


def test_hevc_pipeline_odd_size_slices_gop(pkg):
    _torch()
    _hevc_sequence(pkg, 200, 120, 5, 37, 3, 32, 2, seed=32, nrect=2, gop=3)


def test_hevc_pipeline_main10(pkg):
    """10-bit P010 pictures (4 v + dither, as the 4K Main10 workload makes
    them), odd size, slices, a GOP."""
    _torch()
    _hevc_sequence(pkg, 200, 120, 5, 27, 3, 16, 1, seed=41, nrect=2, gop=3, bd=10)


def test_hevc_pipeline_ctu32_slices_gop(pkg):
    """CTU32: odd size (partial CTU32s), CTU32-row slices, I pictures in a GOP."""
    _torch()
    _hevc_sequence(pkg, 200, 120, 5, 32, 3, 16, 1, seed=43, nrect=2, gop=3, ctu_log2=5)


def test_hevc_pipeline_ctu32_main10(pkg):
    _torch()
    _hevc_sequence(pkg, 240, 144, 4, 27, 2, 16, 1, seed=44, nrect=3, gop=2, bd=10, ctu_log2=5)


def test_hevc_pipeline_main10_no_deblock(pkg):
    _torch()
    _hevc_sequence(pkg, 128, 96, 3, 37, 1, 16, 1, seed=42, nrect=3, deblock=False, bd=10)


def test_hevc_main10_4k_full_size(pkg):
    """The 4K Main10 workload at its real shape: 3840x2160 P010, CTU32, CQP
    27, 4 slices, +-64 search, deblocking.  The I picture is compared whole with
    the oracle.  For the P picture the oracle codes the frame with the GPU's
    vectors (levels and deblocked reconstruction must match exactly), and a
    spread sample of 64 MBs is searched and refined by the oracle from
    scratch (vectors and SADs must match)."""
    _torch()
    w, h, qp, ns, rng = 3840, 2160, 27, 4, 64
    say("hevc main10 4k: 2 frames")
    g = pkg.FrameGen(2160, w, h, 12)
    rs = np.random.default_rng(2160)
    vis = [tuple(_ten(a, rs) for a in g.frame(i)) for i in range(2)]
    g.close()
    enc = pkg.Encoder(w, h, search_range=rng, lambda8=1, codec=pkg.GPU_CODEC_HEVC, bit_depth=10, ctu_log2=5)
    try:
        cods = [pkg.pad_coded(*v) for v in vis]
        im = enc.new_image()
        enc.upload(im, *vis[0])
        got = enc.results(enc.encode(im, qp, True, ns))
        cy, cuv = cods[0]
        exp = orc.hevc_encode_frame(cy, cuv, cy, cuv, qp, True, ns, rng, 1, True, 10, 5)
        assert np.array_equal(got["hevc_mode"], exp["modes"]) and np.array_equal(got["quant"], exp["coef"])
        my, muv = enc.map_recon()
        assert np.array_equal(my, exp["y"]) and np.array_equal(muv, exp["uv"]), "I recon"
        ry, ruv = exp["y"], exp["uv"]
        say("  I ok")
        im = enc.new_image()
        enc.upload(im, *vis[1])
        got = enc.results(enc.encode(im, qp, False, ns))
        assert enc.error_flags() == 0
        cy, cuv = cods[1]
        mv8 = got["mv8"]
        mv = np.stack([mv8["mvx"], mv8["mvy"]], 1).astype(np.int32)
        cb, cr = orc.split_uv(cuv)
        fb, fr = orc.split_uv(ruv)
        o = orc.hevc_encode_p(cy, cb, cr, ry, fb, fr, qp, mv, orc.HEVC_ROUND_INTER, 10, 5)
        assert np.array_equal(got["quant"], o["coef"]), "P levels"
        nt = o["culog2"].size // 4
        assert np.array_equal(got["hevc_mode"].reshape(nt, 16)[:, :4].reshape(-1), o["culog2"]), "P CU sizes"
        say(f"  P CU8 / CU16 / CU32 share: {[int((o['culog2'] == L).sum()) for L in (3, 4, 5)]}")
        uv = np.empty_like(cuv)
        uv[:, 0::2], uv[:, 1::2] = o["cb"], o["cr"]
        mv4 = np.zeros((mv.shape[0], 4), np.int32)
        mv4[:, :2] = mv
        wy, wuv = orc.hevc_deblock(o["y"], uv, qp, 0, orc.hevc_cu_cbf(o["coef"], o["culog2"], w), mv4, ns, 10, 5,
                                   o["culog2"])
        my, muv = enc.map_recon()
        assert np.array_equal(my, wy) and np.array_equal(muv, wuv), "P recon"
        c8, r8 = (cy >> 2).astype(np.uint8), (ry >> 2).astype(np.uint8)
        mbw, nmb = w // 16, (w // 16) * ((h + 15) // 16)
        for m in np.linspace(0, nmb - 1, 64).astype(int):
            _, i8 = orc.me_fullsearch_mb(c8, r8, rng, 1, int(m % mbw), int(m // mbw))
            for z in range(4):
                x0, y0 = (m % mbw) * 16 + (z & 1) * 8, (m // mbw) * 16 + (z >> 1) * 8
                mx, my_, sad = orc.hevc_refine8(cy, ry, int(x0), int(y0), int(i8[z]["dx"]) * 4,
                                                int(i8[z]["dy"]) * 4, 10)
                cu = 4 * m + z
                assert (mv8[cu]["mvx"], mv8[cu]["mvy"], mv8[cu]["sad"]) == (mx, my_, sad), (m, z)
        say("  P ok")
    finally:
        enc.close()


def test_hevc_pipeline_640x480_cqp32_4slices(pkg):
    """The 640x480 workload's shape: CQP 32, 4 slices, +-32 full search."""
    _torch()
    _hevc_sequence(pkg, 640, 480, 3, 32, 4, 32, 1, seed=2024, nrect=6)


@pytest.mark.parametrize("log2n,bd", [(4, 8), (5, 8), (4, 10), (5, 10), (3, 8), (2, 8)])
def test_tqn_dqitn_batch(pkg, log2n, bd):
    torch = _torch()
    n2 = 1 << (2 * log2n)
    rs = np.random.default_rng(log2n * 100 + bd)
    vmax = (1 << bd) - 1
    n = 97
    res = rs.integers(-vmax, vmax + 1, (n, n2)).astype(np.int16)
    res[1::3] //= 17
    for qp in ([0, 22, 27, 37, 51] if bd == 8 else [12, 39, 63]):
        say(f"tqn log2n={log2n} bd={bd} qp={qp}")
        dres, dout = dev(torch, res), torch.zeros((n, n2), dtype=torch.int16, device="cuda")
        assert pkg.hip.bc250_hevc_tqn_batch(P(dres), n, log2n, qp, bd, 2, P(dout), None) == 0
        dback = torch.zeros_like(dout)
        assert pkg.hip.bc250_hevc_dqitn_batch(P(dout), n, log2n, qp, bd, P(dback), None) == 0
        torch.cuda.synchronize()
        lv, back = host(dout), host(dback)
        for b in range(0, n, 4):
            want = orc.hevc_tqn(res[b], log2n, qp, bd, 2)
            assert np.array_equal(lv[b], want), (b, qp)
            assert np.array_equal(back[b], orc.hevc_dqitn(want, log2n, qp, bd)), (b, qp)


@pytest.mark.parametrize("method", ["butterfly", "mfma", "matrix"])
@pytest.mark.parametrize("bd", [8, 10])
def test_t32_forms_match_oracle(pkg, method, bd):
    """The 32x32 forms (VALU butterflies, split-int8 matrix cores, LDS matrix
    product) against the oracle's plain matrix product: residual-range and
    full-int16 inputs (the saturating cases), odd batch sizes (the butterfly
    pairs blocks)."""
    torch = _torch()
    m = {"butterfly": pkg.T32_BUTTERFLY, "mfma": pkg.T32_MFMA, "matrix": pkg.T32_MATRIX}[method]
    rs = np.random.default_rng(3200 + bd)
    vmax = (1 << bd) - 1
    for n in (1, 3, 257):
        res = rs.integers(-vmax, vmax + 1, (n, 1024)).astype(np.int16)
        res[1::3] //= 9
        if n > 3:
            res[2] = rs.integers(-32768, 32768, 1024)     # saturating first pass
            res[5] = 32767
            res[6] = -32768
        for qp in ([0, 27, 51] if bd == 8 else [12, 39, 63]):
            say(f"t32 {method} bd={bd} n={n} qp={qp}")
            dres, dout = dev(torch, res), torch.zeros((n, 1024), dtype=torch.int16, device="cuda")
            assert pkg.hip.bc250_hevc_tq32_batch(P(dres), n, qp, bd, 2, m, P(dout), None) == 0
            coef = rs.integers(-32768, 32768, (n, 1024)).astype(np.int16)
            coef[::2] //= 311
            dcoef, dback = dev(torch, coef), torch.zeros_like(dout)
            assert pkg.hip.bc250_hevc_dqit32_batch(P(dcoef), n, qp, bd, m, P(dback), None) == 0
            torch.cuda.synchronize()
            lv, back = host(dout), host(dback)
            for b in sorted({0, n - 1} | set(range(0, n, 16)) | ({2, 5, 6} if n > 3 else set())):
                assert np.array_equal(lv[b], orc.hevc_tqn(res[b], 5, qp, bd, 2)), (b, qp)
                assert np.array_equal(back[b], orc.hevc_dqitn(coef[b], 5, qp, bd)), (b, qp)
    assert pkg.hip.bc250_hevc_tq32_batch(None, 0, 27, bd, 2, m, None, None) == 0
    assert pkg.hip.bc250_hevc_tq32_batch(None, 4, 27, 9, 2, m, None, None) == -1
    assert pkg.hip.bc250_hevc_dqit32_batch(None, 4, 27, bd, 7, None, None) == -1


def _blocky(rs, h, w, step):
    """Smooth ramps with per-8x8 offsets (block edges the filter acts on) and
    light noise; plane values stay inside 0..255."""
    yy, xx = np.mgrid[0:h, 0:w]
    base = 100 + (xx * 40) // max(w, 1) + (yy * 30) // max(h, 1)
    off = rs.integers(-step, step + 1, (h // 8 + 1, w // 8 + 1))
    v = base + off[yy // 8, xx // 8] + rs.integers(-1, 2, (h, w))
    return np.clip(v, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("w,h,qp,intra,ns", [(64, 48, 37, 1, 1), (64, 48, 30, 0, 1), (192, 112, 45, 1, 2),
                                             (192, 112, 27, 0, 3), (320, 240, 51, 0, 4), (320, 240, 22, 1, 1)])
def test_deblock(pkg, w, h, qp, intra, ns):
    torch = _torch()
    say(f"hevc deblock {w}x{h} qp={qp} intra={intra} slices={ns}")
    rs = np.random.default_rng(w * 7 + qp)
    y = _blocky(rs, h, w, 12)
    uv = _blocky(rs, h // 2, w, 10)
    ncu = (w // 8) * (h // 8)
    cbf = (rs.random(ncu) < 0.3).astype(np.uint8)
    mv4 = np.zeros((ncu, 4), np.int32)
    mv4[:, :2] = rs.integers(-6, 7, (ncu, 2))
    mv4[::3, :2] = 0
    wy, wuv = orc.hevc_deblock(y, uv, qp, intra, cbf, mv4, ns)
    assert not np.array_equal(wy, y)
    pitch = w + 64          # a pitch wider than the picture; the margin must stay untouched
    gy = np.full((h, pitch), 7, np.uint8)
    guv = np.full((h // 2, pitch), 9, np.uint8)
    gy[:, :w], guv[:, :w] = y, uv
    d = [dev(torch, a) for a in (gy, guv, cbf, mv4)]
    assert pkg.hip.bc250_hevc_launch_deblock(P(d[0]), P(d[1]), pitch, pitch, w, h, qp, intra, P(d[2]), P(d[3]), ns,
                                             8, None) == 0
    torch.cuda.synchronize()
    oy, ouv = host(d[0]), host(d[1])
    assert np.array_equal(oy[:, :w], wy)
    assert np.array_equal(ouv[:, :w], wuv)
    assert (oy[:, w:] == 7).all() and (ouv[:, w:] == 9).all()
    assert pkg.hip.bc250_hevc_launch_deblock(P(d[0]), P(d[1]), pitch, pitch, w + 8, h, qp, intra, P(d[2]), P(d[3]),
                                             ns, 8, None) == -1


@pytest.mark.parametrize("w,h,qp,intra,ns", [(192, 112, 45, 1, 2), (320, 240, 27, 0, 3), (208, 144, 37, 1, 3)])
def test_deblock_ctu32_slices(pkg, w, h, qp, intra, ns):
    """Slices as CTU32-row ranges: no filtering across their borders only."""
    torch = _torch()
    say(f"hevc deblock ctu32 {w}x{h} qp={qp} intra={intra} slices={ns}")
    rs = np.random.default_rng(w * 11 + qp)
    y = _blocky(rs, h, w, 12)
    uv = _blocky(rs, h // 2, w, 10)
    ncu = (w // 8) * (h // 8)
    cbf = (rs.random(ncu) < 0.3).astype(np.uint8)
    mv4 = np.zeros((ncu, 4), np.int32)
    mv4[:, :2] = rs.integers(-6, 7, (ncu, 2))
    wy, wuv = orc.hevc_deblock(y, uv, qp, intra, cbf, mv4, ns, 8, 5)
    assert not np.array_equal(wy, orc.hevc_deblock(y, uv, qp, intra, cbf, mv4, ns, 8, 4)[0])
    d = [dev(torch, a) for a in (y, uv, cbf, mv4)]
    assert pkg.hip.bc250_hevc_launch_deblock_ctu(P(d[0]), P(d[1]), w, w, w, h, qp, intra, P(d[2]), P(d[3]), None,
                                                 ns, 5, 8, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(host(d[0]), wy)
    assert np.array_equal(host(d[1]), wuv)
    assert pkg.hip.bc250_hevc_launch_deblock_ctu(P(d[0]), P(d[1]), w, w, w, h, qp, intra, P(d[2]), P(d[3]), None,
                                                 ns, 6, 8, None) == -1


@pytest.mark.parametrize("w,h,qp,intra,ns", [(64, 48, 37, 1, 1), (192, 112, 30, 0, 2), (320, 240, 45, 1, 3),
                                             (128, 96, 22, 0, 1)])
def test_deblock_p010(pkg, w, h, qp, intra, ns):
    """10-bit deblocking on P010 planes (sample << 6) against the oracle's
    10-bit filter on plain samples."""
    torch = _torch()
    say(f"hevc deblock p010 {w}x{h} qp={qp} intra={intra} slices={ns}")
    rs = np.random.default_rng(w * 5 + qp)
    y = _blocky(rs, h, w, 12).astype(np.uint16) * 4 + rs.integers(0, 4, (h, w)).astype(np.uint16)
    uv = _blocky(rs, h // 2, w, 10).astype(np.uint16) * 4 + rs.integers(0, 4, (h // 2, w)).astype(np.uint16)
    ncu = (w // 8) * (h // 8)
    cbf = (rs.random(ncu) < 0.3).astype(np.uint8)
    mv4 = np.zeros((ncu, 4), np.int32)
    mv4[:, :2] = rs.integers(-6, 7, (ncu, 2))
    mv4[::3, :2] = 0
    wy, wuv = orc.hevc_deblock(y, uv, qp, intra, cbf, mv4, ns, 10)
    assert not np.array_equal(wy, y)
    pitch = w + 32          # samples; a margin that must stay untouched
    gy = np.full((h, pitch), 7 << 6, np.uint16)
    guv = np.full((h // 2, pitch), 9 << 6, np.uint16)
    gy[:, :w], guv[:, :w] = y << 6, uv << 6
    d = [dev(torch, a) for a in (gy.view(np.int16), guv.view(np.int16), cbf, mv4)]
    assert pkg.hip.bc250_hevc_launch_deblock(P(d[0]), P(d[1]), pitch * 2, pitch * 2, w, h, qp, intra, P(d[2]),
                                             P(d[3]), ns, 10, None) == 0
    torch.cuda.synchronize()
    oy, ouv = host(d[0]).view(np.uint16), host(d[1]).view(np.uint16)
    assert np.array_equal(oy[:, :w] >> 6, wy) and np.array_equal(ouv[:, :w] >> 6, wuv)
    assert not (oy[:, :w] & 63).any() and not (ouv[:, :w] & 63).any()
    assert (oy[:, w:] == 7 << 6).all() and (ouv[:, w:] == 9 << 6).all()


def test_cu_cbf(pkg):
    torch = _torch()
    rs = np.random.default_rng(4)
    nctu = 37
    coef = np.zeros((nctu, 384), np.int16)
    for cu in range(nctu * 4):
        if rs.random() < 0.5:
            coef[cu // 4, (cu % 4) * 64 + rs.integers(0, 64)] = rs.integers(1, 9) * (1 if rs.random() < .5 else -1)
    coef[:, 256:] = 5          # chroma levels never count
    dc, dcbf = dev(torch, coef), torch.full((nctu * 4,), 7, dtype=torch.uint8, device="cuda")
    assert pkg.hip.bc250_hevc_cu_cbf(P(dc), nctu, P(dcbf), None) == 0
    torch.cuda.synchronize()
    want = orc.hevc_cu_cbf(coef)
    assert np.array_equal(host(dcbf), want) and 0 < want.sum() < nctu * 4
