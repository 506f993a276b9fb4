"""The GPU H.264 CAVLC slice writer (include/bc250_bitstream.h) against the
CPU oracle: every slice's RBSP byte for byte.  Synthetic staging data with
escapes, all-skip and all-zero frames, up to 64 slices and one-MB-wide
pictures; oracle-encoded frames; the real pipeline's 1080p staging, I then
P; output bounds and argument checks."""
import ctypes as C

import numpy as np
import pytest

import orc
from conftest import have_gpu, load_pkg
from test_oracle_cavlc import synth_frame

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def pkg():
    if not have_gpu():
        pytest.skip("no GPU")
    return load_pkg()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _staging(torch, lv, dc, modes, mvs, n):
    mv = np.zeros(n, orc.MV) if mvs is None else np.ascontiguousarray(mvs, orc.MV)
    return (_dev(torch, np.ascontiguousarray(lv, np.int16).reshape(-1)),
            _dev(torch, np.ascontiguousarray(dc, np.int32).reshape(-1)),
            _dev(torch, np.ascontiguousarray(modes, np.uint32).view(np.int32)),
            _dev(torch, mv.view(np.int32).reshape(-1)))


def compare(pkg, lv, dc, modes, mvs, cfg, writer=None):
    import torch
    wmb, hmb, ns = int(cfg[0]), int(cfg[1]), int(cfg[4])
    n = wmb * hmb
    want = orc.h264_cavlc_slices(lv, dc, modes, mvs, cfg)
    w = writer or pkg.CavlcWriter(wmb, hmb)
    got = w.slices(*_staging(torch, lv, dc, modes, mvs, n), cfg)
    assert len(got) == ns
    for s in range(ns):
        if got[s] != want[s]:
            a, b = np.frombuffer(got[s], np.uint8), np.frombuffer(want[s], np.uint8)
            k = min(len(a), len(b))
            d = np.nonzero(a[:k] != b[:k])[0]
            raise AssertionError(f"slice {s}: {len(a)} vs {len(b)} bytes, first difference at "
                                 f"{d[0] if len(d) else k}")
    return got


@pytest.mark.parametrize("wmb,hmb,ns,idr", [(5, 4, 1, 1), (5, 4, 1, 0), (7, 3, 2, 1), (7, 3, 2, 0),
                                            (6, 5, 3, 0), (9, 4, 4, 1), (8, 6, 5, 0), (4, 4, 16, 0),
                                            (1, 7, 2, 0), (1, 5, 1, 1), (13, 9, 64, 0), (13, 9, 64, 1)])
def test_synthetic_frames(pkg, wmb, hmb, ns, idr):
    say(f"synthetic {wmb}x{hmb} MBs, {ns} slices, idr {idr}")
    rs = np.random.default_rng(wmb * 100 + hmb * 10 + ns + idr)
    lv, dc, modes, mvs = synth_frame(rs, wmb, hmb, idr)
    cfg = orc.slice_cfg(wmb, hmb, 27, idr, ns, pps_id=1, pic_init_qp_minus26=-3, frame_num=5,
                        log2_max_frame_num_minus4=2, idr_pic_id=3, poc=14, log2_max_poc_lsb_minus4=1,
                        deblock_idc=ns % 3, relax_chroma_skip=ns & 1)
    compare(pkg, lv, dc, modes, mvs, cfg)


@pytest.mark.parametrize("qp,idr", [(0, 1), (12, 0), (22, 1), (37, 0), (51, 1)])
def test_escapes_and_qps(pkg, qp, idr):
    say(f"big levels qp {qp} idr {idr}")
    rs = np.random.default_rng(qp + 100 * idr)
    lv, dc, modes, mvs = synth_frame(rs, 11, 6, idr, density=0.6, big=True)
    lv[3] = rs.integers(-32767, 32768, (24, 16))  # every coefficient of one MB at the int16 extremes
    compare(pkg, lv, dc, modes, mvs, orc.slice_cfg(11, 6, qp, idr, 3))


@pytest.mark.parametrize("ns,idr", [(4, 0), (16, 1)])
def test_4k_frames(pkg, ns, idr):
    """3840x2160 (240 x 135 MBs): the scan's threads hold 32 MBs each."""
    say(f"4K synthetic, {ns} slices, idr {idr}")
    rs = np.random.default_rng(2160 + ns)
    lv, dc, modes, mvs = synth_frame(rs, 240, 135, idr, density=0.05)
    compare(pkg, lv, dc, modes, mvs, orc.slice_cfg(240, 135, 30, idr, ns))


def test_largest_frame_and_the_limit(pkg):
    """2^17 MBs (512 x 256) is accepted -- an all-skip P frame, 128 MBs per
    scan thread -- and one more MB row is refused."""
    import torch
    say("2^17 MBs")
    wmb, hmb = 512, 256
    n = wmb * hmb
    z = np.zeros((n, 24, 16), np.int16)
    compare(pkg, z, np.zeros((n, 24), np.int32), np.zeros(n, np.uint32), np.zeros(n, orc.MV),
            orc.slice_cfg(wmb, hmb, 30, 0, 3))
    t = torch.zeros(4096, dtype=torch.int32, device="cuda")
    p = C.c_void_p(t.data_ptr())
    cfg = pkg.h264_slice_cfg(wmb, hmb + 1, 30, 0, 1)
    assert pkg.hip.bc250_h264_cavlc_slices(p, p, p, p, cfg.ctypes.data_as(C.c_void_p), p, 4096, p, p, 1 << 40,
                                           None) == -1


def test_all_skip_and_all_zero(pkg):
    say("all-skip P frame, all-zero I frame")
    wmb, hmb = 12, 7
    n = wmb * hmb
    z = np.zeros((n, 24, 16), np.int16)
    compare(pkg, z, np.zeros((n, 24), np.int32), np.zeros(n, np.uint32), np.zeros(n, orc.MV),
            orc.slice_cfg(wmb, hmb, 30, 0, 2))
    compare(pkg, z, np.zeros((n, 24), np.int32), np.full(n, 15, np.uint32), None, orc.slice_cfg(wmb, hmb, 30, 1, 2))


def test_oracle_encoded_frames_one_writer(pkg):
    """I then two P frames of a moving picture through one writer (the
    buffers carry the previous frame's bytes)."""
    say("oracle-encoded 320x192 IPP")
    rs = np.random.default_rng(23)
    w, h = 320, 192
    y = rs.integers(30, 220, (h, w)).astype(np.uint8)
    y = ((y.astype(int) + np.roll(y, 1, 1) + np.roll(y, 1, 0) + np.roll(y, 1, (0, 1))) // 4).astype(np.uint8)
    uv = rs.integers(90, 160, (h // 2, w)).astype(np.uint8)
    writer = pkg.CavlcWriter(w // 16, h // 16)
    f = orc.encode_frame(y, uv, y, uv, 24, True, 3)
    compare(pkg, f["levels"], f["dc"], f["modes"], None, orc.slice_cfg(w // 16, h // 16, 24, 1, 3), writer)
    ry, ruv = f["y"], f["uv"]
    for k in (1, 2):
        y2, uv2 = np.roll(y, (k, 2 * k), (0, 1)), np.roll(uv, (0, 2 * k), (0, 1))
        f = orc.encode_frame(y2, uv2, ry, ruv, 24, False, 3, rng=16)
        compare(pkg, f["levels"], f["dc"], f["modes"], f["mvs"],
                orc.slice_cfg(w // 16, h // 16, 24, 0, 3, frame_num=k, poc=2 * k), writer)
        ry, ruv = f["y"], f["uv"]


def test_pipeline_1080p_staging(pkg):
    """The H.264 pipeline's own staging of a 1080p I and P frame (+-64)."""
    import torch
    say("1080p pipeline staging")
    w, h, qp = 1920, 1080, 27
    g = pkg.FrameGen(1080, w, h, 12)
    enc = pkg.Encoder(w, h, search_range=64, lambda8=1, deblock_idc=0)
    try:
        ims = []
        for f in range(2):
            im = enc.new_image()
            enc.upload(im, *g.frame(f))
            ims.append(im)
        wmb, hmb = 120, 68
        writer = pkg.CavlcWriter(wmb, hmb)
        for f, im in enumerate(ims):
            slot = enc.encode(im, qp, f == 0, 4)
            r = enc.results(slot)
            cfg = orc.slice_cfg(wmb, hmb, qp, int(f == 0), 4, frame_num=f, poc=2 * f)
            got = compare(pkg, r["quant"], r["dc"], r["pred_mode"], r["mv"], cfg, writer)
            say(f"  frame {f}: {sum(len(x) for x in got)} bytes")
    finally:
        enc.close()
        g.close()


def test_output_capacity_is_respected(pkg):
    """A buffer too small for the frame: the overflow flag is set and no
    byte past the capacity changes."""
    import torch
    say("capacity")
    rs = np.random.default_rng(31)
    wmb, hmb = 8, 6
    n = wmb * hmb
    lv, dc, modes, mvs = synth_frame(rs, wmb, hmb, 1, density=0.5, big=True)
    cfg = orc.slice_cfg(wmb, hmb, 20, 1, 2)
    want = orc.h264_cavlc_slices(lv, dc, modes, mvs, cfg)
    total = sum(len(x) for x in want)
    cap = (total // 2) // 4 * 4
    out = torch.full((total + 4096,), 0x5A, dtype=torch.uint8, device="cuda")
    info = torch.zeros(4 * 2 + 2, dtype=torch.int32, device="cuda")
    work = torch.empty(int(pkg.hip.bc250_h264_cavlc_work_size(wmb, hmb)), dtype=torch.uint8, device="cuda")
    st = _staging(torch, lv, dc, modes, mvs, n)
    cfg_c = np.ascontiguousarray(cfg, np.int32)
    p = lambda t: C.c_void_p(t.data_ptr())
    rc = pkg.hip.bc250_h264_cavlc_slices(p(st[0]), p(st[1]), p(st[2]), p(st[3]), cfg_c.ctypes.data_as(C.c_void_p),
                                         p(out), cap, p(info), p(work), work.numel(), None)
    assert rc == 0
    torch.cuda.synchronize()
    i = info.cpu().numpy().view(np.uint32)
    assert i[8] == total and i[9] == 1
    assert (out[cap:].cpu().numpy() == 0x5A).all()


def test_invalid_arguments_are_rejected(pkg):
    import torch
    say("argument checks")
    t = torch.zeros(4096, dtype=torch.int32, device="cuda")
    w = torch.zeros(1 << 20, dtype=torch.uint8, device="cuda")
    p = C.c_void_p(t.data_ptr())
    for bad in (dict(qp=52), dict(num_slices=0), dict(num_slices=65), dict(deblock_idc=3),
                dict(log2_max_frame_num_minus4=13), dict(is_idr=2), dict(pps_id=256), dict(idr_pic_id=65536),
                dict(pic_init_qp_minus26=26), dict(pic_init_qp_minus26=-27), dict(relax_chroma_skip=2),
                dict(poc=-1)):
        c = dict(width_in_mbs=2, height_in_mbs=2, qp=26, is_idr=0, num_slices=1)
        c.update(bad)
        cfg = pkg.h264_slice_cfg(**c)
        rc = pkg.hip.bc250_h264_cavlc_slices(p, p, p, p, cfg.ctypes.data_as(C.c_void_p), p, 4096, p,
                                             C.c_void_p(w.data_ptr()), w.numel(), None)
        assert rc == -1, bad
    cfg = pkg.h264_slice_cfg(2, 2, 26, 0, 1)
    assert pkg.hip.bc250_h264_cavlc_slices(p, p, p, None, cfg.ctypes.data_as(C.c_void_p), p, 4096, p,
                                           C.c_void_p(w.data_ptr()), w.numel(), None) == -1  # P frame without mvs
    assert pkg.hip.bc250_h264_cavlc_slices(p, p, p, p, cfg.ctypes.data_as(C.c_void_p), p, 4096, p,
                                           C.c_void_p(w.data_ptr()), 16, None) == -1  # work too small


@pytest.mark.parametrize("w,h,ns,rng", [(320, 192, 3, 16), (1920, 1080, 4, 64)])
def test_pipeline_entropy_slices(pkg, w, h, ns, rng):
    """gpu_encode_params_t.entropy = 1: each frame's slices, numbered IPPP
    with a second IDR, equal the oracle writer's on the frame's staging."""
    say(f"pipeline slices {w}x{h}, {ns} slices, +-{rng}")
    qp, wmb, hmb = 27, (w + 15) // 16, (h + 15) // 16
    g = pkg.FrameGen(7, w, h, 6)
    enc = pkg.Encoder(w, h, search_range=rng, lambda8=1, deblock_idc=0, readback=1, entropy=1)
    enc.set_slice_header(pps_id=2, pic_init_qp_minus26=1, log2_max_frame_num_minus4=0, log2_max_poc_lsb_minus4=1)
    try:
        fn, idr = 0, -1
        for i in range(7):
            intra = i in (0, 5)
            im = enc.new_image()
            enc.upload(im, *g.frame(i))
            slot = enc.encode(im, qp, intra, ns)
            got, r = enc.slices(slot), enc.results(slot)
            fn, idr = (0, idr + 1) if intra else ((fn + 1) % 16, idr)
            cfg = orc.slice_cfg(wmb, hmb, qp, int(intra), ns, pps_id=2, pic_init_qp_minus26=1, frame_num=fn,
                                idr_pic_id=idr, poc=2 * fn, log2_max_poc_lsb_minus4=1)
            want = orc.h264_cavlc_slices(r["quant"], r["dc"], r["pred_mode"], r["mv"], cfg)
            assert got == want, f"frame {i}"
            say(f"  frame {i}: {sum(len(x) for x in got)} bytes")
    finally:
        enc.close()
        g.close()


def test_pipeline_frame_num_and_poc_counters(pkg):
    """frame_num and POC are separate counters as in the reference (0 at an
    IDR, +1 and +2 per frame, written modulo 2^log2_max: encoder_h264.c:
    3102,3108,3526-3527), so with 4 frame_num bits and 8 POC bits the POC
    keeps rising after frame_num wraps at 16; 22 frames through the
    sequence driver, every frame checked against the oracle."""
    import seqcheck
    w, h, qp, n = 96, 64, 30, 22
    hdr = dict(log2_max_frame_num_minus4=0, log2_max_poc_lsb_minus4=4)
    g = pkg.FrameGen(21, w, h, 2)
    frames = [g.frame(i) for i in range(n)]
    g.close()
    enc = pkg.Encoder(w, h, search_range=16, lambda8=1, deblock_idc=0, readback=1, entropy=1)
    try:
        enc.set_slice_header(**hdr)
        images = []
        for y, uv in frames:
            im = enc.new_image()
            enc.upload(im, y, uv)
            images.append(im)
        _, _, got = enc.encode_sequence(images, qp, gop=0, pipelined=1, collect=True)
    finally:
        enc.close()
    seqcheck.check_h264([pkg.pad_coded(*f) for f in frames], got, qp, 1, 16, 1, sample=1000, slice_hdr=hdr)


def test_pipeline_slices_decode_to_recon(pkg):
    """The pipeline's own slices (entropy = 1, deblocking off), parsed and
    decoded with the reference decoder's prediction / MC / inverse
    transforms, rebuild the pipeline's reconstruction exactly, IPPPP."""
    import os
    if not os.path.exists(orc.REF_H264):
        pytest.skip("oracle/_ref not built")
    from h264_cavlc_parse import SliceParser
    from h264_decode import decode_frame
    w, h, qp, ns = 320, 192, 28, 2
    say(f"pipeline slices decode == recon {w}x{h}")
    g = pkg.FrameGen(11, w, h, 4)
    enc = pkg.Encoder(w, h, search_range=32, lambda8=1, deblock_idc=1, readback=0, entropy=1)
    try:
        ry = ruv = None
        for i in range(5):
            im = enc.new_image()
            enc.upload(im, *g.frame(i))
            slot = enc.encode(im, qp, i == 0, ns)
            ps = SliceParser(w // 16, h // 16, ns)
            for s, rb in enumerate(enc.slices(slot)):
                ps.parse(s, rb)
            dy, duv = decode_frame(ps, qp, ry, ruv)
            gy, guv = enc.map_recon()
            np.testing.assert_array_equal(dy, gy, err_msg=f"frame {i} luma")
            np.testing.assert_array_equal(duv, guv, err_msg=f"frame {i} chroma")
            ry, ruv = gy, guv
            say(f"  frame {i} decodes to the reconstruction")
    finally:
        enc.close()
        g.close()


@pytest.mark.parametrize("qp,idc", [(24, 0), (36, 0), (40, 2)])
def test_pipeline_slices_decode_to_recon_with_loop_filter(pkg, qp, idc):
    """Deblocking on (idc 0: across slice edges; 2: not): the pipeline's
    slices parsed, decoded with the reference decoder's primitives and
    filtered by a literal reading of clause 8.7 give the pipeline's
    (GPU-deblocked) reconstruction exactly, IPPPP with two slices; the
    filter moves samples, so the check is not vacuous."""
    import os
    if not os.path.exists(orc.REF_H264):
        pytest.skip("oracle/_ref not built")
    from h264_cavlc_parse import SliceParser
    from h264_decode import decode_frame, loop_filter
    w, h, ns = 320, 192, 2
    say(f"decode == recon with the loop filter, {w}x{h}, QP {qp}, idc {idc}")
    g = pkg.FrameGen(31, w, h, 4)
    enc = pkg.Encoder(w, h, search_range=32, lambda8=1, deblock_idc=idc, readback=0, entropy=1)
    try:
        ry = ruv = None
        moved = 0
        for i in range(5):
            im = enc.new_image()
            enc.upload(im, *g.frame(i))
            slot = enc.encode(im, qp, i == 0, ns)
            ps = SliceParser(w // 16, h // 16, ns)
            for s, rb in enumerate(enc.slices(slot)):
                ps.parse(s, rb)
            assert {hd["deblock_idc"] for hd in ps.headers} == {idc}
            dy, duv = decode_frame(ps, qp, ry, ruv)
            uy = dy.copy()
            loop_filter(ps, qp, dy, duv)
            moved += int(np.sum(uy != dy))
            gy, guv = enc.map_recon()
            np.testing.assert_array_equal(dy, gy, err_msg=f"frame {i} luma")
            np.testing.assert_array_equal(duv, guv, err_msg=f"frame {i} chroma")
            ry, ruv = gy, guv
        assert moved > 1000, moved
        say(f"  5 frames decode to the deblocked reconstruction ({moved} luma samples moved by the filter)")
    finally:
        enc.close()
        g.close()


def test_pipeline_annexb_stream_decodes(pkg):
    """SPS + PPS + the pipeline's slice NAL units of an IPPPP sequence form
    one Annex B stream; split at its start codes, un-escaped, parsed and
    decoded with the reference decoder's primitives, every frame equals the
    pipeline's reconstruction (deblocking off)."""
    import os
    if not os.path.exists(orc.REF_H264):
        pytest.skip("oracle/_ref not built")
    from h264_cavlc_parse import SliceParser, parse_pps, parse_sps
    from h264_decode import decode_frame
    from test_nal import unescape
    w, h, qp, ns = 320, 192, 28, 2
    wmb, hmb = w // 16, h // 16
    say("Annex B stream")
    g = pkg.FrameGen(13, w, h, 4)
    enc = pkg.Encoder(w, h, search_range=32, lambda8=1, deblock_idc=1, readback=0, entropy=1)
    enc.set_slice_header(log2_max_frame_num_minus4=4, log2_max_poc_lsb_minus4=4, pic_init_qp_minus26=qp - 26)
    stream = pkg.h264_parameter_sets(
        dict(profile_idc=66, level_idc=40, chroma_format_idc=1, bit_depth_luma=8, bit_depth_chroma=8,
             log2_max_frame_num_minus4=4, log2_max_poc_lsb_minus4=4, max_num_ref_frames=1, width_in_mbs=wmb,
             height_in_mbs=hmb, frame_mbs_only=1, direct_8x8_inference=1, vui_present=1, sar_width=1, sar_height=1,
             timing_info_present=1, num_units_in_tick=1, time_scale=60),
        dict(pic_init_qp_minus26=qp - 26, deblocking_filter_control=1))
    recons = []
    try:
        for i in range(5):
            im = enc.new_image()
            enc.upload(im, *g.frame(i))
            slot = enc.encode(im, qp, i == 0, ns)
            stream += pkg.h264_frame_nals(enc.slices(slot), i == 0)
            recons.append(enc.map_recon())
    finally:
        enc.close()
        g.close()
    nals = stream.split(b"\x00\x00\x00\x01")[1:]
    assert [n[0] & 31 for n in nals] == [7, 8] + [5] * ns + [1] * (4 * ns)
    # the parameter sets agree with the encoder's settings, and the slices
    # are parsed and decoded with what they say
    sps, pps = parse_sps(unescape(nals[0][1:])), parse_pps(unescape(nals[1][1:]))
    assert (sps["width_in_mbs"], sps["height_in_map_units"], sps["frame_mbs_only"]) == (wmb, hmb, 1)
    assert (sps["log2_max_frame_num_minus4"], sps["pic_order_cnt_type"], sps["log2_max_poc_lsb_minus4"]) == (4, 0, 4)
    assert (sps["profile_idc"], sps["level_idc"], sps["time_scale"], sps["sar"]) == (66, 40, 60, (1, 1))
    assert (pps["sps_id"], pps["entropy_coding_mode"], pps["pic_init_qp_minus26"]) == (sps["sps_id"], 0, qp - 26)
    assert pps["deblocking_filter_control"] == 1
    ry = ruv = None
    for i in range(5):
        ps = SliceParser(sps["width_in_mbs"], sps["height_in_map_units"], ns, sps["log2_max_frame_num_minus4"],
                         sps["log2_max_poc_lsb_minus4"])
        for s in range(ns):
            ps.parse(s, unescape(nals[2 + i * ns + s][1:]))
        slice_qps = {26 + pps["pic_init_qp_minus26"] + hd["slice_qp_delta"] for hd in ps.headers}
        assert slice_qps == {qp}, slice_qps
        assert all(hd["pps_id"] == pps["pps_id"] for hd in ps.headers)
        assert all(hd["frame_num"] == i and hd["poc_lsb"] == 2 * i for hd in ps.headers)
        dy, duv = decode_frame(ps, slice_qps.pop(), ry, ruv)
        np.testing.assert_array_equal(dy, recons[i][0], err_msg=f"frame {i} luma")
        np.testing.assert_array_equal(duv, recons[i][1], err_msg=f"frame {i} chroma")
        ry, ruv = recons[i]
    say(f"  {len(stream)} bytes, 5 frames decode to the reconstruction")
