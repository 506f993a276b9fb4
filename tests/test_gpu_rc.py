"""Rate-controlled H.264 sequences on the GPU (bc250_encode_sequence_rc):
every picture's QP is what the host controller gives when replayed on the
CPU with the bits the driver reported, in the driver's call order
(pipelined: picture i's QP before picture i-1's bits); the stream lands
near its bitrate; and a small stream whose QP changes from picture to
picture parses (slice QP = 26 + pic_init_qp_minus26 + slice_qp_delta) and
decodes to the encoder's reconstruction."""
import numpy as np
import pytest

import orc
from conftest import have_gpu, load_pkg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    if not have_gpu():
        pytest.skip("no GPU")
    return load_pkg()


@pytest.fixture(autouse=True)
def _no_qp_min_env(monkeypatch):
    monkeypatch.delenv("BC250_QP_MIN", raising=False)


def say(m):
    print(m, flush=True)


def _replay(pkg, mk, qps, bits, pipelined, gop):
    r = mk()
    n = len(qps)
    for i in range(n):
        intra = i == 0 or (gop > 0 and i % gop == 0)
        q = r.frame_qp(0, intra, False)
        assert q == qps[i], f"picture {i}: driver QP {qps[i]}, replay {q}"
        if pipelined and i >= 1:
            r.coded(int(bits[i - 1]), int(qps[i - 1]), i - 1 == 0 or (gop > 0 and (i - 1) % gop == 0), False)
        elif not pipelined:
            r.coded(int(bits[i]), int(qps[i]), intra, False)
    if pipelined:
        r.coded(int(bits[-1]), int(qps[-1]), (n - 1) == 0 or (gop > 0 and (n - 1) % gop == 0), False)
    return r


@pytest.mark.parametrize("model,pipelined", [(False, 1), (True, 0), (True, 1)])
def test_rc_sequence_replays_and_holds_the_rate(pkg, model, pipelined):
    w, h, n, gop, fps, rate = 1920, 1080, 48, 24, 60.0, 20_000_000
    say(f"rc: 1080p low-latency {rate / 1e6:.0f} Mbit/s, model {model}, pipelined {pipelined}")
    g = pkg.FrameGen(1080, w, h, 12)
    enc = pkg.Encoder(w, h, search_range=64, lambda8=1, deblock_idc=0, readback=0, entropy=1)
    try:
        ims = []
        for i in range(n):
            im = enc.new_image()
            enc.upload(im, *g.frame(i))
            ims.append(im)
        mk = lambda: pkg.RateControl(pkg.RC_LOW_LATENCY, rate, fps, w, h, model=model)
        stream, qps, bits, _, _ = enc.encode_sequence_rc(ims, mk(), gop=gop, num_slices=4, pipelined=pipelined)
    finally:
        enc.close()
        g.close()
    assert len(stream) == int(bits.sum()) // 8
    _replay(pkg, mk, qps, bits, pipelined, gop)
    assert len(set(qps.tolist())) > 1
    got = bits[12:].mean() * fps
    say(f"  QP {qps.min()}..{qps.max()}, {got / 1e6:.2f} Mbit/s over pictures 12..{n - 1}")
    if model:
        # the model aims at the rate: within a tenth after the settling
        assert abs(got / rate - 1) < 0.1, got
    else:
        # the buffer loop's two-picture bucket clips an I picture's
        # overspend, so it runs over on this content (as the reference's
        # does); it still steers within a factor of two
        assert 0.5 < got / rate < 2.0, got


def test_rc_stream_parses_and_decodes(pkg):
    import os
    if not os.path.exists(orc.REF_H264):
        pytest.skip("oracle/_ref not built")
    from h264_cavlc_parse import SliceParser
    from h264_decode import decode_frame
    from test_nal import unescape
    w, h, n, ns = 320, 192, 10, 2
    wmb, hmb = w // 16, h // 16
    say("rc: 320x192 stream with per-picture QPs")
    g = pkg.FrameGen(21, w, h, 4)
    enc = pkg.Encoder(w, h, search_range=32, lambda8=1, deblock_idc=1, readback=0, entropy=1)
    try:
        enc.set_slice_header(log2_max_frame_num_minus4=4, log2_max_poc_lsb_minus4=4, pic_init_qp_minus26=0)
        ims = []
        for i in range(n):
            im = enc.new_image()
            enc.upload(im, *g.frame(i))
            ims.append(im)
        rc = pkg.RateControl(pkg.RC_LOW_LATENCY, 400_000, 30.0, w, h)
        stream, qps, bits, _, _ = enc.encode_sequence_rc(ims, rc, num_slices=ns, pipelined=0)
        final_y, final_uv = enc.map_recon()
    finally:
        enc.close()
        g.close()
    assert len(set(qps.tolist())) > 2, qps
    nals = stream.split(b"\x00\x00\x00\x01")[1:]
    assert len(nals) == n * ns
    ry = ruv = None
    for i in range(n):
        ps = SliceParser(wmb, hmb, ns, 4, 4)
        for s in range(ns):
            ps.parse(s, unescape(nals[i * ns + s][1:]))
        sq = {26 + hd["slice_qp_delta"] for hd in ps.headers}
        assert sq == {int(qps[i])}, (i, sq, qps[i])
        ry, ruv = decode_frame(ps, int(qps[i]), ry, ruv)
    np.testing.assert_array_equal(ry, final_y)
    np.testing.assert_array_equal(ruv, final_uv)
