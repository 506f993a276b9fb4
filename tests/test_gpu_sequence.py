"""The path bench.py times, checked: 1920x1080 frames of the bench's own
generator (seed 1080, 12 rectangles) through the host sequence driver
(bc250_encode_sequence_cb, pipelined over the two slots, readback on), a
warm-up sequence and then a longer one on the same context as the bench
runs them, every frame's outputs read in the driver's per-frame consumer and
checked against the oracle frame by frame (tests/seqcheck.py: the whole
frame's coding, slices and the +-64 search on a sample of MBs)."""
import numpy as np
import pytest

import seqcheck

pytestmark = pytest.mark.gpu

W, H, QP = 1920, 1080, 27


@pytest.fixture(scope="module")
def frames_1080p(pkg):
    g = pkg.FrameGen(1080, W, H, 12)
    try:
        return [g.frame(i) for i in range(8)]
    finally:
        g.close()


def _run(pkg, frames, codec, **kw):
    enc = pkg.Encoder(W, H, search_range=64, lambda8=1, deblock_idc=0, me_8x8=1, readback=1, codec=codec, **kw)
    try:
        images = []
        for y, uv in frames:
            im = enc.new_image()
            enc.upload(im, y, uv)
            images.append(im)
        print(f"encoding {len(frames)} frames, codec {codec}", flush=True)
        _, _, warm = enc.encode_sequence(images[:2], QP, gop=0, pipelined=1, collect=True)
        _, _, timed = enc.encode_sequence(images[2:], QP, gop=0, pipelined=1, collect=True)
        last = enc.map_recon()
        errs = enc.error_flags()
    finally:
        enc.close()
    return warm, timed, last, errs


def test_h264_benched_sequence_1080p(pkg, frames_1080p):
    warm, timed, (gy, guv), errs = _run(pkg, frames_1080p, pkg.GPU_CODEC_H264, entropy=1)
    assert errs == 0
    coded = [pkg.pad_coded(*f) for f in frames_1080p]
    seqcheck.check_h264(coded[:2], warm, QP, 1, 64, 1)
    y, uv = seqcheck.check_h264(coded[2:], timed, QP, 1, 64, 1, last_idr_pic_id=0, seed=1)
    assert np.array_equal(gy, y) and np.array_equal(guv, uv)


def test_hevc_benched_sequence_1080p(pkg, frames_1080p):
    warm, timed, (gy, guv), errs = _run(pkg, frames_1080p, pkg.GPU_CODEC_HEVC)
    assert errs == 0
    coded = [pkg.pad_coded(*f) for f in frames_1080p]
    seqcheck.check_hevc(coded[:2], warm, QP, 1, 64, 1)
    y, uv = seqcheck.check_hevc(coded[2:], timed, QP, 1, 64, 1, seed=1)
    assert np.array_equal(gy, y) and np.array_equal(guv, uv)


def test_nal_sink_writes_the_frames_slices(pkg, frames_1080p):
    """The C consumer bench.py times with H.264 slices: each frame's slices as
    Annex B NAL units, byte-equal to packing the slices the context hands out."""
    frames = frames_1080p[:4]
    buf = np.zeros(8 << 20, np.uint8)
    sizes = np.zeros(len(frames), np.uint64)
    sink = pkg.NalSink(0, buf.ctypes.data, buf.size, 0, sizes.ctypes.data)
    got = None
    for consumer in ("collect", "sink"):  # a fresh context each: the same frame_num / idr_pic_id
        enc = pkg.Encoder(W, H, search_range=64, lambda8=1, deblock_idc=0, readback=1, entropy=1)
        try:
            images = []
            for y, uv in frames:
                im = enc.new_image()
                enc.upload(im, y, uv)
                images.append(im)
            if consumer == "collect":
                _, _, got = enc.encode_sequence(images, QP, gop=2, num_slices=4, pipelined=1, collect=True)
            else:
                enc.encode_sequence(images, QP, gop=2, num_slices=4, pipelined=1, nal_sink=sink)
        finally:
            enc.close()
    want = [pkg.h264_frame_nals(g["slices"], g["is_intra"]) for g in got]
    assert [int(s) for s in sizes] == [len(w) for w in want]
    assert buf[:sink.used].tobytes() == b"".join(want)
