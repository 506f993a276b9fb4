"""The seeded synthetic-frame generator (bc250-vaapi_amd/host/framegen.c)
against its committed hashes (tests/golden/framegen_sha256.json, written by
tests/golden/make_framegen_hashes.py) and the value ranges it promises."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


def test_frames_match_committed_hashes():
    import make_framegen_hashes as mk
    with open(os.path.join(HERE, "golden", "framegen_sha256.json")) as f:
        want = json.load(f)
    assert mk.compute() == want


def test_frame_ranges_and_motion(pkg):
    g = pkg.FrameGen(2024, 640, 480, 6)
    try:
        y0, uv0 = g.frame(0)
        assert y0.shape == (480, 640) and uv0.shape == (240, 640)
        assert y0.min() >= 16 and y0.max() <= 235
        assert uv0.min() >= 108 and uv0.max() <= 148
        # the global motion is quarter-pel and at most 4 px per frame
        prev = g.offset_qpel(0)
        for f in range(1, 32):
            cur = g.offset_qpel(f)
            assert all(abs(c - p) <= 16 for c, p in zip(cur, prev))
            prev = cur
        # frames differ (motion), but are not unrelated noise
        y1, _ = g.frame(1)
        assert (y1 != y0).mean() > 0.5
        assert np.abs(y1.astype(int) - y0.astype(int)).mean() < 40
    finally:
        g.close()
