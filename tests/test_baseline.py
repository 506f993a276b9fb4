"""The CPU baseline encoder (oracle/orc_baseline.c) is the oracle's
computation: its AVX2 search gives the scalar oracle's records, and its
threaded IPPP sequence ends on the reconstruction the oracle's per-frame
chain (orc.hevc_encode_frame) gives, so the baseline times the same
encoder the GPU runs."""
import numpy as np
import pytest

import orc


@pytest.mark.skipif(not orc.baseline_has_avx2(), reason="no AVX2 on this CPU")
@pytest.mark.parametrize("rng", [16, 32])
def test_avx2_search_equals_oracle(rng):
    rs = np.random.default_rng(rng)
    h, w = 64, 96
    ref = np.clip(128 + rs.normal(0, 30, (h, w)), 0, 255).astype(np.uint8)
    ref[10:40, 20:70] = 77          # a flat patch: ties broken by the raster index
    cur = np.roll(ref, (2, -5), axis=(0, 1))
    a16, a8 = orc.me_fullsearch(cur, ref, rng, 1)
    for nt in (1, 3):
        b16, b8 = orc.baseline_fullsearch(cur, ref, rng, 1, nthreads=nt, simd=1)
        assert (a16 == b16).all() and (a8 == b8).all()
    c16, c8 = orc.baseline_fullsearch(cur, ref, rng, 1, nthreads=2, simd=0)
    assert (a16 == c16).all() and (a8 == c8).all()


def test_sequence_equals_the_oracle_chain():
    g_h, g_w, n, qp, ns = 64, 96, 3, 30, 2
    rs = np.random.default_rng(5)
    yy, xx = np.mgrid[0:g_h, 0:g_w]
    ys, uvs = [], []
    for f in range(n):
        y = 128 + 40 * np.sin((xx + 3 * f) / 9.0) * np.cos((yy - 2 * f) / 7.0) + rs.normal(0, 3, (g_h, g_w))
        ys.append(np.clip(y, 0, 255).astype(np.uint8))
        uvs.append(np.clip(128 + rs.normal(0, 6, (g_h // 2, g_w)), 0, 255).astype(np.uint8))
    ms, gy, guv = orc.baseline_hevc_seq(np.stack(ys), np.stack(uvs), qp, 16, 1, ns, nthreads=2)
    assert (ms > 0).all()
    ry = ruv = None
    for f in range(n):
        o = orc.hevc_encode_frame(ys[f], uvs[f], ry, ruv, qp, f == 0, ns, 16, 1)
        ry, ruv = o["y"], o["uv"]
    np.testing.assert_array_equal(gy, ry)
    np.testing.assert_array_equal(guv, ruv)
