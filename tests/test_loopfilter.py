"""The oracle's H.264 picture deblocking (oracle/orc_h264.c:
orc_h264_deblock -- bS derivation, macroblock and edge order, chroma edges,
slice edges with disable_deblocking_filter_idc 0 and 2) against a literal
reading of clause 8.7 (tests/h264_loopfilter.py) on small pictures with
random content, coded blocks, vectors, slices and QPs; and that reading's
tables (8-15, 8-16, 8-17) against the reference decoder's own
(h264_dec_tables.c, built into oracle/_ref)."""
import ctypes as C
import os

import numpy as np
import pytest

import h264_loopfilter as lf
import orc


def _slice_of(total, ns):
    return [max(s for s in range(ns) if s * total // ns <= m) for m in range(total)]


def _picture(rs, mbw, mbh):
    h, w = mbh * 16, mbw * 16
    yy, xx = np.mgrid[0:h, 0:w]
    y = 128 + 20 * np.sin(xx / 7.0) * np.cos(yy / 9.0) + rs.normal(0, 2, (h, w))
    y += np.kron(rs.integers(-8, 9, (h // 4, w // 4)), np.ones((4, 4)))   # steps at block edges
    steps = np.kron(rs.integers(-6, 7, (h // 8, w // 8)), np.ones((4, 4)))              # per 4x4 chroma block
    uv = 128 + rs.normal(0, 2, (h // 2, w)) + np.repeat(steps, 2, axis=1)
    return np.clip(y, 0, 255).astype(np.uint8), np.clip(uv, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("seed", range(12))
def test_oracle_picture_deblock_equals_clause_8_7(seed):
    rs = np.random.default_rng(seed)
    mbw, mbh = 4, 3
    total = mbw * mbh
    is_intra = seed % 4 == 0
    qp = int(rs.integers(20, 52))
    ns = int(rs.integers(1, 4))
    idc = (0, 2, 0, 2, 1)[seed % 5]
    y, uv = _picture(rs, mbw, mbh)
    nz = (rs.random(total * 24) < 0.3).astype(np.uint32) * rs.integers(1, 5, total * 24).astype(np.uint32)
    mvs = np.zeros(total, orc.MV)
    mvs["mvx"] = rs.integers(-8, 9, total)
    mvs["mvy"] = rs.integers(-8, 9, total)
    if seed % 3 == 1:
        mvs["mvx"][:] = 5   # equal vectors: bS 0 where no coefficients
        mvs["mvy"][:] = -2
    gy, guv = orc.deblock(y, uv, qp, is_intra, nz, mvs, ns, idc)

    sl = _slice_of(total, ns)
    pic = lf.Picture(mbw, mbh, intra=[is_intra] * total, qp=[qp] * total,
                     nnz=[[int(nz[m * 24 + b]) for b in range(16)] for m in range(total)],
                     mv=[[(int(mvs["mvx"][m]), int(mvs["mvy"][m]))] * 16 for m in range(total)],
                     ref=[0] * total, slice_of=sl, idc=[idc] * ns, off_a=[0] * ns, off_b=[0] * ns)
    wy, cb, cr = y.copy(), np.ascontiguousarray(uv[:, 0::2]), np.ascontiguousarray(uv[:, 1::2])
    lf.deblock(pic, wy, cb, cr)
    wuv = np.empty_like(uv)
    wuv[:, 0::2], wuv[:, 1::2] = cb, cr
    np.testing.assert_array_equal(gy, wy)
    np.testing.assert_array_equal(guv, wuv)
    changed = np.mean(gy != y)
    if idc == 1:
        assert changed == 0
    elif qp >= 30:
        assert changed > 0.02, changed   # the case filters something


def test_filter_offsets_and_per_mb_qp_move_the_thresholds():
    """What the oracle fixes at zero offsets and one QP, the reading takes
    from the slice and the macroblocks: FilterOffsetA raises indexA (more
    filtering), and qPav averages the two macroblocks' QPs."""
    rs = np.random.default_rng(7)
    mbw, mbh = 2, 2
    y, uv = _picture(rs, mbw, mbh)
    base = dict(intra=[False] * 4, nnz=[[1] * 16] * 4, mv=[[(0, 0)] * 16] * 4, ref=[0] * 4, slice_of=[0] * 4,
                idc=[0], off_b=[0])
    counts = []
    for oa in (-12, 0, 12):
        pic = lf.Picture(mbw, mbh, qp=[28] * 4, off_a=[oa], **base)
        wy = y.copy()
        lf.deblock(pic, wy, uv[:, 0::2].copy(), uv[:, 1::2].copy())
        counts.append(int(np.sum(wy != y)))
    assert counts[0] < counts[1] < counts[2], counts
    # qPav of the shared edge: on a picture flat but for a step of 10 at the
    # vertical MB edge x = 16 (every other edge sees equal samples), QPs
    # 27 | 29 filter it as 28 | 28 does, and 21 | 23 as 22 | 22, where
    # alpha (9) is below the step and nothing moves
    flat = np.where(np.arange(32)[None, :] < 16, 100, 110).repeat(32, axis=0).astype(np.uint8)
    c8 = np.full((16, 16), 128, np.uint8)

    def run(qps):
        yy = flat.copy()
        lf.deblock(lf.Picture(mbw, mbh, qp=qps, off_a=[0], **base), yy, c8.copy(), c8.copy())
        return yy
    assert np.array_equal(run([27, 29, 27, 29]), run([28] * 4))
    assert not np.array_equal(run([28] * 4), flat)
    assert np.array_equal(run([21, 23, 21, 23]), flat) and np.array_equal(run([22] * 4), flat)


@pytest.mark.skipif(not os.path.exists(orc.REF_H264), reason="oracle/_ref not built (reference tree absent)")
def test_tables_match_the_reference_decoder():
    lib = C.CDLL(orc.REF_H264)

    def table(name, n):
        return list((C.c_uint8 * n).in_dll(lib, name))
    assert table("h264d_alpha", 52) == lf.ALPHA
    assert table("h264d_beta", 52) == lf.BETA
    t = table("h264d_tc0", 156)
    assert [tuple(t[3 * i:3 * i + 3]) for i in range(52)] == lf.TC0
    assert table("h264d_chroma_qp", 52) == [lf.qpc_of(q) for q in range(52)]
