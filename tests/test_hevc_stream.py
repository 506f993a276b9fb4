"""The HEVC bitstream writer (include/bc250_hevc_stream.h) on the CPU: the
oracle's pictures (tests run the oracle encoder, as the GPU pipeline's
staging is checked against it elsewhere) written as a stream and decoded by
the reference's own H.265 decoding functions (tests/hevc_refdec.py) must give
back the oracle's reconstruction bit for bit, deblocking included; the
parameter sets read back through the reference's parser; planted mistakes in
the staging show in the decode; invalid configurations are refused."""
import numpy as np
import pytest

import orc
import hevc_refdec as rd
from conftest import load_pkg

needs_ref = pytest.mark.skipif(not rd.available(), reason="reference decoder not built (oracle/_ref)")


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


def _frames(pkg, w, h, n, bd, seed):
    g = pkg.FrameGen(seed, w, h, 3)
    fr = [pkg.pad_coded(*g.frame(i)) for i in range(n)]
    g.close()
    if bd == 10:
        fr = [((y.astype(np.uint16) << 2) | (y >> 6), (uv.astype(np.uint16) << 2) | (uv >> 6)) for y, uv in fr]
    return fr


def _sequence(pkg, w, h, n, qp, ns, bd, ctu, deblock, seed=3, rng=16):
    """(stream, oracle pictures) of an IPPP sequence."""
    st = pkg.HevcStream(w, h, bd, ctu, ns, deblock, init_qp=26)
    s, recs, ry, ruv = st.headers(), [], None, None
    for i, (cy, cuv) in enumerate(_frames(pkg, w, h, n, bd, seed)):
        o = orc.hevc_encode_frame(cy, cuv, ry, ruv, qp, i == 0, ns, rng, 1, deblock, bd, ctu)
        s += st.picture(i == 0, i, qp, o["coef"], o["modes"], o["mv8"])
        ry, ruv = o["y"], o["uv"]
        recs.append(o)
    return s, recs


def _assert_decodes_to(s, recs, w, h, bd):
    pics = rd.decode(s, w, h, bd, len(recs))
    assert len(pics) == len(recs)
    for i, ((y, cb, cr), o) in enumerate(zip(pics, recs)):
        np.testing.assert_array_equal(y, o["y"], err_msg=f"picture {i} luma")
        np.testing.assert_array_equal(cb, o["uv"][:, 0::2], err_msg=f"picture {i} Cb")
        np.testing.assert_array_equal(cr, o["uv"][:, 1::2], err_msg=f"picture {i} Cr")


CASES = [  # w, h, frames, qp, slices, bit depth, ctu log2, deblock
    (64, 48, 3, 30, 1, 8, 4, False),
    (64, 48, 3, 30, 1, 8, 4, True),
    (128, 96, 4, 22, 2, 8, 4, True),
    (96, 64, 2, 0, 1, 8, 4, True),      # large levels: long remaining-level codes
    (96, 64, 2, 51, 1, 8, 4, True),
    (176, 144, 4, 27, 3, 8, 5, True),   # CTU32 with partial CTUs, CU8 / CU16 / CU32
    (176, 144, 3, 32, 1, 10, 5, True),
    (160, 96, 3, 37, 2, 10, 4, True),
]


@needs_ref
@pytest.mark.parametrize("w,h,n,qp,ns,bd,ctu,deblock", CASES)
def test_stream_decodes_to_the_reconstruction(pkg, w, h, n, qp, ns, bd, ctu, deblock):
    s, recs = _sequence(pkg, w, h, n, qp, ns, bd, ctu, deblock)
    _assert_decodes_to(s, recs, w, h, bd)
    if ctu == 5:  # the case covers every CU size
        sizes = set(np.concatenate([o["modes"].reshape(-1, 16)[:, :4].ravel() for o in recs[1:]]).tolist())
        assert {3, 4}.issubset(sizes), sizes


@needs_ref
def test_emulation_prevention_and_nal_types(pkg):
    s, recs = _sequence(pkg, 128, 96, 3, 0, 2, 8, 4, True)
    units = rd.nal_units(s)
    assert [t for t, _ in units] == [32, 33, 34] + [19, 19] + [1, 1] * 2
    for _, u in units:
        b = bytes(u)
        for k in range(len(b) - 2):   # no 00 00 00..02; 00 00 03 only as an escape of 00..03
            if b[k] == 0 and b[k + 1] == 0:
                assert b[k + 2] >= 3, k
                assert b[k + 2] != 3 or k + 3 == len(b) or b[k + 3] <= 3, k
    assert b"\x00\x00\x03" in s     # QP 0 levels do produce escapes


@needs_ref
@pytest.mark.parametrize("bd,ctu,deblock,crop", [(8, 4, True, (0, 0)), (10, 5, False, (1920, 1080))])
def test_parameter_sets_through_the_reference_parser(pkg, bd, ctu, deblock, crop):
    st = pkg.HevcStream(1920, 1088, bd, ctu, 4, deblock, init_qp=30, crop_width=crop[0], crop_height=crop[1])
    units = dict(rd.nal_units(st.headers()))
    sps, pps = rd.read_sps(units[33]), rd.read_pps(units[34])
    assert (sps["width"], sps["height"], sps["bit_depth_luma"], sps["bit_depth_chroma"]) == (1920, 1088, bd, bd)
    assert (sps["crop_bottom"], sps["crop_right"]) == ((8, 0) if crop[1] else (0, 0))
    assert (sps["log2_min_cb"], sps["log2_ctb"], sps["log2_min_tb"]) == (3, ctu, 2)
    assert sps["log2_max_tb"] == (3 if ctu == 4 else 5)
    assert (sps["max_th_depth_inter"], sps["max_th_depth_intra"], sps["num_st_rps"]) == (0, 0, 1)
    assert (sps["temporal_mvp"], sps["amp"], sps["sao"], sps["chroma_format_idc"], sps["log2_max_poc_lsb"]) == \
        (0, 0, 0, 1, 8)
    assert pps["init_qp"] == 30 and pps["deblocking_disabled"] == (0 if deblock else 1)
    assert (pps["loop_filter_across_slices"], pps["cu_qp_delta"], pps["sign_data_hiding"], pps["tiles"],
            pps["entropy_coding_sync"], pps["num_ref_idx_l0_default"]) == (0, 0, 0, 0, 0, 1)


@needs_ref
def test_planted_mistakes_show_in_the_decode(pkg):
    """A changed level, vector or intra mode in the staging decodes to a
    different picture: the comparison above would see a wrong stream."""
    w, h = 64, 48
    fr = _frames(pkg, w, h, 2, 8, 5)
    o0 = orc.hevc_encode_frame(fr[0][0], fr[0][1], None, None, 30, True, 1, 16, 1, True, 8, 4)
    o1 = orc.hevc_encode_frame(fr[1][0], fr[1][1], o0["y"], o0["uv"], 30, False, 1, 16, 1, True, 8, 4)
    st = pkg.HevcStream(w, h, 8, 4, 1, True, init_qp=30)
    head = st.headers() + st.picture(True, 0, 30, o0["coef"], o0["modes"])

    def p1(coef=o1["coef"], mv8=o1["mv8"]):
        return rd.decode(head + st.picture(False, 1, 30, coef, o1["modes"], mv8), w, h, 8, 2)[1][0]

    assert np.array_equal(p1(), o1["y"])
    c = o1["coef"].copy()
    c[np.flatnonzero(c)[0]] += 1
    assert not np.array_equal(p1(coef=c), o1["y"])
    m = o1["mv8"].copy()
    m["mvx"][5] += 4
    assert not np.array_equal(p1(mv8=m), o1["y"])
    md = o0["modes"].copy()
    md[3] = 26 if md[3] != 26 else 10
    assert not np.array_equal(rd.decode(st.headers() + st.picture(True, 0, 30, o0["coef"], md), w, h, 8, 1)[0][0],
                              o0["y"])


def test_writer_refuses_invalid_configurations(pkg):
    for kw in (dict(width=40), dict(bit_depth=9), dict(ctu_log2=6), dict(num_slices=4), dict(init_qp=52),
               dict(crop_width=70)):
        a = dict(width=64, height=48, bit_depth=8, ctu_log2=4, num_slices=1, init_qp=26)
        a.update(kw)
        st = pkg.HevcStream(**a)
        with pytest.raises(ValueError):
            st.headers()
    st = pkg.HevcStream(64, 48)
    z = np.zeros(12 * 384, np.int16)
    with pytest.raises(RuntimeError):   # QP outside 0..51
        st.picture(True, 0, 52, z, np.zeros(12 * 16, np.uint8))
    with pytest.raises(ValueError):     # staging of another size
        st.picture(True, 0, 30, z[:100], np.zeros(12 * 16, np.uint8))


@needs_ref
@pytest.mark.parametrize("bd", [8, 10])
def test_still_picture_codes_cu32(pkg, bd):
    """A repeated picture: one vector everywhere, so the CTU32 P pictures
    are CU32s (with CU16s and CU8s at the partial CTUs of the edges), and
    their CUs go as merges and skips: the P pictures cost a few percent of
    the I picture."""
    w, h = 144, 112
    f0 = _frames(pkg, w, h, 1, bd, 9)[0]
    st = pkg.HevcStream(w, h, bd, 5, 2, True, init_qp=26)
    s, recs, ry, ruv, nbytes = st.headers(), [], None, None, []
    for i in range(3):
        o = orc.hevc_encode_frame(f0[0], f0[1], ry, ruv, 24, i == 0, 2, 16, 1, True, bd, 5)
        pic = st.picture(i == 0, i, 24, o["coef"], o["modes"], o["mv8"])
        s += pic
        nbytes.append(len(pic))
        ry, ruv = o["y"], o["uv"]
        recs.append(o)
    sizes = set(recs[1]["modes"].reshape(-1, 16)[:, :4].ravel().tolist())
    assert 5 in sizes and 4 in sizes, sizes
    assert max(nbytes[1:]) < nbytes[0] / 10, nbytes
    _assert_decodes_to(s, recs, w, h, bd)
