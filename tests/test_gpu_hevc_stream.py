"""HEVC bitstreams of GPU-coded pictures: the pipeline's own staging (levels,
modes, vectors, as the readback leaves them) written by the host writer
(include/bc250_hevc_stream.h) and decoded by the reference's own H.265
decoding functions (tests/hevc_refdec.py) gives, picture by picture, the
GPU's reconstruction bit for bit - deblocked, at CTU16 and CTU32, 8 and 10
bits, with slices and IDR pictures inside the sequence."""
import numpy as np
import pytest

import hevc_refdec as rd
from conftest import have_gpu, load_pkg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    if not have_gpu():
        pytest.skip("no GPU")
    if not rd.available():
        pytest.skip("reference decoder not built (oracle/_ref)")
    return load_pkg()


def say(msg):
    print(msg, flush=True)


def _run(pkg, w, h, n, qp, ns, rng, bd, ctu, deblock=True, gop=0, seed=7, nrect=3):
    say(f"hevc stream {w}x{h} frames={n} qp={qp} slices={ns} R={rng} bd={bd} ctu={ctu} deblock={deblock} gop={gop}")
    g = pkg.FrameGen(seed, w, h, nrect)
    enc = pkg.Encoder(w, h, search_range=rng, lambda8=1, deblock_idc=0 if deblock else 1, codec=pkg.GPU_CODEC_HEVC,
                      bit_depth=bd, ctu_log2=ctu)
    st = pkg.HevcStream(enc.wc, enc.hc, bd, ctu, ns, deblock, init_qp=26,
                        crop_width=w if w != enc.wc else 0, crop_height=h if h != enc.hc else 0)
    rs = np.random.default_rng(seed)
    s, recs, poc = st.headers(), [], 0
    try:
        for i in range(n):
            vy, vuv = g.frame(i)
            if bd == 10:
                vy = np.asarray(vy, np.uint16) * 4 + rs.integers(0, 4, vy.shape).astype(np.uint16)
                vuv = np.asarray(vuv, np.uint16) * 4 + rs.integers(0, 4, vuv.shape).astype(np.uint16)
            im = enc.new_image()
            enc.upload(im, vy, vuv)
            intra = i == 0 or bool(gop and i % gop == 0)
            poc = 0 if intra else poc + 1
            r = enc.results(enc.encode(im, qp, intra, ns))
            assert enc.error_flags() == 0
            s += st.picture(intra, poc, qp, r["quant"], r["hevc_mode"], r["mv8"])
            recs.append(enc.map_recon())
    finally:
        enc.close()
        g.close()
    pics = rd.decode(s, enc.wc, enc.hc, bd, n)
    assert len(pics) == n
    for i, ((y, cb, cr), (my, muv)) in enumerate(zip(pics, recs)):
        assert np.array_equal(y, my), f"picture {i} luma"
        assert np.array_equal(cb, muv[:, 0::2]) and np.array_equal(cr, muv[:, 1::2]), f"picture {i} chroma"
    say(f"  {n} pictures, {len(s)} bytes, decode == recon")
    return s


def test_stream_640x480_ctu16_slices(pkg):
    _run(pkg, 640, 480, 4, 32, 4, 32, 8, 4)


def test_stream_1080p_ctu16(pkg):
    """The headline's picture size and coding (1080p, +-64, CQP 27)."""
    _run(pkg, 1920, 1080, 3, 27, 1, 64, 8, 4, nrect=12)


def test_stream_ctu32_gop_odd_size(pkg):
    _run(pkg, 200, 120, 5, 32, 3, 16, 8, 5, gop=3)


def test_stream_main10_ctu32(pkg):
    _run(pkg, 240, 144, 4, 27, 2, 16, 10, 5, gop=2)


def test_stream_main10_no_deblock(pkg):
    _run(pkg, 128, 96, 3, 37, 1, 16, 10, 4, deblock=False)
