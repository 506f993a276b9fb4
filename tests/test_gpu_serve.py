"""The serving runtime (include/bc250_serve.h) on the GPU: every session's
frames are bit-identical alone, in a mix of sessions of different sizes
and codecs, and through the single-session pipeline (plain and learned
integer search: the batched launches change nothing but scheduling);
admission follows the device-memory footprint; paced Poisson requests all
complete, in each session's order, after they arrive."""
import numpy as np
import pytest

from conftest import have_gpu, load_pkg

pytestmark = pytest.mark.gpu

MIX = [  # width, height, codec, qp
    (320, 240, 1, 32),
    (640, 480, 0, 30),
    (1280, 720, 1, 27),
    (352, 288, 1, 30),
]
NFRAMES, RING = 5, 3


@pytest.fixture(scope="module")
def pkg():
    if not have_gpu():
        pytest.skip("no GPU")
    return load_pkg()


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def sources(pkg):
    out = []
    for k, (w, h, _, _) in enumerate(MIX):
        g = pkg.FrameGen(500 + k, w, h, 3)
        out.append([g.frame(i) for i in range(RING)])
        g.close()
    return out


def _requests(sids, nframes, t=0.0):
    """Every session's frames, interleaved, all arrived at t."""
    return [(sid, f, f % RING, t) for f in range(nframes) for sid in sids]


def _serve(pkg, sources, which, learned):
    srv = pkg.Server(learned=learned, max_batch=32)
    try:
        sids = []
        for k in which:
            w, h, codec, qp = MIX[k]
            sid = srv.open(w, h, codec=codec, qp=qp, ring=RING)
            assert sid >= 0
            for i, (y, uv) in enumerate(sources[k]):
                srv.upload(sid, i, y, uv)
            sids.append(sid)
        res, _ = srv.run(_requests(sids, NFRAMES), checksums=True)
        st = srv.stats()
    finally:
        srv.close()
    by = {}
    for r in res:
        by[(which[sids.index(int(r["session"]))], int(r["frame"]))] = int(r["checksum"])
    return by, st


def _pipeline(pkg, sources, k, learned):
    """The same session through one gpu_compute context, frame by frame."""
    w, h, codec, qp = MIX[k]
    enc = pkg.Encoder(w, h, search_range=64, lambda8=1, deblock_idc=0, readback=1, codec=codec)
    try:
        if learned:
            enc.set_proposer(pkg.load_proposer_weights(), **pkg.PROPOSER_DEFAULT)
        ims = []
        for y, uv in sources[k]:
            im = enc.new_image()
            enc.upload(im, y, uv)
            ims.append(im)
        out = {}
        for f in range(NFRAMES):
            slot = enc.encode(ims[f % RING], qp, f == 0, 1)
            out[(k, f)] = pkg.staging_checksum(enc.results(slot), codec)
    finally:
        enc.close()
    return out


@pytest.mark.parametrize("learned", [False, True])
def test_each_session_alone_equals_in_the_mix(pkg, sources, learned):
    say(f"serve: mix of {len(MIX)} sessions, learned={learned}")
    mix, st = _serve(pkg, sources, list(range(len(MIX))), learned)
    assert st["max_batch"] >= 2  # the sessions really shared launches
    for k in range(len(MIX)):
        alone, _ = _serve(pkg, sources, [k], learned)
        want = {key: v for key, v in mix.items() if key[0] == k}
        assert alone == want, f"session {k} differs alone and in the mix"
        assert _pipeline(pkg, sources, k, learned) == want, f"session {k} differs from the single-context pipeline"


def test_admission_by_footprint(pkg):
    probe = pkg.Server()
    one = probe.session_bytes(width=640, height=480, codec=1, qp=30, num_slices=1, gop=0, ring=2)
    probe.close()
    srv = pkg.Server(hbm_budget=2 * one + one // 2)
    try:
        a = srv.open(640, 480, qp=30, ring=2)
        b = srv.open(640, 480, qp=30, ring=2)
        assert a >= 0 and b >= 0
        assert srv.open(640, 480, qp=30, ring=2) == -1   # a third does not fit
        assert srv.stats()["hbm_used"] == 2 * one
        srv.close_session(a)
        assert srv.open(640, 480, qp=30, ring=2) >= 0    # and fits once one closed
    finally:
        srv.close()


def test_paced_poisson_requests(pkg, sources):
    rs = np.random.default_rng(5)
    srv = pkg.Server(learned=True)
    try:
        sids = []
        for k in (0, 1, 3):
            w, h, codec, qp = MIX[k]
            sid = srv.open(w, h, codec=codec, qp=qp, ring=RING)
            for i, (y, uv) in enumerate(sources[k]):
                srv.upload(sid, i, y, uv)
            sids.append(sid)
        reqs = []
        for sid in sids:
            t = 0.0
            for f in range(20):
                t += rs.exponential(1 / 60)
                reqs.append((sid, f, f % RING, t))
        reqs.sort(key=lambda r: r[3])
        res, wall = srv.run(reqs)
    finally:
        srv.close()
    assert (res["t_start"] >= res["t_arrival"]).all()
    assert (res["t_done"] > res["t_start"]).all()
    for sid in sids:
        r = res[res["session"] == sid]
        order = np.argsort(r["frame"])
        assert (np.diff(r["t_done"][order]) > 0).all()
        assert (r["t_start"][order][1:] >= r["t_done"][order][:-1]).all()  # low delay: one frame at a time
    say(f"serve: {len(res)} frames in {wall:.3f} s, p99 latency "
        f"{1e3 * np.percentile(res['t_done'] - res['t_arrival'], 99):.2f} ms")
