"""The learned proposer's gfx950 kernels against the CPU oracle
(oracle/orc_cnn.py): each bf16 MFMA convolution on the GPU's own input
(equal to the oracle's bf16 rounding of a float64 sum, up to one bf16 step
where the fp32 sum lands near a rounding boundary), the block sums and head
in fp32, the backward pass (input gradients with flipped weights, weight
gradients, head and loss) against the oracle's adjoint formulation, and the
candidate search against a brute force of the full search's costs."""
import ctypes as C

import numpy as np
import pytest

import orc
import orc_cnn as oc
from conftest import have_gpu, load_pkg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    if not have_gpu():
        pytest.skip("no GPU")
    return load_pkg()


@pytest.fixture(scope="module")
def learned(pkg):
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location("bc250_learned_py", os.path.join(os.path.dirname(pkg.__file__),
                                                                                    "learned.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def say(msg):
    print(msg, flush=True)


def bf(t):
    """int16 device tensor of bf16 bits -> float32 numpy."""
    a = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return a.view(np.float32)


def _setup(pkg, learned, n, seed, train=False):
    import torch
    rs = np.random.default_rng(seed)
    p = learned.init_params(seed)
    p += rs.normal(0, 0.02, p.shape).astype(np.float32)
    for l, (cin, cout) in enumerate(learned.LAYERS):   # padding columns stay zero
        kp = learned.kpad(cin)
        m = p[learned.OFFSETS[l]:learned.OFFSETS[l] + cout * (kp + 1)].reshape(cout, kp + 1)
        m[:, 9 * cin:kp] = 0
    net = learned.Net(pkg, n, p, train=train)
    h, w = 128, 192
    yy, xx = np.mgrid[0:h, 0:w]
    cur = np.clip(128 + 60 * np.sin(xx / 5.0) * np.cos(yy / 7.0) + rs.normal(0, 8, (h, w)), 0, 255).astype(np.uint8)
    ref = np.roll(cur, (2, -3), axis=(0, 1))
    dc, dr = torch.from_numpy(cur).cuda(), torch.from_numpy(ref).cuda()
    tiles = np.array([(0, i % 3, i // 3 % 2) for i in range(n)], np.int32)
    net.set_input(dc, dr, w, w, h, tiles)
    mvin = rs.normal(0, 0.5, (n, 16, 32)).astype(np.float32)
    dmv = torch.from_numpy(mvin).cuda()
    x = oc.input_tiles(cur, ref, [(int(t[1]), int(t[2])) for t in tiles])
    return net, p, x, mvin, dmv


def _close_bf16(got, want, what, frac=0.01):
    """Equal, except a small fraction of entries one bf16 step apart (plus
    the fp32 summation error, 1e-5 of the layer's largest value, for values
    that cancel to near zero)."""
    got, want = np.asarray(got, np.float32), np.asarray(want, np.float32)
    diff = got != want
    step = np.abs(want) * 2 ** -7 + 1e-5 * np.abs(want).max()
    assert (np.abs(got - want)[diff] <= 1.01 * step[diff]).all(), f"{what}: a difference above one bf16 step"
    assert diff.mean() <= frac, f"{what}: {diff.mean():.4f} of entries differ"


def test_input_tiles_match_oracle(pkg, learned):
    net, p, x, _, _ = _setup(pkg, learned, 6, 1)
    got = bf(net.act[0])[:, 1:65, 1:65, :]
    np.testing.assert_array_equal(got, x)
    assert not bf(net.act[0])[:, 0, :, :].any()  # the border stays zero


def test_forward_layers_match_oracle(pkg, learned):
    say("learned: forward, 6 tiles")
    net, p, x, mvin, dmv = _setup(pkg, learned, 6, 2)
    net.forward(6, dmv)
    import torch
    torch.cuda.synchronize()
    prev = x
    for l in range(4):
        got = bf(net.act[l + 1])[:, 1:-1, 1:-1, :]
        w, b = oc.layer_params(p, l)
        w = oc.bf16(w.astype(np.float32)).astype(np.float64)
        want = oc.bf16(np.maximum(oc.conv(prev, w, b, oc.STRIDES[l]), 0).astype(np.float32))
        assert not bf(net.act[l + 1])[:, [0, -1]].any() and not bf(net.act[l + 1])[:, :, [0, -1]].any()
        _close_bf16(got, want, f"conv{l + 1}")
        assert (got > 0).mean() > 0.2, f"conv{l + 1} mostly dead"
        prev = got
    pool = net.pool.cpu().numpy()
    a4 = prev.astype(np.float64)
    want_pool = a4.reshape(6, 4, 4, 4, 4, 32).sum(axis=(2, 4)).reshape(6, 16, 32)
    np.testing.assert_allclose(pool, want_pool, rtol=1e-5, atol=1e-4)
    wh, bh = oc.head_params(p)
    want_z = np.concatenate([want_pool / 16, mvin], axis=2) @ wh.T + bh
    np.testing.assert_allclose(net.z.cpu().numpy(), want_z, rtol=1e-4, atol=1e-4)


def test_fused_forward_equals_the_layers(pkg, learned):
    """The one-launch inference (a tile per workgroup, activations in LDS,
    input rows reused across output rows) gives the cell sums and head outputs of
    the four separate layer launches bit for bit: random weights on 7 tiles
    at the picture's edges, then the trained weights."""
    import torch
    say("learned: fused forward against the layer launches, 7 tiles")
    for trained in (False, True):
        net, p, x, mvin, dmv = _setup(pkg, learned, 7, 11)
        if trained:
            net.params.copy_(torch.from_numpy(pkg.load_proposer_weights()))
            net.pack()
        want_z = net.forward(7, dmv).clone()
        want_pool = net.pool[:7].clone()
        for variant in (0, 1, 2):   # every LDS layout
            assert pkg.hip.bc250_cnn_set_forward_variant(variant) == 0
            net.pool.zero_()
            net.z.zero_()
            got_z = net.forward_fused(7, dmv)
            torch.cuda.synchronize()
            assert torch.equal(net.pool[:7], want_pool), f"block sums differ (trained={trained}, variant {variant})"
            assert torch.equal(got_z, want_z), f"head outputs differ (trained={trained}, variant {variant})"
        assert pkg.hip.bc250_cnn_set_forward_variant(0) == 0
        assert pkg.hip.bc250_cnn_set_forward_variant(3) == -1
        assert (want_pool > 0).float().mean() > 0.2


def test_backward_matches_oracle(pkg, learned):
    import torch
    say("learned: backward, 4 tiles")
    n = 4
    net, p, x, mvin, dmv = _setup(pkg, learned, n, 3, train=True)
    net.forward(n, dmv)
    rs = np.random.default_rng(9)
    target = rs.integers(-24, 25, (n, 16, 2)).astype(np.float32)
    scale = 1.0 / (n * 16)
    loss = net.backward(n, dmv, torch.from_numpy(target).cuda(), scale).cpu().numpy()
    torch.cuda.synchronize()
    acts = [x] + [bf(net.act[l])[:, 1:-1, 1:-1, :] for l in range(1, 5)]
    pool = net.pool.cpu().numpy().astype(np.float64)
    wh, bh = oc.head_params(p)
    z = np.concatenate([pool / 16, mvin], axis=2) @ wh.T + bh
    lb, dzh = oc.loss_and_dz(z, target)
    np.testing.assert_allclose(loss, lb.sum(axis=1), rtol=1e-4, atol=1e-4)
    g, dzs = oc.backward(p, acts, pool, mvin, dzh * scale, rounding="bf16")
    gg = net.grad.cpu().numpy().astype(np.float64)
    # the pre-activation gradients the kernels stored
    for l in (4, 3, 2, 1):
        got = bf(net.dz[l])[:, 1:-1, 1:-1, :]
        want = dzs[l]
        cos = (got * want).sum() / (np.linalg.norm(got) * np.linalg.norm(want) + 1e-30)
        assert cos > 0.999, f"dz{l}: cosine {cos}"
    for l in range(5):
        lo, hi = oc.OFFSETS[l], (oc.OFFSETS[l + 1] if l < 4 else oc.NPARAM)
        a, b = gg[lo:hi], g[lo:hi]
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        assert rel < 2e-2, f"layer {l + 1} gradient: relative error {rel}"


def test_training_steps_reduce_the_loss(pkg, learned):
    import torch
    say("learned: 60 Adam steps on a fixed batch")
    n = 6
    net, p, x, mvin, dmv = _setup(pkg, learned, n, 4, train=True)
    tv = np.array([-3.0, 2.0], np.float32)
    target = torch.from_numpy(np.tile(tv[None, :], (n, 16, 1))).cuda()
    losses = []
    for _ in range(60):
        net.forward(n, dmv)
        losses.append(float(net.backward(n, dmv, target, 1.0 / (n * 16)).sum()))
        net.adam(2e-2)
    z = net.forward(n, dmv).cpu().numpy()
    err = np.abs(z[..., 0:8].reshape(n, 16, 4, 2) * 16 - tv).max(axis=-1).min(axis=-1)
    assert losses[-1] < 0.6 * losses[0], losses   # the floor at the start is mostly log(4) per block
    assert (err < 0.5).mean() > 0.9, err          # the nearest hypothesis within half a pel


def _keys(cur, ref, mbx, mby, dx, dy, R, lam):
    """The full search's 16x16 key and four 8x8 keys at one vector."""
    h, w = ref.shape
    x0, y0 = mbx * 16, mby * 16
    ys = np.clip(np.arange(y0 + dy, y0 + dy + 16), 0, h - 1)
    xs = np.clip(np.arange(x0 + dx, x0 + dx + 16), 0, w - 1)
    d = np.abs(ref[np.ix_(ys, xs)].astype(int) - cur[y0:y0 + 16, x0:x0 + 16])
    q = [d[:8, :8].sum(), d[:8, 8:].sum(), d[8:, :8].sum(), d[8:, 8:].sum()]
    l1 = abs(dx) + abs(dy)
    idx = (dy + R) * (2 * R + 1) + (dx + R)
    return ((sum(q) + 4 * lam * l1) << 15) | idx, [((v + lam * l1) << 15) | idx for v in q]


def test_candidate_search_brute_force_and_full_search(pkg):
    """Every MB's result is the minimum key over the +-4 windows of its
    first ncand candidates (positions outside +-R skipped), and when a
    candidate is the full search's own winner, the result is the full
    search's record, bit for bit."""
    import torch
    rs = np.random.default_rng(12)
    w, h, R, lam = 96, 64, 16, 1
    ref = rs.integers(0, 256, (h, w)).astype(np.uint8)
    cur = np.clip(np.roll(ref, (3, -5), axis=(0, 1)).astype(int) + rs.integers(-4, 5, (h, w)), 0, 255).astype(np.uint8)
    nmb = (w // 16) * (h // 16)
    i16, i8 = orc.me_fullsearch(cur, ref, R, lam)
    cand = rs.integers(-R - 2, R + 3, (nmb, 4, 2)).astype(np.int32)
    cand = np.clip(cand, -R, R)
    cand[::2, 1] = np.stack([i16["dx"], i16["dy"]], axis=1)[::2]     # the true winner on every other MB
    dcur, dref = torch.from_numpy(cur).cuda(), torch.from_numpy(ref).cuda()
    dc = torch.from_numpy(cand.reshape(-1)).cuda()
    for ncand in (2, 4):
        o16 = torch.zeros(nmb * 4, dtype=torch.int32, device="cuda")
        o8 = torch.zeros(nmb * 16, dtype=torch.int32, device="cuda")
        assert pkg.hip.bc250_launch_cand_search(C.c_void_p(dcur.data_ptr()), w, C.c_void_p(dref.data_ptr()), w, w, h,
                                                R, lam, C.c_void_p(dc.data_ptr()), ncand, C.c_void_p(o16.data_ptr()),
                                                C.c_void_p(o8.data_ptr()), None) == 0
        g16 = o16.cpu().numpy().view(orc.IMV)
        g8 = o8.cpu().numpy().view(orc.IMV)
        for m in range(nmb):
            mbx, mby = m % (w // 16), m // (w // 16)
            best16, bq = None, [None] * 4
            for k in range(ncand):
                cx, cy = cand[m, k]
                for dy in range(cy - 4, cy + 5):
                    for dx in range(cx - 4, cx + 5):
                        if abs(dx) > R or abs(dy) > R:
                            continue
                        k16, k8 = _keys(cur, ref, mbx, mby, dx, dy, R, lam)
                        best16 = k16 if best16 is None else min(best16, k16)
                        bq = [v if b is None else min(b, v) for b, v in zip(bq, k8)]
            for key, rec in [(best16, g16[m])] + list(zip(bq, g8[4 * m:4 * m + 4])):
                idx = key & 0x7FFF
                assert (int(rec["dx"]), int(rec["dy"]), int(rec["cost"])) == (idx % (2 * R + 1) - R,
                                                                               idx // (2 * R + 1) - R, key >> 15), m
            if m % 2 == 0 and ncand >= 2:
                assert g16[m] == i16[m]


def test_fallback_select_and_list_search(pkg):
    import torch
    rs = np.random.default_rng(13)
    mbw, mbh = 9, 5
    nmb = mbw * mbh
    r16 = np.zeros(nmb, orc.IMV)
    r16["sad"] = rs.integers(200, 3000, nmb)
    r16["sad"][[7, 20, 33]] = [20000, 9000, 60000]
    d = torch.from_numpy(r16.view(np.int32).reshape(-1).copy()).cuda()
    lst = torch.full((nmb,), -1, dtype=torch.int32, device="cuda")
    cnt = torch.full((1,), 77, dtype=torch.int32, device="cuda")
    alpha, beta, high = 24, 300, 8000
    assert pkg.hip.bc250_fallback_select(C.c_void_p(d.data_ptr()), mbw, mbh, alpha, beta, high,
                                         C.c_void_p(lst.data_ptr()), C.c_void_p(cnt.data_ptr()), None) == 0
    n = int(cnt.item())
    got = sorted(lst.cpu().numpy()[:n].tolist())
    want = []
    for m in range(nmb):
        mx, my = m % mbw, m // mbw
        nb = sorted(int(r16["sad"][k]) for k, ok in ((m - 1, mx > 0), (m + 1, mx + 1 < mbw), (m - mbw, my > 0),
                                                     (m + mbw, my + 1 < mbh)) if ok)
        med = nb[(len(nb) - 1) // 2]
        s = int(r16["sad"][m])
        if s > (alpha * med >> 4) + beta or s > high:
            want.append(m)
    assert got == want and {7, 20, 33} <= set(got)
    # the full search on the list rewrites exactly those MBs
    w, h = mbw * 16, mbh * 16
    ref = rs.integers(0, 256, (h, w)).astype(np.uint8)
    cur = np.roll(ref, (1, 2), axis=(0, 1))
    i16, i8 = orc.me_fullsearch(cur, ref, 16, 1)
    o16 = torch.full((nmb * 4,), -7, dtype=torch.int32, device="cuda")
    o8 = torch.full((nmb * 16,), -7, dtype=torch.int32, device="cuda")
    dcur, dref = torch.from_numpy(cur).cuda(), torch.from_numpy(ref).cuda()
    assert pkg.hip.bc250_launch_me_fullsearch_list(C.c_void_p(dcur.data_ptr()), w, C.c_void_p(dref.data_ptr()), w, w,
                                                   h, 16, 1, C.c_void_p(o16.data_ptr()), C.c_void_p(o8.data_ptr()),
                                                   C.c_void_p(lst.data_ptr()), C.c_void_p(cnt.data_ptr()), None) == 0
    g16 = o16.cpu().numpy().view(orc.IMV)
    for m in range(nmb):
        if m in got:
            assert g16[m] == i16[m]
        else:
            assert g16[m]["dx"] == -7


def test_split_list_search_equals_full_search(pkg):
    """The fallback's split search (each listed MB's candidate rows over
    several workgroups, keys met with atomicMin) writes the oracle full
    search's 16x16 and 8x8 records for exactly the listed MBs, at every
    range, for lists in any order, an empty list and the whole picture."""
    import torch
    rs = np.random.default_rng(29)
    mbw, mbh = 12, 8
    nmb, w, h = mbw * mbh, mbw * 16, mbh * 16
    ref = rs.integers(0, 256, (h, w)).astype(np.uint8)
    ref[40:90, 50:140] = 128 + (np.arange(50, 140)[None, :] % 9) * 3   # a periodic patch: ties
    cur = np.roll(ref, (3, -5), axis=(0, 1))
    dcur, dref = torch.from_numpy(cur).cuda(), torch.from_numpy(ref).cuda()
    keys = torch.zeros(nmb * 5, dtype=torch.int32, device="cuda")
    P = lambda t: C.c_void_p(t.data_ptr())
    for R in (16, 32, 64):
        say(f"learned: split list search, range {R}")
        i16, i8 = orc.me_fullsearch(cur, ref, R, 1)
        sub = [int(m) for m in rs.permutation(nmb)[:nmb // 3]] + [0, mbw - 1, nmb - mbw, nmb - 1]
        sub = list(dict.fromkeys(sub))
        for lst_py in (sub, [], list(range(nmb))):
            lst = torch.tensor(lst_py + [-1] * (nmb - len(lst_py)), dtype=torch.int32, device="cuda")
            cnt = torch.tensor([len(lst_py)], dtype=torch.int32, device="cuda")
            o16 = torch.full((nmb * 4,), -7, dtype=torch.int32, device="cuda")
            o8 = torch.full((nmb * 16,), -7, dtype=torch.int32, device="cuda")
            assert pkg.hip.bc250_launch_me_fullsearch_list_split(P(dcur), w, P(dref), w, w, h, R, 1, P(o16), P(o8),
                                                                 P(lst), P(cnt), P(keys), 0, None) == 0
            g16 = o16.cpu().numpy().view(orc.IMV)
            g8 = o8.cpu().numpy().view(orc.IMV).reshape(nmb, 4)
            on = set(lst_py)
            for m in range(nmb):
                if m in on:
                    assert g16[m] == i16[m], (R, m)
                    assert (g8[m] == i8.reshape(nmb, 4)[m]).all(), (R, m)
                else:
                    assert g16[m]["dx"] == -7 and (g8[m]["dx"] == -7).all()


def test_picture_chain_equals_the_stages(pkg, learned):
    """The learned search of one picture gives, bit for bit, what the
    separate stages give: the network read straight from the pictures
    (pitches above the width, partial 64x64 tiles at the right and bottom)
    equals input tiles + the fused forward in every launch shape; the
    one-launch proposer equals neighbour inputs -> head -> candidates with
    and without a previous picture's vectors, and so does the proposer run
    inside the network's launch; the candidate search, fallback list and
    split search equal the separate launches (records, count and list), on
    two calls in a row with different inputs."""
    import torch
    say("learned: picture chain against the stages, 208x144")
    rs = np.random.default_rng(41)
    w, h, R, lam = 208, 144, 32, 1
    mbw, mbh = w // 16, h // 16
    nmb, tw = mbw * mbh, (mbw + 3) // 4
    nt = tw * ((mbh + 3) // 4)
    P = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    net = learned.Net(pkg, nt, pkg.load_proposer_weights())
    tiles = np.array([(0, i % tw, i // tw) for i in range(nt)], np.int32)
    cfg = (4, 18, 128, 3000)
    cand_a = torch.zeros(nmb * 8, dtype=torch.int32, device="cuda")
    cand_b = torch.zeros_like(cand_a)
    mvin = torch.zeros((nt, 16, 32), dtype=torch.float32, device="cuda")
    z = torch.zeros((nt, 16, learned.NOUT), dtype=torch.float32, device="cuda")
    pool_b = torch.zeros((nt, 16, 32), dtype=torch.float32, device="cuda")
    lists = [torch.full((nmb,), -1, dtype=torch.int32, device="cuda") for _ in range(2)]
    counts = [torch.full((1,), 77, dtype=torch.int32, device="cuda") for _ in range(2)]
    keys = [torch.zeros(nmb * 5, dtype=torch.int32, device="cuda") for _ in range(2)]
    outs = [(torch.zeros(nmb * 4, dtype=torch.int32, device="cuda"), torch.zeros(nmb * 16, dtype=torch.int32,
                                                                                   device="cuda")) for _ in range(2)]
    wh = C.c_void_p(net.params.data_ptr() + 4 * learned.OFFSETS[4])
    for call in range(2):
        yy, xx = np.mgrid[0:h, 0:w]
        cur = np.clip(128 + 50 * np.sin(xx / (4.0 + call)) * np.cos(yy / 6.0) + rs.normal(0, 6, (h, w)), 0,
                      255).astype(np.uint8)
        cur[40:80, 96:160] = rs.integers(0, 256, (40, 64))          # noise: MBs for the fallback
        ref = np.roll(cur, (3 + call, -6), axis=(0, 1))
        cpad = np.zeros((h, 224), np.uint8)
        rpad = np.zeros((h, 256), np.uint8)
        cpad[:, :w], rpad[:, :w] = cur, ref
        dc, dr = torch.from_numpy(cur).cuda(), torch.from_numpy(ref).cuda()
        dcp, drp = torch.from_numpy(cpad).cuda(), torch.from_numpy(rpad).cuda()
        prev = None
        if call == 1:
            v = np.zeros((nmb, 4), np.int32)
            v[:, 0], v[:, 1] = rs.integers(-R, R + 1, nmb), rs.integers(-R, R + 1, nmb)
            prev = torch.from_numpy(v.reshape(-1)).cuda()
        # the network: input tiles + fused forward against the picture form
        net.set_input(dc, dr, w, w, h, tiles)
        assert net.hip.bc250_cnn_forward(P(net.act[0]), P(net.packed), P(net.params), P(net.pool), nt, None) == 0
        for variant in (0, 1, 2):
            assert pkg.hip.bc250_cnn_set_forward_variant(variant) == 0
            pool_b.fill_(-1.0)
            assert pkg.hip.bc250_cnn_forward_picture(P(dcp), 224, P(drp), 256, w, h, P(net.packed), P(net.params),
                                                     P(pool_b), None) == 0
            torch.cuda.synchronize()
            assert torch.equal(pool_b, net.pool), (call, variant)
        assert pkg.hip.bc250_cnn_set_forward_variant(0) == 0
        # the proposer: mvin -> head -> candidates against one launch
        assert pkg.hip.bc250_cnn_mvin(P(prev), mbw, mbh, P(mvin), None) == 0
        assert pkg.hip.bc250_cnn_head(wh, P(net.pool), P(mvin), P(z), nt, None) == 0
        assert pkg.hip.bc250_cnn_candidates(P(z), mbw, mbh, R, P(cand_a), None) == 0
        assert pkg.hip.bc250_cnn_propose(P(net.pool), P(net.params), P(prev), mbw, mbh, R, P(cand_b), None) == 0
        torch.cuda.synchronize()
        assert torch.equal(cand_a, cand_b), call
        # the search: separate stages (0) against the chain (1)
        (a16, a8), (b16, b8) = outs
        assert pkg.hip.bc250_launch_cand_search(P(dc), w, P(dr), w, w, h, R, lam, P(cand_a), cfg[0], P(a16), P(a8),
                                                None) == 0
        assert pkg.hip.bc250_fallback_select(P(a16), mbw, mbh, cfg[1], cfg[2], cfg[3], P(lists[0]), P(counts[0]),
                                             None) == 0
        assert pkg.hip.bc250_launch_me_fullsearch_list_split(P(dc), w, P(dr), w, w, h, R, lam, P(a16), P(a8),
                                                             P(lists[0]), P(counts[0]), P(keys[0]), 0, None) == 0
        assert pkg.hip.bc250_learned_search_picture(P(dcp), 224, P(drp), 256, w, h, R, lam, P(net.params),
                                                    P(net.packed), P(prev), *cfg, P(pool_b), P(cand_b),
                                                    P(lists[1]), P(counts[1]), P(keys[1]), P(b16), P(b8), None) == 0
        torch.cuda.synchronize()
        # the chain's proposer runs inside the network's launch: the same candidates
        assert torch.equal(cand_a, cand_b), call
        n = int(counts[0].item())
        assert int(counts[1].item()) == n and 0 < n < nmb, (call, n)
        assert sorted(lists[0][:n].tolist()) == sorted(lists[1][:n].tolist())
        assert torch.equal(a16, b16) and torch.equal(a8, b8), call
        # and the records of the listed MBs are the oracle's full search
        i16, _ = orc.me_fullsearch(cur, ref, R, lam)
        g16 = b16.cpu().numpy().view(orc.IMV)
        for m in lists[1][:n].tolist():
            assert g16[m] == i16[m], (call, m)


def _fixed_candidate_params(learned, cands):
    """Parameters whose head ignores its input: every block's hypotheses are
    `cands` (pel) and the confidences fall in their listed order."""
    p = np.zeros(learned.NPARAM, np.float32)
    h = p[learned.OFFSETS[4]:].reshape(learned.NOUT, learned.NFEAT + 1)
    for k, (dx, dy) in enumerate(cands):
        h[2 * k, -1], h[2 * k + 1, -1] = dx / 16.0, dy / 16.0
        h[8 + k, -1] = 4.0 - k
    return p


def _learned_int_search(cur, ref, R, lam, cands, ncand, alpha, beta, high):
    """CPU restatement of the learned integer search with fixed candidates:
    per MB the best keys over the +-4 windows, then the fallback rule on the
    16x16 SADs, the oracle's full search for the listed MBs."""
    h, w = cur.shape
    mbw, mbh = w // 16, h // 16
    nmb = mbw * mbh
    r16 = np.zeros(nmb, orc.IMV)
    r8 = np.zeros(nmb * 4, orc.IMV)
    for m in range(nmb):
        mbx, mby = m % mbw, m // mbw
        b16, bq = None, [None] * 4
        for cx, cy in cands[:ncand]:
            for dy in range(cy - 4, cy + 5):
                for dx in range(cx - 4, cx + 5):
                    if abs(dx) > R or abs(dy) > R:
                        continue
                    k16, k8 = _keys(cur, ref, mbx, mby, dx, dy, R, lam)
                    b16 = k16 if b16 is None else min(b16, k16)
                    bq = [v if b is None else min(b, v) for b, v in zip(bq, k8)]
        for rec, key, mul in [(r16[m:m + 1], b16, 4)] + [(r8[4 * m + q:4 * m + q + 1], bq[q], 1) for q in range(4)]:
            idx, cost = key & 0x7FFF, key >> 15
            dx, dy = idx % (2 * R + 1) - R, idx // (2 * R + 1) - R
            rec[0] = (dx, dy, cost, cost - mul * lam * (abs(dx) + abs(dy)))
    fall = []
    for m in range(nmb):
        mx, my = m % mbw, m // mbw
        nb = sorted(int(r16["sad"][k]) for k, ok in ((m - 1, mx > 0), (m + 1, mx + 1 < mbw), (m - mbw, my > 0),
                                                     (m + mbw, my + 1 < mbh)) if ok)
        med = nb[(len(nb) - 1) // 2] if nb else 0
        s = int(r16["sad"][m])
        if s > (alpha * med >> 4) + beta or s > high:
            fall.append(m)
    for m in fall:
        a, b = orc.me_fullsearch_mb(cur, ref, R, lam, m % mbw, m // mbw)
        r16[m], r8[4 * m:4 * m + 4] = a, b
    return r16, r8, fall


@pytest.mark.parametrize("codec", ["h264", "hevc"])
def test_pipeline_learned_search_with_fixed_candidates(pkg, learned, codec):
    """The learned integer search inside the pipeline (set_proposer): with
    the network's output pinned to four known candidates, every P frame's
    vectors are the CPU restatement's (windows, fallback rule, full search
    of the fallback MBs, then the codec's own sub-pel refinement), the
    fallback count reported per frame is the restatement's, and the coding
    after it is the oracle's (tests/seqcheck.py with the GPU's vectors)."""
    import seqcheck
    w, h, R, lam, qp = 128, 96, 16, 1, 30
    cands = [(0, 0), (5, -3), (-9, 6), (12, 12)]
    cfg = dict(ncand=3, alpha_q4=20, beta=200, t_high=3000)
    g = pkg.FrameGen(33, w, h, 3)
    frames = [g.frame(i) for i in range(4)]
    g.close()
    c = pkg.GPU_CODEC_HEVC if codec == "hevc" else pkg.GPU_CODEC_H264
    enc = pkg.Encoder(w, h, search_range=R, lambda8=lam, deblock_idc=0, readback=1, codec=c)
    fb_counts = []
    try:
        enc.set_proposer(_fixed_candidate_params(learned, cands), **cfg)
        images = []
        for y, uv in frames:
            im = enc.new_image()
            enc.upload(im, y, uv)
            images.append(im)
        got = []
        for i, im in enumerate(images):
            slot = enc.encode(im, qp, i == 0, 1)
            r = enc.results(slot)
            r["is_intra"] = i == 0
            got.append(r)
            fb_counts.append(enc.fallbacks(slot))
    finally:
        enc.close()
    coded = [pkg.pad_coded(*f) for f in frames]
    assert fb_counts[0] == -1  # the I frame has no search
    ref_y, ref_uv = coded[0]
    # the oracle chain: the reconstruction each P frame searched is the
    # previous frame's as the coding check rebuilds it
    recon = []
    if codec == "hevc":
        seqcheck.check_hevc(coded[:1], got[:1], qp, 1, R, lam)
    for i in range(1, len(frames)):
        cy, cuv = coded[i]
        if codec == "hevc":
            ry, ruv = seqcheck.check_hevc(coded[:i], got[:i], qp, 1, R, lam, sample=0)
        else:
            ry, ruv = seqcheck.check_h264(coded[:i], got[:i], qp, 1, R, lam, sample=0)
        r16, r8, fall = _learned_int_search(cy, ry, R, lam, cands, cfg["ncand"], cfg["alpha_q4"], cfg["beta"],
                                            cfg["t_high"])
        assert fb_counts[i] == len(fall), (i, fb_counts[i], len(fall))
        mbw = w // 16
        if codec == "hevc":
            for t in range((w // 16) * (h // 16)):
                for z in range(4):
                    x0, y0 = (t % mbw) * 16 + (z & 1) * 8, (t // mbw) * 16 + (z >> 1) * 8
                    mx, my, sad = orc.hevc_refine8(cy, ry, x0, y0, int(r8[4 * t + z]["dx"]) * 4,
                                                   int(r8[4 * t + z]["dy"]) * 4)
                    gv = got[i]["mv8"][4 * t + z]
                    assert (int(gv["mvx"]), int(gv["mvy"]), int(gv["sad"])) == (mx, my, sad), (i, t, z)
        else:
            for m in range((w // 16) * (h // 16)):
                m16, m8 = orc.me_refine_mb(cy, ry, m % mbw, m // mbw, r16[m], r8[4 * m:4 * m + 4])
                assert got[i]["mv"][m] == m16, (i, m)
                assert (got[i]["mv8"][4 * m:4 * m + 4] == m8).all(), (i, m)
    # and the coding of every frame with those vectors
    if codec == "hevc":
        seqcheck.check_hevc(coded, got, qp, 1, R, lam, sample=0)
    else:
        seqcheck.check_h264(coded, got, qp, 1, R, lam, sample=0)


def test_trained_network_gains_over_no_network(pkg, learned):
    """The committed weights on held-out sequences (seeds 9100.., not seen
    in training): the share of MBs whose candidate search alone (before any
    fallback) finds exactly the +-64 full search's 16x16 result is higher
    with the network's candidates than with the same search's no-network
    candidates (zero, the previous picture's co-located, left and top
    vectors), and misses fall by at least a third."""
    import importlib.util
    import os
    import torch
    here = os.path.dirname(pkg.__file__)
    spec = importlib.util.spec_from_file_location("bc250_train_py", os.path.join(here, "train_proposer.py"))
    tp = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(tp)
    say("learned: held-out hit rates, 2 sequences x 7 frames at 1080p")
    ds = tp.Dataset(pkg, torch, [9100, 9101], 7, seed_offset=1)
    hits = tp.hit_rates(ds, pkg, learned, torch, pkg.load_proposer_weights())
    say(f"learned: {hits}")
    assert hits["network"] > hits["no_network"]
    assert (1 - hits["network"]) < (2 / 3) * (1 - hits["no_network"]), hits
