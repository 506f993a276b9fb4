#!/usr/bin/env python3
"""Trains the learned motion proposer on the GPU with the project's own
forward and backward kernels (learned.Net over csrc/me_cnn.hip).

Data: seeded synthetic 1080p sequences (host/framegen.c, 12 rectangles),
coded by the H.264 pipeline at a QP drawn from 22 / 27 / 32 / 37 so that the
references are real reconstructions; for every P frame the targets are the
16x16 integer vectors of the plain +-64 full search on (frame, previous
reconstruction) -- the vectors the plain path itself produces -- and the
neighbour input is the previous P frame's vectors (none after the I frame).

Writes the weights as raw float32 (weights/me_proposer.f32) and a JSON log
(loss curve, held-out hit rates against a no-network baseline).

Usage: python train_proposer.py [--seeds 40] [--frames 51] [--steps 3000]
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
W, H, RANGE, LAMBDA8 = 1920, 1080, 64, 1
QPS = (22, 27, 32, 37)


def load():
    name = "bc250_vaapi_amd"
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"))
        m = importlib.util.module_from_spec(spec)
        sys.modules[name] = m
        spec.loader.exec_module(m)
    spec = importlib.util.spec_from_file_location("bc250_learned_py", os.path.join(HERE, "learned.py"))
    ln = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ln)
    return sys.modules[name], ln


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Dataset:
    """Device tensors of every P frame of the given seeds: cur / ref luma
    [N, hc, wc] uint8, full-search 16x16 vectors [N, nmb, 4] int32 ({dx, dy,
    cost, sad}), the previous P frame's vectors (zero after the I frame)."""

    def __init__(self, pkg, torch, seeds, frames, seed_offset=0):
        self.torch = torch
        wc, hc = (W + 15) // 16 * 16, (H + 15) // 16 * 16
        self.wc, self.hc, self.mbw, self.mbh = wc, hc, wc // 16, hc // 16
        self.nmb = self.mbw * self.mbh
        n = len(seeds) * (frames - 1)
        dev = "cuda"
        self.cur = torch.empty((n, hc, wc), dtype=torch.uint8, device=dev)
        self.ref = torch.empty((n, hc, wc), dtype=torch.uint8, device=dev)
        self.fs = torch.zeros((n, self.nmb, 4), dtype=torch.int32, device=dev)
        self.prev = torch.zeros((n, self.nmb, 4), dtype=torch.int32, device=dev)
        self.has_prev = np.zeros(n, bool)
        self.qp = np.zeros(n, int)
        i8 = torch.zeros(self.nmb * 16, dtype=torch.int32, device=dev)
        k = 0
        t0 = time.time()
        for si, seed in enumerate(seeds):
            qp = QPS[(seed + seed_offset) % len(QPS)]
            g = pkg.FrameGen(seed, W, H, 12)
            enc = pkg.Encoder(W, H, search_range=RANGE, lambda8=LAMBDA8, deblock_idc=0, readback=0)
            try:
                for f in range(frames):
                    y, uv = g.frame(f)
                    im = enc.new_image()
                    enc.upload(im, y, uv)
                    if f > 0:
                        cy, _ = pkg.pad_coded(y, uv)
                        self.cur[k].copy_(torch.from_numpy(cy))
                        self.ref[k].copy_(rec)
                        rc = pkg.hip.bc250_launch_me_fullsearch(
                            C.c_void_p(self.cur[k].data_ptr()), wc, C.c_void_p(self.ref[k].data_ptr()), wc, wc, hc,
                            RANGE, LAMBDA8, C.c_void_p(self.fs[k].data_ptr()), C.c_void_p(i8.data_ptr()), None)
                        if rc != 0:
                            raise RuntimeError("full search failed")
                        if f > 1:
                            self.prev[k].copy_(self.fs[k - 1])
                            self.has_prev[k] = True
                        self.qp[k] = qp
                        k += 1
                    enc.encode(im, qp, f == 0, 1)
                    rec = enc.recon_luma_device(torch)
                    # images are kept only as long as the frame needs them
                    enc.images.remove(im)
                    pkg.hip.gpu_compute_destroy_image(C.byref(enc.ctx), im[0], im[1])
            finally:
                enc.close()
                g.close()
            log(f"data: seed {seed} qp {qp} ({si + 1}/{len(seeds)}, {time.time() - t0:.1f} s)")
        torch.cuda.synchronize()
        self.n = k


def tile_targets(ds, ln, torch, pkg):
    """Per P frame: the head's neighbour input [ntiles, 16, 32] and the
    targets [ntiles, 16, 2] of every 64x64 tile."""
    tw, th = (ds.mbw + 3) // 4, (ds.mbh + 3) // 4
    nt = tw * th
    mvin = torch.zeros((ds.n, nt, 16, 32), dtype=torch.float32, device="cuda")
    for k in range(ds.n):
        prev = C.c_void_p(ds.prev[k].data_ptr()) if ds.has_prev[k] else None
        if pkg.hip.bc250_cnn_mvin(prev, ds.mbw, ds.mbh, C.c_void_p(mvin[k].data_ptr()), None) != 0:
            raise RuntimeError("mvin failed")
    ty, tx = np.divmod(np.arange(nt), tw)
    j = np.arange(16)
    mx = np.minimum(tx[:, None] * 4 + (j % 4)[None, :], ds.mbw - 1)
    my = np.minimum(ty[:, None] * 4 + (j // 4)[None, :], ds.mbh - 1)
    mb = torch.from_numpy((my * ds.mbw + mx).reshape(-1)).cuda()
    tgt = ds.fs[:, mb, :2].float().reshape(ds.n, nt, 16, 2)
    return mvin, tgt, tw, th


def hit_rates(ds, pkg, ln, torch, params, ncand=4, window_only=True):
    """Fraction of MBs whose learned-search 16x16 record equals the full
    search's, with the network's candidates and with the no-network ones
    (zero, previous co-located, previous left, previous top)."""
    net = ln.Net(pkg, ((ds.mbw + 3) // 4) * ((ds.mbh + 3) // 4), params)
    mvin, _, tw, th = tile_targets(ds, ln, torch, pkg)
    nt = tw * th
    tiles = np.stack([np.zeros(nt, int), np.arange(nt) % tw, np.arange(nt) // tw], 1).astype(np.int32)
    cand = torch.zeros(ds.nmb * 8, dtype=torch.int32, device="cuda")
    o16 = torch.zeros(ds.nmb * 4, dtype=torch.int32, device="cuda")
    o8 = torch.zeros(ds.nmb * 16, dtype=torch.int32, device="cuda")
    out = {"network": [], "no_network": []}
    P = lambda t: C.c_void_p(t.data_ptr())
    for k in range(ds.n):
        for mode in out:
            if mode == "network":
                net.set_input(ds.cur[k], ds.ref[k], ds.wc, ds.wc, ds.hc, tiles)
                z = net.forward(nt, mvin[k], keep=False)
                if pkg.hip.bc250_cnn_candidates(P(z), ds.mbw, ds.mbh, RANGE, P(cand), None) != 0:
                    raise RuntimeError("candidates failed")
            else:
                pv = ds.prev[k, :, :2].reshape(ds.mbh, ds.mbw, 2) if ds.has_prev[k] else \
                    torch.zeros((ds.mbh, ds.mbw, 2), dtype=torch.int32, device="cuda")
                left = torch.cat([pv[:, :1], pv[:, :-1]], 1)
                top = torch.cat([pv[:1], pv[:-1]], 0)
                c = torch.stack([torch.zeros_like(pv), pv, left, top], 2).reshape(-1)
                cand.copy_(c.int())
            if pkg.hip.bc250_launch_cand_search(P(ds.cur[k]), ds.wc, P(ds.ref[k]), ds.wc, ds.wc, ds.hc, RANGE,
                                                LAMBDA8, P(cand), ncand, P(o16), P(o8), None) != 0:
                raise RuntimeError("candidate search failed")
            got = o16.view(ds.nmb, 4)[:, :3]
            out[mode].append(float((got == ds.fs[k, :, :3]).all(1).float().mean()))
    return {m: float(np.mean(v)) for m, v in out.items()}


def train(pkg, ln, torch, ds, steps, batch, lr, seed, log_every=100):
    mvin, tgt, tw, th = tile_targets(ds, ln, torch, pkg)
    nt = tw * th
    net = ln.Net(pkg, batch, ln.init_params(seed), train=True)
    rs = np.random.default_rng(seed)
    curve = []
    t0 = time.time()
    slot_stride = ds.hc * ds.wc
    for step in range(steps):
        pair = rs.integers(0, ds.n, batch)
        tile = rs.integers(0, nt, batch)
        tiles = np.stack([pair, tile % tw, tile // tw], 1).astype(np.int32)
        net.set_input(ds.cur, ds.ref, ds.wc, ds.wc, ds.hc, tiles, slot_stride=slot_stride)
        pi, ti = torch.from_numpy(pair).cuda(), torch.from_numpy(tile).cuda()
        mv = mvin[pi, ti].contiguous()
        tg = tgt[pi, ti].contiguous()
        net.forward(batch, mv)
        loss = net.backward(batch, mv, tg, 1.0 / (batch * 16))
        # step decay: the last third at a tenth of the rate
        net.adam(lr if step < 2 * steps // 3 else lr * 0.1)
        if step % log_every == 0 or step == steps - 1:
            l = float(loss.sum()) / (batch * 16)
            curve.append([step, round(l, 5)])
            log(f"train: step {step} loss/block {l:.4f} ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize()
    return net.params_numpy(), curve, time.time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=40, help="training sequences (seeds 1000..)")
    ap.add_argument("--frames", type=int, default=51, help="frames per sequence (I + P)")
    ap.add_argument("--eval-seeds", type=int, default=4, help="held-out sequences (seeds 9000..)")
    ap.add_argument("--eval-frames", type=int, default=51)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "me_proposer.f32"))
    ap.add_argument("--log", default=os.path.join(HERE, "weights", "me_proposer_train.json"))
    args = ap.parse_args()
    import torch
    pkg, ln = load()
    ds = Dataset(pkg, torch, list(range(1000, 1000 + args.seeds)), args.frames)
    log(f"data: {ds.n} P frames, {ds.n * ds.nmb} MB targets")
    params, curve, secs = train(pkg, ln, torch, ds, args.steps, args.batch, args.lr, seed=1)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    params.astype("<f4").tofile(args.out)
    del ds
    torch.cuda.empty_cache()
    ev = Dataset(pkg, torch, list(range(9000, 9000 + args.eval_seeds)), args.eval_frames, seed_offset=1)
    hits = hit_rates(ev, pkg, ln, torch, params)
    res = {"train_sequences": args.seeds, "frames_per_sequence": args.frames,
           "train_p_frames": args.seeds * (args.frames - 1), "steps": args.steps, "batch_tiles": args.batch,
           "lr": args.lr, "train_seconds": round(secs, 1), "loss_per_block": curve,
           "heldout_sequences": args.eval_seeds, "heldout_p_frames": ev.n,
           "heldout_hit_rate": hits}
    with open(args.log, "w") as f:
        json.dump(res, f, indent=1)
    log(json.dumps({k: v for k, v in res.items() if k != "loss_per_block"}))


if __name__ == "__main__":
    main()
