#!/usr/bin/env python3
"""Benchmark of the encoder compute pipeline on one MI355X.

One step = one 1920x1080 NV12 frame through the whole GPU pipeline, frames
already resident in HBM.  The headline (--codec hevc, the default) is the
HEVC Main pipeline: +-64 integer full search on 16x16 and 8x8 blocks,
quarter-sample refinement of every 8x8 CU with the HEVC interpolation, CU
coding (8x8 luma / 4x4 chroma DCT, quant, recon), I picture as the CTU
intra wavefront, then 8.7.2 deblocking.  The H.264 pipeline (the reference's gpu_compute stages,
with deblocking) runs on the same frames and is reported beside it.  The
timed region is an IPPP sequence of --steps frames driven by the host C
sequence driver (host/h264_host.c) through the gpu_compute_* ABI, pipelined
over the two slots, with readback on: every frame's outputs (levels, modes,
vectors) reach the slot's pinned host staging, as the reference's
gpu_compute.c:3328-3353 copies them, before the frame counts as done.  The
H.264 line with slices instead ends each frame with its CAVLC slices written
on the GPU and packed as Annex B NAL units on the host by the driver's C
consumer, so its output is the elementary stream.  Each timed sequence runs
--passes times (3 by default); the value is the run of median fps, and
every run's fps is listed beside it.

Prints ONE JSON line.  Besides the contract fields it carries the per-stage
GPU times from HIP events recorded on the pipeline's own streams in one more
run of the same frames, the same sequence at CQP 22 / 27 / 32 / 37, the roofline of
the dominant kernel (the integer full search), and a CPU baseline: the
oracle timed on a bounded sample of the same workload on this host, in the
same leg that checks a short run of every timed pipeline, frame by frame,
against the oracle (the run fails if any frame differs).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "1080p HEVC fps on 1 GPU, plain vs learned-ME at ≤1% BD-rate; p99 frame ms"
W, H = 1920, 1080
HBM_PEAK_GBS = 8000.0
# 256 CUs x 4 SIMD16 x 2.4 GHz x 4 byte-absdiffs per v_sad_u8 lane-op
SAD_PEAK_T = 256 * 64 * 2.4e9 * 4 / 1e12


def load_pkg():
    name = "bc250_vaapi_amd"
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "bc250-vaapi_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sad_peak(pkg, torch):
    """Sustained v_sad_u8 rate (T byte-absdiffs/s) from the calibration kernel."""
    import ctypes as C
    blocks, iters = 256 * 8, 64
    out = torch.zeros(blocks * 256, dtype=torch.int32, device="cuda")
    f = pkg.hip.bc250_calib_sad_peak
    f.restype, f.argtypes = C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    f(out.data_ptr(), blocks, iters, None)  # warm
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    f(out.data_ptr(), blocks, iters, None)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    return blocks * 256 * iters * 128 * 4 / (ms * 1e-3) / 1e12


def copy_calibration(pkg, torch, nbytes=64 << 20, reps=3):
    """A plain copy of a known size, launched so that a counter run over this
    process (FETCH_SIZE / WRITE_SIZE) has a known-bytes kernel to scale by."""
    import ctypes as C
    src = torch.full((nbytes,), 7, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    f = pkg.hip.bc250_calib_copy
    f.restype, f.argtypes = C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    for _ in range(reps):
        if f(src.data_ptr(), dst.data_ptr(), nbytes, None) != 0:
            raise RuntimeError("calibration copy failed")
    torch.cuda.synchronize()
    return nbytes


def _cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline_hevc(pkg, args):
    """The repo's CPU encoder (oracle/orc_baseline.c: the oracle's HEVC
    stages, AVX2 integer search, search and refine over OpenMP threads) on
    the host cores, timed per picture: the 640x480 CPU configuration (HEVC
    Main 8-bit, IPPP, CQP 32, 4 slices, +-32, 6 moving rectangles) and the
    1080p headline shape (+-64, CQP of the headline), each on all the
    threads the box gives this process (OMP_NUM_THREADS, else the CPU count,
    at most 16) and on one thread.  value: 1080p fps on all threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    model, ncpu = _cpu_info()
    nt = max(1, min(int(os.environ.get("OMP_NUM_THREADS", ncpu or 1)), ncpu or 1, 16))

    def seq(w, h, nrect, n, rng, qp, threads):
        g = pkg.FrameGen(480 if w == 640 else 1080, w, h, nrect)
        fr = [pkg.pad_coded(*g.frame(i)) for i in range(n)]
        g.close()
        ms, _, _ = orc.baseline_hevc_seq(np.stack([f[0] for f in fr]), np.stack([f[1] for f in fr]), qp, rng,
                                         args.lambda8, 4, threads)
        p = ms[1:]   # the P pictures
        return {"threads": threads, "frames": n, "fps": round(n / (ms.sum() / 1e3), 3),
                "mpixel_per_s": round(n * w * h / 1e6 / (ms.sum() / 1e3), 2),
                "p50_frame_ms": round(float(np.percentile(p, 50)), 2),
                "p99_frame_ms": round(float(np.percentile(p, 99)), 2), "i_frame_ms": round(float(ms[0]), 2)}

    log(f"cpu baseline: {model}, {ncpu} CPUs, {nt} threads")
    small = {"all": seq(640, 480, 6, 32, 32, 32, nt), "one": seq(640, 480, 6, 8, 32, 32, 1)}
    big = {"all": seq(W, H, 12, 4, args.range, args.qp, nt), "one": seq(W, H, 12, 2, args.range, args.qp, 1)}
    return {"value": big["all"]["fps"], "unit": "fps", "cores": nt, "kind": "port",
            "sample": f"the repo's CPU encoder (oracle stages, AVX2 +-{args.range} search over {nt} OpenMP threads) "
                      f"on 4 pictures (I + 3 P) of the 1080p sequence; 640x480 CPU configuration and 1-thread runs "
                      f"beside it",
            "cpu_model": model, "nproc": ncpu, "avx2": orc.baseline_has_avx2(),
            "hevc_640x480_ippp_cqp32_4slices_r32": small, "hevc_1080p": big}


def cpu_baseline(frames_coded, qp, rng, lam, budget_s, codec):
    """Oracle (single thread) on a bounded sample: the integer search of a
    spread sample of macroblocks, scaled to the frame, plus the rest of a
    P frame on the whole frame (HEVC: quarter-sample refine of every 8x8 CU,
    CU coding and deblocking; H.264: sub-pel refine, P encode and deblock)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    (ry, ruv), (cy, cuv) = frames_coded[0], frames_coded[1]
    hc, wc = cy.shape
    mbw, nmb = wc // 16, (wc // 16) * (hc // 16)
    t = time.perf_counter()
    orc.me_fullsearch_mb(cy, ry, rng, lam, mbw // 2, (hc // 16) // 2)
    per_mb = time.perf_counter() - t
    n = int(max(8, min(nmb, budget_s / max(per_mb, 1e-6))))
    sample = np.linspace(0, nmb - 1, n).astype(int)
    i16 = np.zeros(nmb, orc.IMV)
    i8 = np.zeros(nmb * 4, orc.IMV)
    t = time.perf_counter()
    for m in sample:
        a, b = orc.me_fullsearch_mb(cy, ry, rng, lam, m % mbw, m // mbw)
        i16[m], i8[4 * m:4 * m + 4] = a, b
    t_int = (time.perf_counter() - t) / n * nmb
    t = time.perf_counter()
    if codec == "hevc":
        mv = np.zeros((nmb * 4, 2), np.int32)
        for cu in range(nmb * 4):
            ctu, z = cu // 4, cu % 4
            x0, y0 = (ctu % mbw) * 16 + (z & 1) * 8, (ctu // mbw) * 16 + (z >> 1) * 8
            mx, my, _ = orc.hevc_refine8(cy, ry, x0, y0, int(i8[cu]["dx"]) * 4, int(i8[cu]["dy"]) * 4)
            mv[cu] = (mx, my)
        cb, cr = orc.split_uv(cuv)
        fb, fr = orc.split_uv(ruv)
        o = orc.hevc_encode_p(cy, cb, cr, ry, fb, fr, qp, mv, orc.HEVC_ROUND_INTER)
        mv4 = np.zeros((nmb * 4, 4), np.int32)
        mv4[:, :2] = mv
        uv = np.empty_like(cuv)
        uv[:, 0::2], uv[:, 1::2] = o["cb"], o["cr"]
        orc.hevc_deblock(o["y"], uv, qp, 0, orc.hevc_cu_cbf(o["coef"]), mv4, 1)
        rest = "quarter-sample refine of every 8x8 CU, HEVC CU coding and deblocking of the whole frame"
    else:
        mv16, _ = orc.me_refine(cy, ry, i16, i8)
        e = orc.encode_p(cy, cuv, ry, ruv, qp, mv16)
        orc.deblock(e["y"], e["uv"], qp, 0, e["nz"], mv16, 1, 0)
        rest = "sub-pel refine, P encode and deblock of the whole frame"
    t_rest = time.perf_counter() - t
    frame_s = t_int + t_rest
    return {"value": round(1.0 / frame_s, 5), "unit": "fps", "cores": 1, "kind": "port",
            "sample": f"oracle, 1 thread: +-{rng} integer search of {n} of {nmb} MBs of one 1080p P frame "
                      f"scaled to the frame ({t_int:.1f} s), plus {rest} ({t_rest:.1f} s)",
            "frame_s": round(frame_s, 2)}


def run_pipeline(pkg, torch, frames, codec, args, entropy=0, qps=None, proposer=None):
    """Uploads the frames to a fresh context and, for each QP, runs the
    warm-up and the timed IPPP sequence; returns, per QP, fps, per-frame
    latencies and the stage statistics from the pipeline's own HIP events.
    entropy 0: readback on (the staging reaches the host); entropy 1
    (H.264): every frame writes its CAVLC slices, which the driver's C
    consumer packs into an Annex B stream in host memory.  proposer: the
    learned integer search's weights (None: the plain full search)."""
    enc = pkg.Encoder(W, H, search_range=args.range, lambda8=args.lambda8, deblock_idc=0, me_8x8=1,
                      readback=0 if entropy else 1,
                      codec=pkg.GPU_CODEC_HEVC if codec == "hevc" else pkg.GPU_CODEC_H264, entropy=entropy)
    out = {}
    try:
        if proposer is not None:
            enc.set_proposer(proposer, **pkg.PROPOSER_DEFAULT)
        images = []
        for y, uv in frames:
            im = enc.new_image()
            enc.upload(im, y, uv)
            images.append(im)
        stream = np.zeros(64 << 20, np.uint8) if entropy else None
        sizes = np.zeros(args.steps, np.uint64)
        for qp in qps or [args.qp]:
            def seq(imgs):
                if not entropy:
                    return enc.encode_sequence(imgs, qp, gop=0, pipelined=1)
                sink = pkg.NalSink(0, stream.ctypes.data, stream.size, 0, sizes.ctypes.data)
                r = enc.encode_sequence(imgs, qp, gop=0, pipelined=1, nal_sink=sink)
                seq.bytes = sink.used
                return r
            log(f"{codec}{' + slices' if entropy else ''} qp {qp}: warmup")
            enc.set_stage_timing(False)
            seq(images[:args.warmup])
            # `passes` timed runs of the same K frames; the run of median fps
            # is the result (a host stall of a few ms in one run then moves
            # neither the value nor its p99), every run's fps is kept
            runs = []
            for k in range(max(args.passes, 1)):
                log(f"{codec}{' + slices' if entropy else ''} qp {qp}: timed {args.steps} frames IPPP, pass {k + 1}")
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                frame_ms, _ = seq(images[args.warmup:])
                torch.cuda.synchronize()
                wall = time.perf_counter() - t0
                runs.append(dict(wall=wall, fps=args.steps / wall, frame_ms=frame_ms,
                                 stream_bytes=int(seq.bytes) if entropy else 0))
            runs_by_fps = sorted(runs, key=lambda r: r["fps"])
            med = runs_by_fps[len(runs) // 2]
            out[qp] = dict(wall=med["wall"], fps=med["fps"], frame_ms=med["frame_ms"], errs=enc.error_flags(),
                           pass_fps=[round(r["fps"], 2) for r in runs])
            if entropy:
                out[qp]["stream_bytes"] = med["stream_bytes"]
            # the same frames again with a HIP event at every stage boundary
            # on the encoder's stream: per-stage GPU times (and the integer
            # search's launch time for the roofline), and what the events cost
            enc.set_stage_timing(True)
            enc.stage_stats(reset=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            seq(images[args.warmup:])
            torch.cuda.synchronize()
            out[qp]["fps_stage_events"] = args.steps / (time.perf_counter() - t0)
            out[qp]["stats"] = enc.stage_stats(reset=True)
            out[qp]["errs"] |= enc.error_flags()
        return out
    finally:
        enc.close()


def bd_rate(r0, p0, r1, p1):
    """Bjontegaard rate difference (%) of curve 1 against curve 0: cubic fits
    of log10(rate) over PSNR, averaged over the common PSNR interval."""
    c0, c1 = np.polyfit(p0, np.log10(r0), 3), np.polyfit(p1, np.log10(r1), 3)
    lo, hi = max(min(p0), min(p1)), min(max(p0), max(p1))
    i0, i1 = np.polyint(c0), np.polyint(c1)
    a0 = (np.polyval(i0, hi) - np.polyval(i0, lo)) / (hi - lo)
    a1 = (np.polyval(i1, hi) - np.polyval(i1, lo)) / (hi - lo)
    return (10 ** (a1 - a0) - 1) * 100


def learned_quality(pkg, frames, args, weights, qps=(22, 27, 32, 37), n=24):
    """Rate and quality of the learned integer search against the plain one
    on the bench's frames (seed 1080, outside the training and held-out
    seeds), I + (n-1) P frames per QP: the HEVC pipeline's pictures written
    by the host HEVC writer (the stream's bytes, VPS/SPS/PPS included) and
    PSNR-Y of every reconstruction -- the BD-rate of the metric -- and the
    same on the H.264 pipeline's CAVLC slices beside it; plus the fraction of
    MBs the learned form searched in full and of 16x16 vectors equal to the
    plain form's (the final vectors: HEVC 8x8 CUs, H.264 16x16 MBs)."""
    out = {}
    for codec in ("hevc", "h264"):
        res = {"plain": {}, "learned": {}}
        vec_same, fb = [], []
        for qp in qps:
            mvs = {}
            for mode in ("plain", "learned"):
                if codec == "hevc":
                    enc = pkg.Encoder(W, H, search_range=args.range, lambda8=args.lambda8, readback=1,
                                      codec=pkg.GPU_CODEC_HEVC)
                    st = pkg.HevcStream(enc.wc, enc.hc, 8, 4, 1, True, init_qp=26,
                                        crop_width=W if W != enc.wc else 0, crop_height=H if H != enc.hc else 0)
                else:
                    enc = pkg.Encoder(W, H, search_range=args.range, lambda8=args.lambda8, deblock_idc=0,
                                      readback=1, entropy=1)
                try:
                    if mode == "learned":
                        enc.set_proposer(weights, **pkg.PROPOSER_DEFAULT)
                    nbytes = len(st.headers()) if codec == "hevc" else 0
                    psnr, mv = [], []
                    for i, (y, uv) in enumerate(frames[:n]):
                        im = enc.new_image()
                        enc.upload(im, y, uv)
                        slot = enc.encode(im, qp, i == 0, 1)
                        if codec == "hevc":
                            r = enc.results(slot)
                            nbytes += len(st.picture(i == 0, i, qp, r["quant"], r["hevc_mode"], r["mv8"]))
                        else:
                            nbytes += len(pkg.h264_frame_nals(enc.slices(slot), i == 0))
                        if mode == "learned" and i > 0:
                            fb.append(enc.fallbacks(slot) / enc.nmb)
                        mv.append(enc.staging(slot, "mv8" if codec == "hevc" else "mv"))
                        ry, _ = enc.recon()
                        mse = np.mean((ry.astype(np.float64) - y) ** 2)
                        psnr.append(10 * np.log10(255.0 ** 2 / max(mse, 1e-10)))
                        enc.images.remove(im)
                        pkg.hip.gpu_compute_destroy_image(pkg.C.byref(enc.ctx), im[0], im[1])
                finally:
                    enc.close()
                res[mode][qp] = {"bytes": nbytes, "psnr_y": float(np.mean(psnr))}
                mvs[mode] = mv
            vec_same += [float(np.mean((a["mvx"] == b["mvx"]) & (a["mvy"] == b["mvy"])))
                         for a, b in zip(mvs["plain"][1:], mvs["learned"][1:])]
            log(f"learned quality {codec} qp {qp}: plain {res['plain'][qp]}, learned {res['learned'][qp]}")
        r0 = [res["plain"][q]["bytes"] for q in qps]
        p0 = [res["plain"][q]["psnr_y"] for q in qps]
        r1 = [res["learned"][q]["bytes"] for q in qps]
        p1 = [res["learned"][q]["psnr_y"] for q in qps]
        out[codec] = {
            "bd_rate_pct": round(bd_rate(r0, p0, r1, p1), 3),
            "per_qp": {str(q): {"plain_bytes": res["plain"][q]["bytes"], "learned_bytes": res["learned"][q]["bytes"],
                                "plain_psnr_y": round(res["plain"][q]["psnr_y"], 3),
                                "learned_psnr_y": round(res["learned"][q]["psnr_y"], 3)} for q in qps},
            "fallback_mb_frac": round(float(np.mean(fb)), 4),
            "vectors_equal_to_plain_frac": round(float(np.mean(vec_same)), 4)}
    h = out["hevc"]
    return {"bd_rate_pct": h["bd_rate_pct"],
            "bd_rate_basis": f"HEVC Main stream bytes (the host writer over the pipeline's staging, parameter sets "
                             f"included) and PSNR-Y, CQP {'/'.join(map(str, qps))}, {n} frames IPPP of the bench's "
                             f"1080p sequence",
            "per_qp": h["per_qp"], "fallback_mb_frac": h["fallback_mb_frac"],
            "vectors_equal_to_plain_frac": h["vectors_equal_to_plain_frac"],
            "h264_cavlc": {"bd_rate_pct": out["h264"]["bd_rate_pct"], "per_qp": out["h264"]["per_qp"],
                           "basis": "the H.264 pipeline's CAVLC slices, same frames, QPs and search"}}


def bench_rate_control(pkg, frames, args, fps=60.0, rate=20_000_000):
    """The host rate controller on the 1080p H.264 pipeline with its CAVLC
    slices (bc250_encode_sequence_rc, pipelined): low-latency mode, the
    buffer loop (the reference's H.264 choice) and the complexity model,
    at rate bit/s for fps pictures per second; I every 60."""
    out = {"mode": "low_latency", "target_mbps": rate / 1e6, "fps_nominal": fps}
    for name, model in (("buffer_loop", False), ("model", True)):
        enc = pkg.Encoder(W, H, search_range=args.range, lambda8=args.lambda8, deblock_idc=0, readback=0, entropy=1)
        try:
            ims = []
            for y, uv in frames:
                im = enc.new_image()
                enc.upload(im, y, uv)
                ims.append(im)
            rc = pkg.RateControl(pkg.RC_LOW_LATENCY, rate, fps, W, H, model=model)
            stream, qps, bits, _, tot = enc.encode_sequence_rc(ims, rc, gop=60, num_slices=4, pipelined=1)
        finally:
            enc.close()
        n = len(frames)
        out[name] = {"frames": n, "achieved_mbps": round(float(bits[min(12, n - 1):].mean()) * fps / 1e6, 3),
                     "qp_min": int(qps.min()), "qp_max": int(qps.max()), "qp_mean": round(float(qps.mean()), 2),
                     "encode_fps": round(n / (tot / 1e3), 1), "stream_bytes": len(stream)}
        log(f"rate control {name}: {out[name]}")
    return out


SERVE_MIX = [(640, 480)] * 4 + [(1280, 720)] * 8 + [(1920, 1080)] * 8 + [(2560, 1440)] * 4


def bench_serve(pkg, args, seconds=2.0, ring=4, sat_frames=40):
    """24 concurrent low-delay-P HEVC Main sessions (4 x 640x480, 8 x 1280x720,
    8 x 1920x1080, 4 x 2560x1440), each with its own seeded generator (ring
    source pictures resident on the device, coded in a cycle), CQP 27, +-64.
    Paced: Poisson frame arrivals at 60 fps per session for `seconds`; a
    frame's latency is its completion minus its arrival.  Saturated: every
    session's frames all arrive at once, for the aggregate throughput.  Plain
    and learned integer search, each on a fresh server (the learned one
    autotuned first), warmed by an I and a P picture per session.  Every
    session's first paced frame is an I picture arriving near t = 0, so the
    paced object also gives p99 over the frames that arrive after the last
    session's I picture is done (steady state)."""
    rs = np.random.default_rng(60)
    log(f"serve: generating {ring} source pictures for each of {len(SERVE_MIX)} sessions")
    sources = []
    for k, (w, h) in enumerate(SERVE_MIX):
        g = pkg.FrameGen(7000 + k, w, h, 12 if w >= 1920 else 6)
        sources.append([g.frame(i) for i in range(ring)])
        g.close()
    paced = []
    for k in range(len(SERVE_MIX)):
        t, f = 0.0, 0
        while True:
            t += rs.exponential(1 / 60)
            if t > seconds:
                break
            paced.append((k, f, f % ring, t))
            f += 1
    paced.sort(key=lambda r: r[3])
    sat = sorted([(k, f, f % ring, 0.0) for f in range(sat_frames) for k in range(len(SERVE_MIX))],
                 key=lambda r: r[1])
    mpx = {k: w * h / 1e6 for k, (w, h) in enumerate(SERVE_MIX)}
    out = {}
    for mode in ("plain", "learned"):
        srv = pkg.Server(learned=mode == "learned", max_batch=32)
        try:
            sids = []
            for k, (w, h) in enumerate(SERVE_MIX):
                sid = srv.open(w, h, codec=pkg.GPU_CODEC_HEVC, qp=27, ring=ring)
                if sid < 0:
                    raise RuntimeError(f"session {k} ({w}x{h}) not admitted")
                for i, (y, uv) in enumerate(sources[k]):
                    srv.upload(sid, i, y, uv)
                sids.append(sid)
            if mode == "learned":
                srv.autotune()
            # warm-up: an I and a P picture per session (kernels loaded, buffers touched)
            srv.run([(sids[k], f, f % ring, 0.0) for f in range(2) for k in range(len(SERVE_MIX))])
            log(f"serve {mode}: paced run, {len(paced)} frames over {seconds} s")
            res, wall = srv.run([(sids[k], f, src, t) for k, f, src, t in paced])
            lat = (res["t_done"] - res["t_arrival"]) * 1e3
            # every session opens with an I picture near t = 0; steady state
            # starts once the last of those is done
            t_steady = max(float(res["t_done"][(res["session"] == sids[k]) & (res["frame"] == 0)].max())
                           for k in range(len(SERVE_MIX)))
            st_lat = lat[res["t_arrival"] > t_steady]
            log(f"serve {mode}: saturated run, {len(sat)} frames")
            res2, wall2 = srv.run([(sids[k], f, src, t) for k, f, src, t in sat])
            st = srv.stats()
        finally:
            srv.close()
        px = sum(mpx[k] for k, _, _, _ in sat)
        out[mode] = {
            "paced": {"frames": len(paced), "wall_s": round(wall, 3), "fps": round(len(paced) / wall, 1),
                      "p50_frame_ms": round(float(np.percentile(lat, 50)), 3),
                      "p99_frame_ms": round(float(np.percentile(lat, 99)), 3),
                      "max_frame_ms": round(float(lat.max()), 3),
                      "steady_from_s": round(t_steady, 4),
                      "steady_p99_frame_ms": round(float(np.percentile(st_lat, 99)), 3),
                      "steady_max_frame_ms": round(float(st_lat.max()), 3),
                      "within_16_7_ms": round(float((lat <= 16.7).mean()), 4),
                      "p99_by_size_ms": {f"{w}x{h}": round(float(np.percentile(
                          lat[np.isin(res["session"], [sids[k] for k, sz in enumerate(SERVE_MIX) if sz == (w, h)])],
                          99)), 3) for (w, h) in sorted(set(SERVE_MIX))}},
            "saturated": {"frames": len(sat), "wall_s": round(wall2, 3), "fps": round(len(sat) / wall2, 1),
                          "mpixel_per_s": round(px / wall2, 1)},
            "batches": st["batches"], "mean_batch": round(st["mean_batch"], 2), "max_batch": st["max_batch"],
            "hbm_bytes": st["hbm_used"],
        }
        if mode == "learned":
            out[mode]["autotune"] = {"forward_variant": st["forward_variant"],
                                     "us_by_variant_0_1_2": [round(x, 1) for x in st["autotune_us"]]}
            fb = res["fallbacks"][res["fallbacks"] >= 0]
            out[mode]["fallback_mbs_per_batch"] = round(float(fb.mean()), 1) if fb.size else None
        log(f"serve {mode}: {out[mode]['paced']}")
    return {"workload": "24 sessions HEVC Main 8-bit (4x640x480, 8x1280x720, 8x1920x1080, 4x2560x1440), "
                        f"low-delay P, CQP 27, +-64, Poisson arrivals at 60 fps each for {seconds} s; "
                        f"saturated: {sat_frames} frames per session at once", **out}


def verify_pipelines(pkg, frames, args, codecs, n=4):
    """The checker half of the CPU leg: a fresh context per timed pipeline
    encodes the first n frames exactly as timed (sequence driver, pipelined,
    the same parameters) with a consumer that reads every frame's outputs,
    and tests/seqcheck.py compares them with the oracle frame by frame.
    Raises on the first difference."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import seqcheck
    coded = [pkg.pad_coded(*f) for f in frames[:n]]
    done = []
    for codec, entropy in codecs:
        enc = pkg.Encoder(W, H, search_range=args.range, lambda8=args.lambda8, deblock_idc=0, me_8x8=1, readback=1,
                          codec=pkg.GPU_CODEC_HEVC if codec == "hevc" else pkg.GPU_CODEC_H264, entropy=entropy)
        try:
            images = []
            for y, uv in frames[:n]:
                im = enc.new_image()
                enc.upload(im, y, uv)
                images.append(im)
            _, _, got = enc.encode_sequence(images, args.qp, gop=0, pipelined=1, collect=True)
        finally:
            enc.close()
        log(f"checking {codec}{' + slices' if entropy else ''} ({n} frames) against the oracle")
        if codec == "hevc":
            seqcheck.check_hevc(coded, got, args.qp, 1, args.range, args.lambda8, sample=32)
        else:
            seqcheck.check_h264(coded, got, args.qp, 1, args.range, args.lambda8, sample=32)
        done.append(codec + (" with CAVLC slices" if entropy else ""))
    return (f"{', '.join(done)}: {n} frames (I + {n - 1} P) each through the timed path, every frame's outputs "
            "equal to the oracle's (whole-frame coding and slices; the +-64 search on 32 sampled MBs + the corners)")


def bench_main10_4k(pkg, torch, args):
    """The 4K Main10 workload (not the headline): 3840x2160 P010 (10-bit
    samples = 4 x the generator's 8-bit output + uniform [0, 3]), CTU32,
    IPPP, CQP 27, 4 slices, +-64 full search, deblocking; one JSON line."""
    w, h = 3840, 2160
    n = args.warmup + args.steps
    log(f"generating {n} frames {w}x{h} P010")
    g = pkg.FrameGen(2160, w, h, 12)
    rs = np.random.default_rng(2160)
    frames = []
    for i in range(n):
        y, uv = g.frame(i)
        frames.append((y.astype(np.uint16) * 4 + rs.integers(0, 4, y.shape, dtype=np.uint16),
                       uv.astype(np.uint16) * 4 + rs.integers(0, 4, uv.shape, dtype=np.uint16)))
    g.close()
    enc = pkg.Encoder(w, h, search_range=args.range, lambda8=args.lambda8, deblock_idc=0, me_8x8=1, readback=1,
                      codec=pkg.GPU_CODEC_HEVC, bit_depth=10, ctu_log2=5)
    try:
        images = []
        for y, uv in frames:
            im = enc.new_image()
            enc.upload(im, y, uv)
            images.append(im)
        del frames
        log("main10 4k: warmup")
        enc.encode_sequence(images[:args.warmup], args.qp, gop=0, num_slices=4, pipelined=1)
        log(f"main10 4k: timed {args.steps} frames IPPP")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        frame_ms, _ = enc.encode_sequence(images[args.warmup:], args.qp, gop=0, num_slices=4, pipelined=1)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        errs = enc.error_flags()
        enc.set_stage_timing(True)   # stage times from a second pass with events
        enc.stage_stats(reset=True)
        enc.encode_sequence(images[args.warmup:], args.qp, gop=0, num_slices=4, pipelined=1)
        st, errs = enc.stage_stats(reset=True), errs | enc.error_flags()
    finally:
        enc.close()
    names = ["me_integer", "me_subpel", "encode", "deblock", "staging"]
    res = {
        "metric": "4K HEVC Main10 fps on 1 GPU", "value": round(args.steps / wall, 2), "unit": "fps", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": f"3840x2160 P010, {args.steps} frames IPPP, CQP {args.qp}, 4 slices: HEVC Main10 "
                               f"pipeline (CTU32), +-{args.range} full search on the top 8 bits, 10-bit refine, "
                               "CU coding, CTU intra wavefront, 8.7.2 deblocking",
                   "codec": "hevc", "width": w, "height": h, "bit_depth": 10, "frames": args.steps, "qp": args.qp,
                   "search_range": args.range, "num_slices": 4, "ctu": 32, "seed": 2160, "rectangles": 12},
        "p50_frame_ms": round(float(np.percentile(frame_ms, 50)), 4),
        "p99_frame_ms": round(float(np.percentile(frame_ms, 99)), 4),
        "mpixel_per_s": round(args.steps / wall * w * h / 1e6, 1),
        "stage_ms": {k: round(st["sum_ms"][i] / max(st["frames"], 1), 4) for i, k in enumerate(names)},
        "wavefront_wait_errors": errs,
    }
    print(json.dumps(res))


def bench_t32(pkg, torch, args):
    """The 32x32 transform forms compared (not the headline): 64800 blocks
    (eight 3840x2160 luma planes as 32x32 TUs) of 10-bit residuals, forward
    + quant and dequant + inverse at Qp' 39, each form launched --steps times
    after --warmup launches, timed with HIP events on the launch stream.
    Algorithmic bytes per block: 2048 read + 2048 written."""
    import ctypes as C
    n, bd, qp = 64800, 10, 39
    rs = np.random.default_rng(32)
    res = rs.integers(-1023, 1024, (n, 1024)).astype(np.int16)
    res[1::2] //= 13
    dres = torch.from_numpy(res).cuda()
    forms = {"butterfly": pkg.T32_BUTTERFLY, "mfma": pkg.T32_MFMA, "matrix": pkg.T32_MATRIX}
    outs, coefs, out = {}, {}, {}
    for name, m in forms.items():
        lv = torch.empty_like(dres)
        back = torch.empty_like(dres)
        fwd = lambda: pkg.hip.bc250_hevc_tq32_batch(C.c_void_p(dres.data_ptr()), n, qp, bd, 2, m,
                                                    C.c_void_p(lv.data_ptr()), None)
        inv = lambda: pkg.hip.bc250_hevc_dqit32_batch(C.c_void_p(lv.data_ptr()), n, qp, bd, m,
                                                      C.c_void_p(back.data_ptr()), None)
        row = {}
        for d, f in (("fwd", fwd), ("inv", inv)):
            log(f"t32 {name} {d}: {args.warmup} warmup + {args.steps} timed launches")
            for _ in range(args.warmup):
                if f() != 0:
                    raise RuntimeError(f"{name} {d} launch failed")
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(args.steps):
                f()
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / args.steps
            gbs = n * 4096 / (us * 1e-6) / 1e9
            row[d] = {"us_per_launch": round(us, 2), "blocks_per_s": round(n / (us * 1e-6)),
                      "GB_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        outs[name], coefs[name], out[name] = lv, back, row
    same = all(torch.equal(outs[k], outs["matrix"]) and torch.equal(coefs[k], coefs["matrix"]) for k in forms)
    best = min(forms, key=lambda k: out[k]["fwd"]["us_per_launch"])
    fw = out[best]["fwd"]
    print(json.dumps({
        "metric": "HEVC 32x32 transform + quant blocks/s on 1 GPU", "value": fw["blocks_per_s"],
        "unit": "blocks/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(fw["us_per_launch"] / 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int16", "data": "synthetic",
        "config": {"workload": f"{n} 32x32 blocks of 10-bit residuals (eight 3840x2160 luma planes), Qp' {qp}",
                   "blocks": n, "bit_depth": bd, "qp": qp},
        "forms": out, "fastest_forward": best, "forms_identical": bool(same),
        "roofline": {"bound": "hbm", "achieved": fw["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": fw["hbm_frac"], "traffic": None},
    }))


def bench_cavlc(pkg, torch, args):
    """The GPU H.264 CAVLC slice writer (not the headline): the 1080p H.264
    pipeline's own staging (levels, DC terms, modes, vectors; +-64, CQP
    --qp, 4 slices) of an I frame and 7 P frames, resident in HBM; each
    frame's slices are checked byte for byte against the oracle, then
    --steps writer calls cycling over the 8 frames are timed with HIP events
    on the launch stream.  Algorithmic bytes per frame: the staging read once
    (nmb x (768 + 96 + 16 + 4)) plus the RBSP written."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    import seqcheck
    wmb, hmb, ns, nf = (W + 15) // 16, (H + 15) // 16, 4, 8
    nmb = wmb * hmb
    log(f"cavlc: encoding {nf} frames {W}x{H} for their staging")
    g = pkg.FrameGen(1080, W, H, 12)
    pics = [g.frame(f) for f in range(nf)]
    g.close()
    enc = pkg.Encoder(W, H, search_range=args.range, lambda8=args.lambda8, deblock_idc=0, readback=1)
    try:
        images = []
        for y, uv in pics:
            im = enc.new_image()
            enc.upload(im, y, uv)
            images.append(im)
        _, _, outs = enc.encode_sequence(images, args.qp, gop=0, num_slices=ns, pipelined=1, collect=True)
    finally:
        enc.close()
    # the staging must be the pipeline's true output before the writer is timed on it
    log("cavlc: checking the staging against the oracle")
    seqcheck.check_h264([pkg.pad_coded(*f) for f in pics], outs, args.qp, ns, args.range, args.lambda8, sample=32)
    frames = [(r, pkg.h264_slice_cfg(wmb, hmb, args.qp, f == 0, ns, frame_num=f, poc=2 * f))
              for f, r in enumerate(outs)]
    dev = [tuple(torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in
                 (r["quant"], r["dc"], r["pred_mode"].view(np.int32), r["mv"].view(np.int32).reshape(-1)))
           for r, _ in frames]
    writer = pkg.CavlcWriter(wmb, hmb)
    sizes = []
    for (r, cfg), d in zip(frames, dev):
        got = writer.slices(*d, cfg)
        want = orc.h264_cavlc_slices(r["quant"], r["dc"], r["pred_mode"], r["mv"], cfg)
        if got != want:
            raise RuntimeError("GPU CAVLC slices differ from the oracle's")
        sizes.append(sum(len(x) for x in got))
    log(f"cavlc: bit-exact on {nf} frames ({', '.join(str(x) for x in sizes)} bytes); "
        f"{args.warmup} warmup + {args.steps} timed calls")
    st = torch.cuda.current_stream()
    for i in range(args.warmup):
        writer.launch(*dev[i % nf], frames[i % nf][1])
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for i in range(args.steps):
        writer.launch(*dev[i % nf], frames[i % nf][1])
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    out_bytes = sum(sizes[i % nf] for i in range(args.steps)) / args.steps
    alg = nmb * (768 + 96 + 16 + 4) + out_bytes
    gbs = alg / (ms * 1e-3) / 1e9
    traffic = None  # HBM bytes per frame from the FETCH_SIZE / WRITE_SIZE profile
    tfile = os.path.join(ROOT, "profiles", "cavlc_1080p_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f)["hbm_bytes_per_frame"]
    cpu = None
    if not args.no_cpu:
        log("cavlc: timing the oracle writer on the host")
        t0, k = time.perf_counter(), 0
        while time.perf_counter() - t0 < min(args.cpu_budget, 10.0):
            r, cfg = frames[k % nf]
            orc.h264_cavlc_slices(r["quant"], r["dc"], r["pred_mode"], r["mv"], cfg)
            k += 1
        cpu_s = (time.perf_counter() - t0) / k
        cpu = {"value": round(1.0 / cpu_s, 2), "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"oracle CAVLC writer, 1 thread, {k} calls over the same {nf} staged 1080p frames"}
    print(json.dumps({
        "metric": "1080p H.264 CAVLC slices written per second on 1 GPU (staging resident in HBM)",
        "value": round(1e3 / ms, 1), "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int16", "data": "synthetic",
        "config": {"workload": f"CAVLC slices of the 1080p H.264 pipeline's staging, I + 7 P frames, CQP {args.qp}, "
                               f"{ns} slices, +-{args.range}", "frames": nf, "slices": ns, "qp": args.qp,
                   "mean_frame_bytes": round(out_bytes)},
        "roofline": {"bound": "hbm", "kernel": "cavlc writer (5 launches)", "achieved": round(gbs, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "algorithmic_bytes": round(alg)},
        "cpu_baseline": cpu,
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--passes", type=int, default=3,
                    help="timed runs of the K frames per pipeline and QP; the value is the run of median fps")
    ap.add_argument("--qp", type=int, default=27, help="QP of the headline value")
    ap.add_argument("--qps", default="22,27,32,37", help="QPs of the qp_sweep object (comma list; '' for none)")
    ap.add_argument("--range", type=int, default=64)
    ap.add_argument("--lambda8", type=int, default=1)
    ap.add_argument("--codec", choices=("hevc", "h264"), default="hevc",
                    help="pipeline the headline value measures (the other one is reported beside it)")
    ap.add_argument("--single", action="store_true", help="skip the second codec's run")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of oracle integer search")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-ME line")
    ap.add_argument("--no-serve", action="store_true", help="skip the 24-session serving object")
    ap.add_argument("--workload", choices=("1080p", "4k-main10", "t32", "cavlc", "serve"), default="1080p",
                    help="4k-main10: the 3840x2160 P010 HEVC Main10 workload instead of the 1080p headline; "
                         "t32: the 32x32 transform forms compared; cavlc: the H.264 CAVLC slice writer")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the pipeline shards by session (one GPU per group of sessions); "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import torch
    pkg = load_pkg()
    if args.workload == "4k-main10":
        bench_main10_4k(pkg, torch, args)
        return
    if args.workload == "t32":
        bench_t32(pkg, torch, args)
        return
    if args.workload == "cavlc":
        bench_cavlc(pkg, torch, args)
        return
    if args.workload == "serve":
        sv = bench_serve(pkg, args)
        lp = sv["learned"]["paced"]
        print(json.dumps({
            "metric": "24-session HEVC serving on 1 GPU: p99 frame latency (ms) at 60 fps per session",
            "value": lp["p99_frame_ms"], "unit": "ms", "n_gpus": 1, "steps": lp["frames"], "warmup": 0,
            "ms_per_step": round(lp["wall_s"] * 1e3 / lp["frames"], 4), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": sv["workload"]}, "serving": sv}))
        return
    n = args.warmup + args.steps
    log(f"generating {n} frames {W}x{H}")
    g = pkg.FrameGen(1080, W, H, 12)
    frames = [g.frame(i) for i in range(n)]
    g.close()
    other = "h264" if args.codec == "hevc" else "hevc"
    qps = sorted(set([args.qp] + [int(q) for q in args.qps.split(",") if q]))
    sweeps = {args.codec: run_pipeline(pkg, torch, frames, args.codec, args, qps=qps)}
    if not args.single:
        sweeps[other] = run_pipeline(pkg, torch, frames, other, args, qps=qps)
        sweeps["h264_cavlc"] = run_pipeline(pkg, torch, frames, "h264", args, entropy=1, qps=qps)
    weights = None
    if not args.no_learned and os.path.exists(pkg.WEIGHTS):
        weights = pkg.load_proposer_weights()
        sweeps["hevc_learned"] = run_pipeline(pkg, torch, frames, "hevc", args, qps=[args.qp], proposer=weights)
    runs = {k: v[args.qp] for k, v in sweeps.items()}
    head = runs[args.codec]
    st = head["stats"]

    wc, hc = (W + 15) // 16 * 16, (H + 15) // 16 * 16
    nmb = (wc // 16) * (hc // 16)
    p_frames = max(st["p_frames"], 1)
    me_ms = st["me_int_ms_p"] / p_frames  # average integer-search kernel time per launch
    stage_names = ["me_integer", "me_subpel", "encode", "deblock", "staging"]

    def stage_ms(stats):
        return {k: round(stats["sum_ms"][i] / max(stats["frames"], 1), 4) for i, k in enumerate(stage_names)}

    cand = (2 * args.range + 1) ** 2
    sad_ops = cand * nmb * 256                 # byte-absdiffs per launch
    hbm_bytes = 2 * wc * hc + nmb * 16 * 5     # cur + ref luma read, 5 MV records per MB written
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "me_fullsearch_1080p_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    peak_meas = sad_peak(pkg, torch)
    copy_calibration(pkg, torch)
    if args.codec == "hevc":
        desc = ("HEVC Main 8-bit pipeline: I picture as CTU16/CU8/PU4 intra (Planar/DC/H/V by SAD, 4x4 DST), "
                "P pictures as 8x8 inter CUs with +-{r} integer full search (16x16 and 8x8 costs) refined to "
                "quarter samples with the HEVC 8-tap filter, 8x8 luma / 4x4 chroma DCT, quant, recon, "
                "8.7.2 deblocking").format(r=args.range)
    else:
        desc = ("H.264 pipeline (the reference's gpu_compute stages): +-{r} integer full search on 16x16 and 8x8 "
                "+ quarter-pel refine, P16x16 / I16x16 encode with 4x4 transform and quant, "
                "H.264 8.7 deblock").format(r=args.range)
    res = {
        "metric": METRIC,
        "value": round(head["fps"], 2),
        "unit": "fps",
        "passes": head["pass_fps"],
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["wall"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"1920x1080 NV12, {args.steps} frames IPPP, CQP {args.qp}: " + desc,
            "codec": args.codec, "width": W, "height": H, "frames": args.steps, "qp": args.qp,
            "search_range": args.range, "lambda8": args.lambda8, "seed": 1080, "rectangles": 12,
        },
        "p50_frame_ms": round(float(np.percentile(head["frame_ms"], 50)), 4),
        "p99_frame_ms": round(float(np.percentile(head["frame_ms"], 99)), 4),
        "mpixel_per_s": round(head["fps"] * W * H / 1e6, 1),
        "readback": "on: every frame's levels (P frames packed on the GPU: the 16-level groups holding a nonzero "
                    "level), modes and vectors copied to pinned host staging inside the timed region; the HEVC "
                    "writer (host CABAC) turns them into the stream outside it; the H.264 line with slices ends "
                    "in the Annex B stream in host memory",
        "stage_ms": stage_ms(st),
        "stage_events": "stage_ms and the roofline's kernel time come from HIP events at every stage boundary on "
                        "the encoder's stream, recorded in a second pass over the same timed frames; the first "
                        "pass (value) runs without them",
        "fps_with_stage_events": round(head["fps_stage_events"], 2),
        "wavefront_wait_errors": head["errs"],
        "qp_sweep": {name: {str(q): {"fps": round(r["fps"], 2),
                                     "p99_frame_ms": round(float(np.percentile(r["frame_ms"], 99)), 4),
                                     **({"stream_bytes_per_frame": round(r["stream_bytes"] / args.steps)}
                                        if "stream_bytes" in r else {})}
                             for q, r in sw.items()} for name, sw in sweeps.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": "me_fullsearch_kernel<64>",
            "achieved": round(hbm_bytes / (me_ms * 1e-3) / 1e9, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(hbm_bytes / (me_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 6),
            "traffic": traffic,
            "kernel_ms": round(me_ms, 4),
            "algorithmic_bytes": hbm_bytes,
        },
        "roofline_valu": {
            "bound": "valu (v_sad_u8)",
            "achieved": round(sad_ops / (me_ms * 1e-3) / 1e12, 3),
            "peak": round(SAD_PEAK_T, 1),
            "peak_measured": round(peak_meas, 1),
            "unit": "T byte-absdiff/s",
            "frac": round(sad_ops / (me_ms * 1e-3) / 1e12 / SAD_PEAK_T, 4),
            "algorithmic_ops": sad_ops,
        },
    }
    if other in runs:
        o = runs[other]
        res[f"{other}_pipeline"] = {"fps": round(o["fps"], 2), "ms_per_step": round(o["wall"] * 1e3 / args.steps, 4),
                                    "p99_frame_ms": round(float(np.percentile(o["frame_ms"], 99)), 4),
                                    "stage_ms": stage_ms(o["stats"]), "wavefront_wait_errors": o["errs"]}
    if "h264_cavlc" in runs:
        # the H.264 pipeline ending in its CAVLC slices (entropy = 1): the
        # "staging" stage is the slice writer plus the copy of the slices;
        # the host packs them into the Annex B stream inside the timed region
        o = runs["h264_cavlc"]
        res["h264_pipeline_with_cavlc"] = {"fps": round(o["fps"], 2),
                                           "ms_per_step": round(o["wall"] * 1e3 / args.steps, 4),
                                           "p99_frame_ms": round(float(np.percentile(o["frame_ms"], 99)), 4),
                                           "stream_bytes_per_frame": round(o["stream_bytes"] / args.steps),
                                           "stage_ms": stage_ms(o["stats"]), "wavefront_wait_errors": o["errs"]}
    if weights is not None:
        lr, pl = runs["hevc_learned"], runs["hevc"]
        res["learned_me"] = {
            "weights": os.path.relpath(pkg.WEIGHTS, ROOT),
            "hevc_fps": round(lr["fps"], 2), "plain_hevc_fps": round(pl["fps"], 2),
            "speedup": round(lr["fps"] / pl["fps"], 3),
            "passes": lr["pass_fps"],
            "p99_frame_ms": round(float(np.percentile(lr["frame_ms"], 99)), 4),
            "stage_ms": stage_ms(lr["stats"]),
            "fps_with_stage_events": round(lr["fps_stage_events"], 2),
            "search": "network over the picture's 64x64 tiles, 4 candidates per MB searched in +-4 windows, "
                      "MBs above the neighbour SAD bound searched in full (+-64)",
        }
        log("learned ME: rate and quality against plain")
        res["learned_me"].update(learned_quality(pkg, frames, args, weights))
    if not args.single:
        res["rate_control"] = bench_rate_control(pkg, frames, args)
    if not args.no_serve:
        res["serving"] = bench_serve(pkg, args)
    if not args.no_cpu:
        log("cpu leg: checking the timed pipelines against the oracle")
        pipes = [(args.codec, 0)] + ([(other, 0), ("h264", 1)] if not args.single else [])
        checked = verify_pipelines(pkg, frames, args, pipes)
        log("cpu baseline (oracle)")
        coded = [pkg.pad_coded(*frames[args.warmup]), pkg.pad_coded(*frames[args.warmup + 1])]
        if args.codec == "hevc":
            res["cpu_baseline"] = cpu_baseline_hevc(pkg, args)
        else:
            res["cpu_baseline"] = cpu_baseline(coded, args.qp, args.range, args.lambda8, args.cpu_budget, args.codec)
        res["checked_against_oracle"] = checked
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res))


if __name__ == "__main__":
    main()
