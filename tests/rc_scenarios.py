"""Seeded rate-control scenarios: a synthetic stream whose pictures cost
bits = complexity * 2^(-(qp - 12) / 5.3) (x4 for I pictures, log-normal
noise, scene cuts, still stretches), with bitrate changes, picture-size
limits and quality levels on the way.  run() drives any controller object
with the reference's call order (QP before the picture, bits after it) and
returns the per-picture trace."""
import ctypes as C
import os

import numpy as np

MODES = {"cbr": 0, "vbr": 1, "low_latency": 2, "cqp": 3}
SHAPES = [(640, 480, 30.0, 1_500_000), (1280, 720, 60.0, 10_000_000), (1920, 1080, 60.0, 20_000_000),
          (2560, 1440, 60.0, 40_000_000), (1920, 1080, 30.0, 2_000_000)]


def scenario(seed, n=240):
    rs = np.random.default_rng(seed)
    w, h, fps, rate = SHAPES[seed % len(SHAPES)]
    gop = int(rs.choice([0, 30, 60]))
    base = w * h * float(rs.uniform(0.5, 8.0))  # bits of a P picture at QP 12
    pics = []
    cplx = base
    still_left = 0
    for i in range(n):
        if rs.random() < 0.02:
            cplx = base * float(rs.uniform(0.2, 6.0))   # a cut
        if still_left == 0 and rs.random() < 0.03:
            still_left = int(rs.integers(3, 20))
        still = still_left > 0 and i > 0
        still_left = max(0, still_left - 1)
        ev = None
        u = rs.random()
        if u < 0.015:
            ev = ("bitrate", int(rate * rs.uniform(0.3, 3.0)))
        elif u < 0.025:
            ev = ("max_frame", int(rate / fps * rs.uniform(0.5, 3.0)))
        elif u < 0.03:
            ev = ("quality", int(rs.integers(0, 9)))
        pics.append({"intra": i == 0 or (gop > 0 and i % gop == 0), "still": bool(still),
                     "sad": int(cplx / 64 * rs.uniform(0.5, 1.5)), "noise": float(rs.lognormal(0, 0.25)),
                     "cplx": cplx, "event": ev})
    return {"w": w, "h": h, "fps": fps, "rate": rate, "pics": pics}


def bits_of(p, qp):
    b = p["cplx"] * 2.0 ** (-(qp - 12) / 5.3) * p["noise"] * (4.0 if p["intra"] else 1.0)
    if p["still"]:
        b *= 0.05
    return max(1, min(int(b), 2**31 - 1))


def run(ctl, sc, model):
    """ctl: init(mode-less), frame_qp(sad, intra, still), coded(bits, qp,
    intra, still), update_bitrate(rate, w, h), set_max_frame_size(bits),
    set_quality_level(level), state() -> (fullness, integral, debt)."""
    out = []
    for p in sc["pics"]:
        ev = p["event"]
        if ev and ev[0] == "bitrate":
            ctl.update_bitrate(ev[1], sc["w"], sc["h"])
        elif ev and ev[0] == "max_frame":
            ctl.set_max_frame_size(ev[1])
        elif ev and ev[0] == "quality":
            ctl.set_quality_level(ev[1])
        qp = ctl.frame_qp(p["sad"], p["intra"], p["still"])
        bits = bits_of(p, qp)
        ctl.coded(bits, qp, p["intra"], p["still"])
        out.append([qp, bits] + list(ctl.state()))
    return out


class RefRc:
    """The reference's rate_control.c (oracle/_ref/libref_rc.so), the
    checker."""
    LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle", "_ref", "libref_rc.so")

    def __init__(self, mode, rate, fps, w, h, model=False, current_qp=0):
        lib = self.lib = C.CDLL(self.LIB)
        lib.ref_rc_new.restype = C.c_void_p
        for f, r, a in (("rc_init", None, [C.c_void_p, C.c_int, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32]),
                        ("rc_update_bitrate", None, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]),
                        ("rc_get_frame_qp", C.c_int, [C.c_void_p, C.c_uint64]),
                        ("rc_update_stats", None, [C.c_void_p, C.c_int]),
                        ("rc_model_frame_qp", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
                        ("rc_model_frame_coded", None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
                        ("rc_set_max_frame_size", None, [C.c_void_p, C.c_uint32]),
                        ("rc_set_quality_level", None, [C.c_void_p, C.c_uint32]),
                        ("ref_rc_free", None, [C.c_void_p]), ("ref_rc_set_model", None, [C.c_void_p, C.c_int]),
                        ("ref_rc_set_current_qp", None, [C.c_void_p, C.c_int]),
                        ("ref_rc_fullness", C.c_longlong, [C.c_void_p]),
                        ("ref_rc_integral", C.c_longlong, [C.c_void_p]), ("ref_rc_debt", C.c_double, [C.c_void_p])):
            getattr(lib, f).restype = r
            getattr(lib, f).argtypes = a
        self.p = lib.ref_rc_new()
        self.model = model
        lib.ref_rc_set_current_qp(self.p, current_qp)
        lib.rc_init(self.p, mode, rate, fps, w, h)
        lib.ref_rc_set_model(self.p, int(model))

    def __del__(self):
        self.lib.ref_rc_free(self.p)

    def frame_qp(self, sad, intra, still):
        if self.model:
            return self.lib.rc_model_frame_qp(self.p, int(intra), int(still))
        return self.lib.rc_get_frame_qp(self.p, sad)

    def coded(self, bits, qp, intra, still):
        self.lib.rc_update_stats(self.p, bits)
        if self.model:
            self.lib.rc_model_frame_coded(self.p, int(intra), int(still), qp, bits)

    def update_bitrate(self, r, w, h):
        self.lib.rc_update_bitrate(self.p, r, w, h)

    def set_max_frame_size(self, b):
        self.lib.rc_set_max_frame_size(self.p, b)

    def set_quality_level(self, q):
        self.lib.rc_set_quality_level(self.p, q)

    def state(self):
        return (self.lib.ref_rc_fullness(self.p), self.lib.ref_rc_integral(self.p), self.lib.ref_rc_debt(self.p))


class OurRc:
    """The package's RateControl with the same methods."""

    def __init__(self, pkg, mode, rate, fps, w, h, model=False, current_qp=0):
        self.r = pkg.RateControl(mode, rate, fps, w, h, model=model, current_qp=current_qp)

    def frame_qp(self, sad, intra, still):
        return self.r.frame_qp(sad, intra, still)

    def coded(self, bits, qp, intra, still):
        self.r.coded(bits, qp, intra, still)

    def update_bitrate(self, r, w, h):
        self.r.update_bitrate(r, w, h)

    def set_max_frame_size(self, b):
        self.r.set_max_frame_size(b)

    def set_quality_level(self, q):
        self.r.set_quality_level(q)

    def state(self):
        return (self.r.st.buffer_fullness, self.r.st.error_integral, self.r.st.debt)


CASES = [(mode, model, seed) for mode in ("cbr", "vbr", "low_latency", "cqp") for model in (False, True)
         for seed in range(5)]
