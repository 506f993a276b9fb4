"""The host rate controller (host/bc250_rc.c) against the reference's own
rate_control.c, built from the reference tree into oracle/_ref (every
picture's QP and the bucket, integral and debt after it, on seeded
scenarios in every mode, buffer loop and model), against traces of that
build committed in tests/golden/rc_traces.json, and against what the
formulas fix on their own: the bits-per-pixel anchor of the starting QP and
the long-run rate of a stream whose cost curve the model knows."""
import json
import os

import numpy as np
import pytest

import rc_scenarios as rs
from conftest import load_pkg


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


@pytest.fixture(autouse=True)
def _no_qp_min_env(monkeypatch):
    monkeypatch.delenv("BC250_QP_MIN", raising=False)


def _ours(pkg, mode, sc, model):
    return rs.OurRc(pkg, rs.MODES[mode], sc["rate"], sc["fps"], sc["w"], sc["h"], model=model, current_qp=30)


@pytest.mark.skipif(not os.path.exists(rs.RefRc.LIB), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("mode,model,seed", rs.CASES)
def test_matches_reference_build(pkg, mode, model, seed):
    sc = rs.scenario(seed)
    want = rs.run(rs.RefRc(rs.MODES[mode], sc["rate"], sc["fps"], sc["w"], sc["h"], model=model, current_qp=30), sc,
                  model)
    got = rs.run(_ours(pkg, mode, sc, model), sc, model)
    for i, (a, b) in enumerate(zip(got, want)):
        assert a == b, f"picture {i}: ours {a}, reference {b}"
    qps = [t[0] for t in got]
    if mode != "cqp":
        assert len(set(qps)) > 3   # the scenario moves the QP


def test_matches_committed_reference_traces(pkg):
    with open(os.path.join(os.path.dirname(__file__), "golden", "rc_traces.json")) as f:
        traces = json.load(f)
    assert len(traces) == 8
    for key, want in traces.items():
        mode, model, seed = key.split("/")
        sc = rs.scenario(int(seed))
        got = rs.run(_ours(pkg, mode, sc, bool(int(model))), sc, bool(int(model)))
        assert got == want, key


def test_starting_qp_anchor(pkg, monkeypatch):
    """26 at 1 Mbit/s of 1280x720 at 30 fps, 6 QP per doubling of the bits
    per pixel, clamped to [BC250_QP_MIN (12), 51]."""
    for rate, want in ((1_000_000, 26), (2_000_000, 20), (4_000_000, 14), (500_000, 32), (250_000, 38),
                       (1_000, 51), (400_000_000, 12)):
        r = pkg.RateControl(pkg.RC_CBR, rate, 30.0, 1280, 720)
        assert (r.st.base_qp, r.st.current_qp) == (want, want), rate
    # the same bits per pixel at another shape and rate
    assert pkg.RateControl(pkg.RC_CBR, 4_000_000, 60.0, 1920, 1080).st.base_qp == 26 - round(
        6 * np.log2(4e6 / (1920 * 1080 * 60) / 0.036169))
    monkeypatch.setenv("BC250_QP_MIN", "5")
    assert pkg.RateControl(pkg.RC_CBR, 400_000_000, 30.0, 1280, 720).st.base_qp == 5
    # buckets: a second, or two pictures' budget in low latency
    r = pkg.RateControl(pkg.RC_LOW_LATENCY, 6_000_000, 60.0, 1280, 720)
    assert (r.st.target_bits_per_frame, r.st.buffer_size, r.st.buffer_fullness) == (100000, 200000, 100000)
    r = pkg.RateControl(pkg.RC_CBR, 6_000_000, 60.0, 1280, 720)
    assert (r.st.buffer_size, r.st.buffer_fullness) == (6_000_000, 3_000_000)
    # constant QP keeps the caller's QP and ignores the bits
    r = pkg.RateControl(pkg.RC_CQP, 6_000_000, 60.0, 1280, 720, current_qp=33)
    assert r.frame_qp(0) == 33
    r.coded(10**7, 33)
    assert r.frame_qp(0) == 33 and r.st.buffer_fullness == 3_000_000


@pytest.mark.parametrize("mode", ["cbr", "low_latency"])
def test_model_reaches_the_rate(pkg, mode):
    """A stream that costs exactly c * 2^(-(qp - 12) / 5) per P picture is
    coded within one QP step of the QP that spends the bitrate, so the mean
    over the last 200 of 300 pictures is within a factor 2^(1/5) of one
    picture's share (QP is whole, and a QP is held until the model moves a
    whole step away from it)."""
    w, h, fps, rate = 1920, 1080, 60.0, 20_000_000
    c = w * h * 3.0
    r = pkg.RateControl(rs.MODES[mode], rate, fps, w, h, model=True)
    spent = []
    for i in range(300):
        intra = i == 0
        qp = r.frame_qp(0, intra, False)
        bits = int(c * 2.0 ** (-(qp - 12) / 5.0) * (4.0 if intra else 1.0))
        r.coded(bits, qp, intra, False)
        spent.append(bits)
    share = rate / fps
    assert 2 ** -0.2 <= np.mean(spent[100:]) / share <= 2 ** 0.2
    # the QP the cost curve gives for one share, within the rounding
    assert abs(qp - (12 + 5 * np.log2(c / share))) <= 1.0


def test_buffer_loop_holds_the_rate(pkg):
    """The buffer loop on a stream of fixed cost curve converges to the
    bitrate: over the last 300 of 600 pictures the mean is within 10 % of
    one share (QP moves in whole steps, so it dithers around the rate)."""
    w, h, fps, rate = 1280, 720, 60.0, 8_000_000
    c = w * h * 2.0
    for mode in (pkg.RC_CBR, pkg.RC_LOW_LATENCY):
        r = pkg.RateControl(mode, rate, fps, w, h)
        spent = []
        for i in range(600):
            qp = r.frame_qp(0)
            bits = int(c * 2.0 ** (-(qp - 12) / 6.0))
            r.coded(bits, qp)
            spent.append(bits)
        assert abs(np.mean(spent[300:]) / (rate / fps) - 1) < 0.10, mode
