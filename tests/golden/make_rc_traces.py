"""Writes rc_traces.json: per-picture QP, bits, bucket fullness, integral
and debt of the reference's rate controller (oracle/_ref/libref_rc.so,
built from approach1-compute-encoder/src/rate_control.c) on the seeded
scenarios of tests/rc_scenarios.py, so the parity test also runs where the
reference tree is absent.  Run from the repo root after `make -C oracle ref`."""
import json
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
import rc_scenarios as rs  # noqa: E402

out = {}
for mode, model, seed in rs.CASES[::5]:
    sc = rs.scenario(seed)
    ctl = rs.RefRc(rs.MODES[mode], sc["rate"], sc["fps"], sc["w"], sc["h"], model=model, current_qp=30)
    out[f"{mode}/{int(model)}/{seed}"] = rs.run(ctl, sc, model)
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rc_traces.json"), "w") as f:
    json.dump(out, f, separators=(",", ":"))
print(len(out), "traces")
