"""Writes framegen_sha256.json: SHA-256 of the synthetic frames (visible NV12
planes, luma then interleaved CbCr) that the tests and bench.py use, for the
first and last frame of each workload.  Run from the repository root after
building the package; the test compares the generator against this file."""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))

WORKLOADS = [  # name, seed, width, height, rectangles, frames
    ("640x480 IPPP", 2024, 640, 480, 6, 32),
    ("1920x1080 IPPP (bench)", 1080, 1920, 1080, 12, 128),
]


def frame_hash(g, f):
    y, uv = g.frame(f)
    h = hashlib.sha256()
    h.update(y.tobytes())
    h.update(uv.tobytes())
    return h.hexdigest()


def compute():
    from conftest import load_pkg
    pkg = load_pkg()
    out = []
    for name, seed, w, h, nrect, n in WORKLOADS:
        g = pkg.FrameGen(seed, w, h, nrect)
        out.append({"workload": name, "seed": seed, "width": w, "height": h, "rectangles": nrect,
                    "frames": {str(f): frame_hash(g, f) for f in (0, n - 1)},
                    "offset_qpel": {str(f): list(g.offset_qpel(f)) for f in (0, 1, n - 1)}})
        g.close()
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "framegen_sha256.json")
    with open(path, "w") as f:
        json.dump(compute(), f, indent=1)
    print("wrote", path)
