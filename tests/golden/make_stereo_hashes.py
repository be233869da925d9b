"""Writes tests/golden/stereo_hashes.json: SHA-256 of the int32 cost volume of
seeded generator instances, computed with the oracle's numpy restatement
(oracle/stereo_gen.py), so the fixture is independent of the product's C++
generator it pins.  Run from the repository root:
    python tests/golden/make_stereo_hashes.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.stereo_gen import gen_stereo  # noqa: E402

CASES = [
    # config-1 instance: 128x96, 16 labels, 8 rectangles, disparity 0..15, truncation 20, seed 1
    dict(seed=1, W=128, H=96, L=16, rects=8, dmax=15, trunc=20),
    # config-2 recipe on a 240x135 crop: 128 labels, 64 rectangles, disparity 0..127, truncation 40, seed 2
    dict(seed=2, W=240, H=135, L=128, rects=64, dmax=127, trunc=40),
]


def main():
    out = []
    for c in CASES:
        g = gen_stereo(c["seed"], c["W"], c["H"], c["L"], c["rects"], c["dmax"], c["trunc"])
        out.append(dict(c, sha256_cost_i32=hashlib.sha256(g["cost_i32"].tobytes()).hexdigest()))
    with open(os.path.join(HERE, "stereo_hashes.json"), "w") as f:
        json.dump({"generator": "SplitMix64 stereo generator (oracle/stereo_gen.py)", "cases": out}, f, indent=1)


if __name__ == "__main__":
    main()
