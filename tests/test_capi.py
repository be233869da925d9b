"""CPU checks of the C ABI library: it loads, exports every symbol the header
declares, and its host-side generator matches the oracle's restatement."""
import ctypes
import hashlib
import json
import os
import re

import numpy as np
import pytest

from gridinfcuda_amd import _lib as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ginf_bp.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ginf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    lib = G.lib()
    syms = declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in ginf_bp.h but not exported"
    assert set(syms) == set(G.EXPORTS)


def test_version_string():
    assert b"gfx950" in G.lib().ginf_version()


@pytest.mark.parametrize("seed,W,H,L,R,dmax,T", [(1, 128, 96, 16, 8, 15, 20), (9, 37, 23, 8, 5, 7, 30),
                                                  (3000, 64, 40, 32, 10, 31, 40)])
def test_generator_matches_numpy_restatement(seed, W, H, L, R, dmax, T):
    from oracle.stereo_gen import gen_stereo
    ref = gen_stereo(seed, W, H, L, R, dmax, T)
    got = G.gen_stereo(seed, W, H, L, R, dmax, T, want=("left", "right", "disp", "cost_i32", "cost_f32"))
    for k in ("left", "right", "disp", "cost_i32"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(got["cost_f32"], ref["cost_i32"].astype(np.float32))


def test_generator_fixture_hashes():
    """Byte hashes of the config-1 instance (seed 1) and a config-2 crop, fixed in
    tests/golden/stereo_hashes.json (written by tests/golden/make_stereo_hashes.py)."""
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "stereo_hashes.json")))
    for case in fx["cases"]:
        got = G.gen_stereo(case["seed"], case["W"], case["H"], case["L"], case["rects"], case["dmax"],
                           case["trunc"], want=("cost_i32",))
        assert hashlib.sha256(got["cost_i32"].tobytes()).hexdigest() == case["sha256_cost_i32"]


def test_generator_properties():
    g = G.gen_stereo(5, 200, 100, 64, 16, 63, 40, want=("left", "right", "disp", "cost_i32"))
    c = g["cost_i32"]
    assert c.min() >= 0 and c.max() <= 40
    for l in range(64):
        assert (c[l, :, :l] == 40).all()
    # where the warp is unambiguous, the true disparity costs 0
    d = g["disp"]
    ys, xs = np.nonzero(d >= 0)
    zero_hits = c[d[ys, xs], ys, xs] == 0
    assert zero_hits.mean() > 0.5


def test_rejects_invalid_sizes():
    assert G.lib().ginf_gen_stereo(ctypes.c_uint64(1), 0, 5, 4, 1, 3, 10, None, None, None, None, None) == 1
