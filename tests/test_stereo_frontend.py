"""Stereo front-end (examples/stereo/stereo.cc) pieces that run on the CPU:
the grey conversion rule and the example's cost volume."""
import os

import numpy as np

from gridinfcuda_amd import stereo as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_gray_rule_known_values(tmp_path):
    from PIL import Image
    px = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [10, 20, 30], [0, 0, 0]]], np.uint8)
    p = tmp_path / "c.png"
    Image.fromarray(px).save(p)
    g = S.load_gray(str(p))
    # BT.601 in 14-bit fixed point: 0.299, 0.587, 0.114 of 255, rounded
    np.testing.assert_array_equal(g[0], [76, 150, 29, 255, round(0.299 * 10 + 0.587 * 20 + 0.114 * 30), 0])


def test_tsukuba_fixture_loads():
    left = S.load_gray(os.path.join(GOLDEN, "tsukuba_left.png"))
    right = S.load_gray(os.path.join(GOLDEN, "tsukuba_right.png"))
    assert left.shape == right.shape == (288, 384)
    assert left.dtype == np.uint8 and 0 < left.mean() < 255


def test_example_costs_brute_force():
    rng = np.random.default_rng(0)
    left = rng.integers(0, 256, size=(5, 23)).astype(np.uint8)
    right = rng.integers(0, 256, size=(5, 23)).astype(np.uint8)
    c = S.example_costs(left, right)
    for y in range(5):
        for x in range(23):
            for k in range(16):
                want = min(abs(int(left[y, x]) - int(right[y, x - k])), 20) if x >= 15 else 0
                assert c[k, y, x] == want
