"""Real-image end-to-end check: the reference's stereo example pipeline on
the Tsukuba pair it ships (examples/stereo/left.png, right.png, kept as
fixtures), through the Python front-end and through the C++ drop-in API on
the same grey images, against the oracle; then ICM refinement."""
import os
import subprocess

import numpy as np
import pytest

from gpu_util import say
from gridinfcuda_amd import _lib as G
from gridinfcuda_amd import stereo as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "gridinfcuda_amd", "bin", "dropin_stereo")


@pytest.fixture(scope="module")
def tsukuba(oracle_mod):
    left = S.load_gray(os.path.join(GOLDEN, "tsukuba_left.png"))
    right = S.load_gray(os.path.join(GOLDEN, "tsukuba_right.png"))
    costs = S.example_costs(left, right)
    table = S.smoothness_table()
    ref = oracle_mod.bp(costs, table, G.SM_TRUNC_LINEAR, 30)
    return left, right, costs, table, ref


def write_pgm(path, img):
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n255\n" % (img.shape[1], img.shape[0]))
        f.write(np.ascontiguousarray(img, np.uint8).tobytes())


def test_python_frontend_matches_oracle(oracle_mod, tsukuba):
    left, right, costs, table, ref = tsukuba
    say("tsukuba python front-end")
    lab = S.disparity(left, right)
    np.testing.assert_array_equal(lab, ref)
    e_bp = oracle_mod.energy(costs, table, lab)
    e_prior = oracle_mod.energy(costs, table, oracle_mod.max_prior(costs))
    say(f"  energy BP {e_bp}, max-prior {e_prior}")
    assert e_bp < e_prior


@pytest.mark.parametrize("icm_mode", ["sync", "async"])
def test_cpp_frontend_matches_oracle(oracle_mod, tsukuba, tmp_path, icm_mode):
    left, right, costs, table, ref = tsukuba
    say(f"tsukuba C++ front-end + ICM {icm_mode}")
    lp, rp = str(tmp_path / "l.pgm"), str(tmp_path / "r.pgm")
    write_pgm(lp, left)
    write_pgm(rp, right)
    out = subprocess.run([EXE, "--pgm", lp, rp, "30", "1", "5", icm_mode], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().split("\n")
    labels = np.array([[int(v) for v in ln.split()] for ln in lines[:-1]], np.int32)
    e_bp, e_icm, e_prior = (int(v) for v in lines[-1].split()[1:])
    want = oracle_mod.icm(costs, table, ref, 5, sync=(icm_mode == "sync"))
    np.testing.assert_array_equal(labels, want)
    assert e_bp == oracle_mod.energy(costs, table, ref)
    assert e_icm == oracle_mod.energy(costs, table, want)
    assert e_prior == oracle_mod.energy(costs, table, oracle_mod.max_prior(costs))
    if icm_mode == "async":
        assert e_icm <= e_bp
