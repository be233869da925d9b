"""The C++ drop-in API, used the way the reference's stereo example uses it
(examples/stereo/stereo.cc:25-43), gives the oracle's labels and energy."""
import os
import subprocess

import numpy as np
import pytest

from gpu_util import say
from gridinfcuda_amd import _lib as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gridinfcuda_amd", "bin", "dropin_stereo")


@pytest.mark.parametrize("W,H,iters,levels", [(96, 64, 30, 1), (131, 77, 6, 3)])
def test_cpp_dropin_matches_oracle(oracle_mod, W, H, iters, levels):
    say(f"C++ drop-in {W}x{H} iters={iters} levels={levels}")
    assert os.path.exists(EXE), "build the library first (python -m gridinfcuda_amd.build)"
    out = subprocess.run([EXE, str(W), str(H), str(iters), str(levels)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().split("\n")
    labels = np.array([[int(v) for v in ln.split()] for ln in lines[:-1]], np.int32)
    energy = int(lines[-1].split()[1])
    g = G.gen_stereo(1, W, H, 16, 8, 15, 20, want=("left", "right"))
    lft, rgt = g["left"].astype(np.int32), g["right"].astype(np.int32)
    data = np.zeros((16, H, W), np.int32)       # Grid<int> starts zeroed
    for k in range(16):
        data[k, :, 15:] = np.minimum(np.abs(lft[:, 15:] - rgt[:, 15 - k:W - k]), 20)
    table = oracle_mod.table_trunc_linear(16, 10, 20, np.int32)
    ref = oracle_mod.hbp(data, table, 2, levels, iters, norm=0)
    np.testing.assert_array_equal(labels, ref)
    assert energy == oracle_mod.energy(data, table, ref)


@pytest.mark.parametrize("refine", ["ab", "async"])
def test_cpp_dropin_refinement(oracle_mod, refine):
    """BP through the C++ API, then alpha-beta swap or asynchronous ICM
    (include/graphcut.h, include/icm.h) from the BP labels."""
    W, H, it = 48, 32, 2
    say(f"C++ drop-in BP + {refine}")
    out = subprocess.run([EXE, str(W), str(H), "10", "1", str(it), refine], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().split("\n")
    labels = np.array([[int(v) for v in ln.split()] for ln in lines[:-1]], np.int32)
    e_bp, e_ref = (int(v) for v in lines[-1].split()[1:3])
    g = G.gen_stereo(1, W, H, 16, 8, 15, 20, want=("left", "right"))
    lft, rgt = g["left"].astype(np.int32), g["right"].astype(np.int32)
    data = np.zeros((16, H, W), np.int32)
    for k in range(16):
        data[k, :, 15:] = np.minimum(np.abs(lft[:, 15:] - rgt[:, 15 - k:W - k]), 20)
    table = oracle_mod.table_trunc_linear(16, 10, 20, np.int32)
    bp = oracle_mod.hbp(data, table, 2, 1, 10, norm=0)
    want = oracle_mod.alpha_beta(data, table, bp, it)[0] if refine == "ab" else oracle_mod.icm(data, table, bp, it,
                                                                                                 sync=False)
    np.testing.assert_array_equal(labels, want)
    assert e_ref == oracle_mod.energy(data, table, want) <= e_bp
