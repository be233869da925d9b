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


MODELS_EXE = os.path.join(ROOT, "gridinfcuda_amd", "bin", "dropin_models")
SM = dict(explicit=0, potts=1, tl=2, tq=3)


@pytest.mark.parametrize("T,model,W,H,L,levels,iters,entry", [
    ("int", "potts", 41, 29, 12, 3, 6, "hbp"),
    ("float", "tl", 37, 23, 20, 1, 9, "bp"),
    ("float", "potts", 33, 35, 8, 2, 5, "gpuhbp"),
    ("int", "tq", 30, 22, 16, 2, 5, "hbp"),
    ("float", "tq", 26, 19, 24, 1, 7, "gpubp"),
    ("int", "explicit", 28, 20, 10, 1, 6, "bp"),
    ("float", "explicit", 31, 17, 33, 3, 4, "hbp"),
])
def test_cpp_dropin_models(oracle_mod, tmp_path, T, model, W, H, L, levels, iters, entry):
    """Grid<int> and Grid<float> through the C++ templates with every
    smoothness setter (setSmoothnessCost selects the explicit table and the
    O(L^2) kernel; useSmoothnessTruncQuad the same kernel with the quadratic
    table), decoded twice with releaseBpDeviceState() in between: labels and
    getLabelingCost equal the oracle's (float costs with a fractional part
    included)."""
    say(f"C++ drop-in models: Grid<{T}> {model} {W}x{H}x{L} levels={levels} iters={iters} {entry}")
    assert os.path.exists(MODELS_EXE), "build the library first (python -m gridinfcuda_amd.build)"
    dtype = np.int32 if T == "int" else np.float32
    rng = np.random.default_rng(W * 1000 + H * 10 + L)
    data = rng.integers(0, 40, size=(L, H, W)).astype(dtype)
    if T == "float":
        data = (data + rng.integers(0, 4, size=data.shape) * np.float32(0.25)).astype(np.float32)
    a = np.arange(L)
    if model == "potts":
        table = (7 * (a[:, None] != a[None, :])).astype(dtype)
    elif model == "tl":
        table = np.minimum(3 * np.abs(a[:, None] - a[None, :]), 12).astype(dtype)
    elif model == "tq":
        table = np.minimum(2 * (a[:, None] - a[None, :]) ** 2, 30).astype(dtype)
    else:
        table = rng.integers(0, 20, size=(L, L)).astype(dtype)  # asymmetric
    cpath, tpath = tmp_path / "costs.bin", tmp_path / "table.bin"
    data.tofile(cpath)
    # the oracle's table is V(a, b) at a + L*b, i.e. table[b, a] in numpy
    np.ascontiguousarray(table).tofile(tpath)
    out = subprocess.run([MODELS_EXE, T, model, str(W), str(H), str(L), str(levels), str(iters), str(cpath),
                          str(tpath), entry], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().split("\n")
    labels = np.array([[int(v) for v in ln.split()] for ln in lines[:-1]], np.int32)
    lv = 1 if entry in ("bp", "gpubp") else levels
    ref = oracle_mod.hbp(data, table, SM[model], lv, iters, norm=0)
    np.testing.assert_array_equal(labels, ref)
    e = lines[-1].split()[1]
    want = oracle_mod.energy(data, table, ref)
    if T == "int":
        assert int(e) == want
    else:
        assert np.float32(float(e)) == np.float32(want)
