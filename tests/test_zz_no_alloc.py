"""ginf_energy_f64_noalloc (the energy entry point that allocates nothing)
against ginf_energy_f64 through the C ABI: equal FP64 energies on two calls in
a row, the second on a smaller grid, for float and int costs, plus the
out-of-range label error and a grid without interior nodes."""
import numpy as np
import pytest

from gpu_util import torch_cuda, say
from gridinfcuda_amd import _lib as G

pytestmark = pytest.mark.gpu


def _case(torch, seed, W, H, L, dtype):
    rng = np.random.default_rng(seed)
    data = rng.integers(0, 40, size=(L, H, W)).astype(dtype)
    if dtype == np.float32:
        data = data + np.float32(0.125)
    a = np.arange(L)
    table = np.minimum(5 * np.abs(a[:, None] - a[None, :]), 17).astype(dtype)
    labels = rng.integers(0, L, size=(H, W)).astype(np.int32)
    return [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (data, table, labels)], data, table, labels


@pytest.mark.parametrize("dtype", [np.float32, np.int32])
def test_noalloc_energy_equals_energy(oracle_mod, dtype):
    torch = torch_cuda()
    s = torch.cuda.current_stream().cuda_stream
    code = G.dtype_code(dtype)
    for k, (W, H, L) in enumerate([(1920, 1080, 16), (301, 97, 40)]):  # second call: a smaller grid
        say(f"no-alloc energy {np.dtype(dtype).name} call {k + 1}: {W}x{H}x{L}")
        (dd, td, ld), data, table, labels = _case(torch, 40 + k, W, H, L, dtype)
        e_ref = G.energy_f64(W, H, L, code, dd.data_ptr(), td.data_ptr(), ld.data_ptr(), s)
        e_new = G.energy_f64(W, H, L, code, dd.data_ptr(), td.data_ptr(), ld.data_ptr(), s, noalloc=True)
        say(f"  ginf_energy_f64 {e_ref!r}  ginf_energy_f64_noalloc {e_new!r}")
        assert e_new == e_ref
        if k == 1:
            assert e_new == oracle_mod.energy_f64(data, table, labels)


def test_noalloc_energy_errors_and_degenerate_grids():
    torch = torch_cuda()
    s = torch.cuda.current_stream().cuda_stream
    (dd, td, ld), _, _, _ = _case(torch, 7, 33, 21, 9, np.float32)
    ld[5, 7] = 9  # label outside [0, L)
    with pytest.raises(G.GinfError) as e:
        G.energy_f64(33, 21, 9, G.F32, dd.data_ptr(), td.data_ptr(), ld.data_ptr(), s, noalloc=True)
    assert e.value.code == 1
    (dd, td, ld), _, _, _ = _case(torch, 8, 2, 2, 3, np.int32)
    assert G.energy_f64(2, 2, 3, G.I32, dd.data_ptr(), td.data_ptr(), ld.data_ptr(), s, noalloc=True) == 0.0
    (dd, td, ld), _, _, _ = _case(torch, 9, 50, 40, 6, np.int32)  # still correct after the error
    assert (G.energy_f64(50, 40, 6, G.I32, dd.data_ptr(), td.data_ptr(), ld.data_ptr(), s, noalloc=True)
            == G.energy_f64(50, 40, 6, G.I32, dd.data_ptr(), td.data_ptr(), ld.data_ptr(), s))
