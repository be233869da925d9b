"""ctypes binding of libginf_bp.so (the C ABI declared in include/ginf_bp.h).

The product path is the HIP library: if it is missing this module raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# GINF_LIB selects another build of the same library (e.g. an experimental variant).
LIB_PATH = os.environ.get("GINF_LIB") or os.path.join(PKG, "libginf_bp.so")

F32, I32 = 0, 1
SM_EXPLICIT, SM_POTTS, SM_TRUNC_LINEAR, SM_TRUNC_QUAD = 0, 1, 2, 3
NORM_MEAN, NORM_MIN = 0, 1
KERNEL_AUTO, KERNEL_CHUNKED, KERNEL_QUAD = 0, 1, 2

# Every symbol include/ginf_bp.h declares.
EXPORTS = [
    "ginf_bp_plan_create", "ginf_bp_plan_run", "ginf_bp_plan_destroy", "ginf_bp_plan_workspace_bytes",
    "ginf_hbp_host", "ginf_energy_f64", "ginf_energy_f64_noalloc", "ginf_bp_plan_dump_msgs", "ginf_bp_plan_level_dims",
    "ginf_gen_stereo", "ginf_last_error", "ginf_version", "ginf_bp_plan_set_sweep_timing",
    "ginf_bp_plan_sweep_stats", "ginf_bp_plan_sweep_level_updates", "ginf_bp_plan_updates_per_run", "ginf_bp_plan_dump_data",
    "ginf_batch_create", "ginf_batch_run", "ginf_batch_set_tiles", "ginf_batch_updates_per_run",
    "ginf_batch_dump_msgs", "ginf_batch_destroy", "ginf_bp_plan_run_schedule", "ginf_bp_plan_seed",
    "ginf_bp_plan_block_mean_msgs", "ginf_conv3x3_fwd", "ginf_conv3x3_bwd_weight", "ginf_conv3x3_flip_weights",
    "ginf_relu_bwd", "ginf_mse", "ginf_adam", "ginf_to_bf16", "ginf_minpool_costs", "ginf_block_mean_msgs",
    "ginf_stereo_costs", "ginf_batch_set_seeded", "ginf_batch_seed", "ginf_hbp_host_release",
    "ginf_icm", "ginf_max_prior", "ginf_icm_host", "ginf_max_prior_host", "ginf_alpha_beta",
    "ginf_alpha_beta_host", "ginf_stream_probe", "ginf_bp_plan_set_compact_data", "ginf_bp_plan_compact_data",
    "ginf_batch_set_compact_data", "ginf_bp_plan_set_pairs", "ginf_bp_plan_pairs", "ginf_convnet_fwd",
    "ginf_bp_plan_run_predicted", "ginf_bp_plan_prepare", "ginf_batch_set_predictions",
]


class GinfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ginf error {code}: {msg}")
        self.code = code


class BpDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("num_labels", ctypes.c_int),
        ("dtype", ctypes.c_int), ("sm_model", ctypes.c_int), ("num_levels", ctypes.c_int),
        ("iters_per_level", ctypes.c_int), ("norm", ctypes.c_int), ("lanes_per_node", ctypes.c_int),
        ("kernel", ctypes.c_int), ("reserved", ctypes.c_int * 6),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m gridinfcuda_amd.build` "
                              "(the HIP library is the only implementation)")
        # PyTorch ships its own libamdhip64 (soname libamdhip64.so.7, but its
        # libraries link it as "libamdhip64.so").  Loading torch first makes
        # this library bind to that same runtime; loading ours first would put
        # two HIP runtimes in the process and one of them sees no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        P, I, VP = ctypes.POINTER, ctypes.c_int, ctypes.c_void_p
        L.ginf_bp_plan_create.argtypes = [P(BpDesc), VP, P(VP)]
        L.ginf_bp_plan_run.argtypes = [VP, VP, VP, VP]
        L.ginf_bp_plan_destroy.argtypes = [VP]
        L.ginf_bp_plan_workspace_bytes.argtypes = [VP, P(ctypes.c_size_t)]
        L.ginf_hbp_host.argtypes = [P(BpDesc), VP, VP, VP]
        L.ginf_energy_f64.argtypes = [I, I, I, I, VP, VP, VP, P(ctypes.c_double), VP]
        L.ginf_energy_f64_noalloc.argtypes = [I, I, I, I, VP, VP, VP, P(ctypes.c_double), VP]
        L.ginf_bp_plan_dump_msgs.argtypes = [VP, I, VP]
        L.ginf_bp_plan_dump_data.argtypes = [VP, I, VP]
        L.ginf_bp_plan_level_dims.argtypes = [VP, I, P(I), P(I)]
        L.ginf_gen_stereo.argtypes = [ctypes.c_uint64, I, I, I, I, I, I, VP, VP, VP, VP, VP]
        L.ginf_bp_plan_set_sweep_timing.argtypes = [VP, I]
        L.ginf_bp_plan_sweep_stats.argtypes = [VP, P(ctypes.c_double), P(ctypes.c_int64), P(ctypes.c_double), I]
        L.ginf_bp_plan_sweep_level_updates.argtypes = [VP, P(ctypes.c_double)]
        L.ginf_bp_plan_updates_per_run.argtypes = [VP, P(ctypes.c_double)]
        I64, F = ctypes.c_int64, ctypes.c_float
        L.ginf_bp_plan_run_schedule.argtypes = [VP, VP, VP, I, P(I), I, VP]
        L.ginf_bp_plan_seed.argtypes = [VP, I, VP, I, I, I, F, VP]
        L.ginf_bp_plan_run_predicted.argtypes = [VP, VP, VP, I, P(I), VP, I, I, I, F, I, VP]
        L.ginf_bp_plan_prepare.argtypes = [VP, VP, I, VP, F, VP]
        L.ginf_bp_plan_block_mean_msgs.argtypes = [VP, I, I, F, VP, VP]
        L.ginf_conv3x3_fwd.argtypes = [VP, VP, VP, I, I, I, I, I, I, VP, VP, VP]
        L.ginf_convnet_fwd.argtypes = [I, VP, VP, VP, VP, I, I, I, VP, VP, VP, VP]
        L.ginf_conv3x3_bwd_weight.argtypes = [VP, VP, I, I, I, I, I, VP, VP, VP]
        L.ginf_conv3x3_flip_weights.argtypes = [VP, I, I, VP, VP]
        L.ginf_relu_bwd.argtypes = [VP, VP, I64, VP, VP]
        L.ginf_mse.argtypes = [VP, VP, I64, VP, VP, P(I), VP]
        L.ginf_adam.argtypes = [VP, VP, VP, VP, VP, I64, F, F, F, F, I, F, VP]
        L.ginf_to_bf16.argtypes = [VP, I64, VP, VP]
        L.ginf_minpool_costs.argtypes = [VP, I, I, I, I, I, F, VP, VP]
        L.ginf_block_mean_msgs.argtypes = [VP, I, I, I, I, F, VP, VP]
        L.ginf_stereo_costs.argtypes = [VP, VP, I, I, I, I, VP, VP]
        L.ginf_stream_probe.argtypes = [I, I, VP, VP, I64, I, VP]
        L.ginf_bp_plan_set_compact_data.argtypes = [VP, I]
        L.ginf_bp_plan_set_pairs.argtypes = [VP, I]
        L.ginf_bp_plan_pairs.argtypes = [VP, P(I)]
        L.ginf_bp_plan_compact_data.argtypes = [VP, P(I)]
        L.ginf_last_error.restype = ctypes.c_char_p
        L.ginf_version.restype = ctypes.c_char_p
        for name in EXPORTS:
            if name not in ("ginf_last_error", "ginf_version"):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise GinfError(rc, lib().ginf_last_error().decode())


def make_desc(width, height, num_labels, dtype=F32, sm_model=SM_TRUNC_LINEAR, num_levels=1, iters=10,
              norm=NORM_MEAN, lanes_per_node=0, kernel=KERNEL_AUTO) -> BpDesc:
    d = BpDesc()
    d.kernel = kernel
    d.width, d.height, d.num_labels = width, height, num_labels
    d.dtype, d.sm_model, d.num_levels = dtype, sm_model, num_levels
    d.iters_per_level, d.norm, d.lanes_per_node = iters, norm, lanes_per_node
    return d


def dtype_code(np_dtype) -> int:
    dt = np.dtype(np_dtype)
    if dt == np.float32:
        return F32
    if dt == np.int32:
        return I32
    raise TypeError(f"unsupported element type {dt}")


class Plan:
    """Device workspace for one problem shape (ginf_bp_plan)."""

    def __init__(self, desc: BpDesc, table: np.ndarray):
        table = np.ascontiguousarray(table)
        assert table.shape == (desc.num_labels, desc.num_labels)
        assert dtype_code(table.dtype) == desc.dtype
        self.desc = desc
        self.table = table
        self._h = ctypes.c_void_p()
        check(lib().ginf_bp_plan_create(ctypes.byref(desc), table.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.byref(self._h)))

    def run(self, data_ptr: int, labels_ptr: int, stream: int = 0) -> None:
        check(lib().ginf_bp_plan_run(self._h, ctypes.c_void_p(data_ptr), ctypes.c_void_p(labels_ptr),
                                     ctypes.c_void_p(stream)))

    def level_dims(self, level: int):
        w, h = ctypes.c_int(), ctypes.c_int()
        check(lib().ginf_bp_plan_level_dims(self._h, level, ctypes.byref(w), ctypes.byref(h)))
        return w.value, h.value

    def dump_msgs(self, level: int = 0) -> np.ndarray:
        w, h = self.level_dims(level)
        L = self.desc.num_labels
        out = np.zeros((L, 4, h, w), np.float32 if self.desc.dtype == F32 else np.int32)
        check(lib().ginf_bp_plan_dump_msgs(self._h, level, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def dump_data(self, level: int = 0) -> np.ndarray:
        """Data costs of a pyramid level as (L, h, w)."""
        w, h = self.level_dims(level)
        L = self.desc.num_labels
        out = np.zeros((L, h, w), np.float32 if self.desc.dtype == F32 else np.int32)
        check(lib().ginf_bp_plan_dump_data(self._h, level, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def run_schedule(self, data_ptr: int, labels_ptr: int, start_level: int, iters, seeded: bool,
                     stream: int = 0) -> None:
        arr = (ctypes.c_int * self.desc.num_levels)(*([int(v) for v in iters] + [0] * (self.desc.num_levels - len(iters))))
        check(lib().ginf_bp_plan_run_schedule(self._h, ctypes.c_void_p(data_ptr), ctypes.c_void_p(labels_ptr),
                                              start_level, arr, 1 if seeded else 0, ctypes.c_void_p(stream)))

    def seed(self, level: int, pred_ptr: int, pw: int, ph: int, shift: int, scale: float, stream: int = 0) -> None:
        check(lib().ginf_bp_plan_seed(self._h, level, ctypes.c_void_p(pred_ptr), pw, ph, shift, scale,
                                      ctypes.c_void_p(stream)))

    def run_predicted(self, data_ptr: int, labels_ptr: int, start_level: int, iters, pred_ptr: int, pw: int,
                      ph: int, shift: int, scale: float, stream: int = 0, prepared: bool = False) -> None:
        """seed(start_level, pred, ...) + run_schedule(..., seeded=True) in one
        call, the first half-sweep reading the prediction directly; with
        prepared, the level data of the last prepare() is used."""
        arr = (ctypes.c_int * self.desc.num_levels)(*([int(v) for v in iters] + [0] * (self.desc.num_levels - len(iters))))
        check(lib().ginf_bp_plan_run_predicted(self._h, ctypes.c_void_p(data_ptr), ctypes.c_void_p(labels_ptr),
                                               start_level, arr, ctypes.c_void_p(pred_ptr), pw, ph, shift, scale,
                                               1 if prepared else 0, ctypes.c_void_p(stream)))

    def prepare(self, data_ptr: int, start_level: int, minpool_ptr: int = 0, minpool_scale: float = 0.0,
                stream: int = 0) -> None:
        """Level data of levels 0..start_level (and, with minpool_ptr, the
        min-pooled network input from the same pass)."""
        check(lib().ginf_bp_plan_prepare(self._h, ctypes.c_void_p(data_ptr), start_level,
                                         ctypes.c_void_p(minpool_ptr) if minpool_ptr else None, minpool_scale,
                                         ctypes.c_void_p(stream)))

    def block_mean_msgs(self, level: int, f: int, scale: float, out_ptr: int, stream: int = 0) -> None:
        check(lib().ginf_bp_plan_block_mean_msgs(self._h, level, f, scale, ctypes.c_void_p(out_ptr),
                                                 ctypes.c_void_p(stream)))

    def set_compact_data(self, enable) -> None:
        """True/1: the narrowest storage that fits (8-bit at levels 0-1, else
        16-bit, else T); 2: at most 16-bit; False/0: always T."""
        check(lib().ginf_bp_plan_set_compact_data(self._h, int(enable)))

    def set_pairs(self, enable) -> None:
        """Fused pairs of half-sweeps (ginf_bp_plan_set_pairs): True, False, or
        None for the default (on for label capacities 16 and 32)."""
        check(lib().ginf_bp_plan_set_pairs(self._h, -1 if enable is None else (1 if enable else 0)))

    def pairs(self) -> bool:
        v = ctypes.c_int()
        check(lib().ginf_bp_plan_pairs(self._h, ctypes.byref(v)))
        return bool(v.value)

    def compact_data(self) -> bool:
        """Whether the last run stored the data costs as 16- or 8-bit integers."""
        return self.data_form() != 0

    def data_form(self) -> int:
        """Storage of the last run's data costs: 0 T, 1 16-bit, 2 8-bit at
        levels 0 and 1 (16-bit above)."""
        v = ctypes.c_int()
        check(lib().ginf_bp_plan_compact_data(self._h, ctypes.byref(v)))
        return v.value

    def set_sweep_timing(self, enable: bool) -> None:
        check(lib().ginf_bp_plan_set_sweep_timing(self._h, 1 if enable else 0))

    def sweep_stats(self, reset: bool = True):
        """(total kernel ms, launches, node-label updates) of the timed half-sweeps."""
        ms, n, up = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        check(lib().ginf_bp_plan_sweep_stats(self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(up),
                                             1 if reset else 0))
        return ms.value, n.value, up.value

    def sweep_level_updates(self):
        """Node-label updates of the timed half-sweeps per level (before a
        resetting sweep_stats)."""
        n = max(32, self.desc.num_levels)
        buf = (ctypes.c_double * n)()
        check(lib().ginf_bp_plan_sweep_level_updates(self._h, buf))
        return [buf[t] for t in range(self.desc.num_levels)]

    def updates_per_run(self) -> float:
        u = ctypes.c_double()
        check(lib().ginf_bp_plan_updates_per_run(self._h, ctypes.byref(u)))
        return u.value

    def workspace_bytes(self) -> int:
        n = ctypes.c_size_t()
        check(lib().ginf_bp_plan_workspace_bytes(self._h, ctypes.byref(n)))
        return n.value

    def close(self) -> None:
        if self._h:
            lib().ginf_bp_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hbp_host(data: np.ndarray, table: np.ndarray, sm_model: int, levels: int, iters: int, norm: int = NORM_MEAN,
             labels_init: np.ndarray | None = None, lanes_per_node: int = 0) -> np.ndarray:
    """Synchronous host-pointer HBP (ginf_hbp_host); data is (L, H, W)."""
    data = np.ascontiguousarray(data)
    L, H, W = data.shape
    table = np.ascontiguousarray(table, dtype=data.dtype)
    labels = (np.zeros((H, W), np.int32) if labels_init is None
              else np.ascontiguousarray(labels_init, dtype=np.int32).copy())
    d = make_desc(W, H, L, dtype_code(data.dtype), sm_model, levels, iters, norm, lanes_per_node)
    check(lib().ginf_hbp_host(ctypes.byref(d), data.ctypes.data_as(ctypes.c_void_p),
                              table.ctypes.data_as(ctypes.c_void_p), labels.ctypes.data_as(ctypes.c_void_p)))
    return labels


ICM_SYNC, ICM_ASYNC = 0, 1


def icm(width, height, num_labels, dtype, data_ptr, table_ptr, initial_ptr, iters, mode, result_ptr,
        stream=0) -> None:
    """ICM on device pointers (ginf_icm)."""
    vp = ctypes.c_void_p
    check(lib().ginf_icm(width, height, num_labels, dtype, vp(data_ptr), vp(table_ptr), vp(initial_ptr), iters, mode,
                         vp(result_ptr), vp(stream)))


def icm_host(data: np.ndarray, table: np.ndarray, initial: np.ndarray, iters: int, mode: int = ICM_SYNC) -> np.ndarray:
    """Synchronous host-array ICM (ginf_icm_host); data is (L, H, W)."""
    data = np.ascontiguousarray(data)
    L, H, W = data.shape
    table = np.ascontiguousarray(table, dtype=data.dtype)
    init = np.ascontiguousarray(initial, dtype=np.int32)
    out = np.zeros((H, W), np.int32)
    vp = ctypes.c_void_p
    check(lib().ginf_icm_host(W, H, L, dtype_code(data.dtype), vp(data.ctypes.data), vp(table.ctypes.data),
                              vp(init.ctypes.data), iters, mode, vp(out.ctypes.data)))
    return out


def alpha_beta(width, height, num_labels, data_ptr, table_ptr, initial_ptr, iters, result_ptr, stream=0) -> int:
    """Alpha-beta swap on device pointers (ginf_alpha_beta, int costs); returns the kept moves."""
    vp = ctypes.c_void_p
    moves = ctypes.c_int()
    check(lib().ginf_alpha_beta(width, height, num_labels, I32, vp(data_ptr), vp(table_ptr), vp(initial_ptr), iters,
                                vp(result_ptr), ctypes.byref(moves), vp(stream)))
    return moves.value


def alpha_beta_host(data: np.ndarray, table: np.ndarray, initial: np.ndarray, iters: int) -> np.ndarray:
    """Synchronous host-array alpha-beta swap (ginf_alpha_beta_host); data is (L, H, W) int32."""
    data = np.ascontiguousarray(data)
    L, H, W = data.shape
    table = np.ascontiguousarray(table, dtype=data.dtype)
    init = np.ascontiguousarray(initial, dtype=np.int32)
    out = np.zeros((H, W), np.int32)
    vp = ctypes.c_void_p
    check(lib().ginf_alpha_beta_host(W, H, L, dtype_code(data.dtype), vp(data.ctypes.data), vp(table.ctypes.data),
                                     vp(init.ctypes.data), iters, vp(out.ctypes.data)))
    return out


def max_prior_host(data: np.ndarray) -> np.ndarray:
    """decodeMaxPrior on the device (ginf_max_prior_host); data is (L, H, W)."""
    data = np.ascontiguousarray(data)
    L, H, W = data.shape
    out = np.zeros((H, W), np.int32)
    check(lib().ginf_max_prior_host(W, H, L, dtype_code(data.dtype), ctypes.c_void_p(data.ctypes.data),
                                    ctypes.c_void_p(out.ctypes.data)))
    return out


def energy_f64(width, height, num_labels, dtype, data_ptr, table_ptr, labels_ptr, stream=0,
               noalloc: bool = False) -> float:
    """Labelling energy in FP64 (ginf_energy_f64, or with noalloc the entry
    point that keeps its partial sums in the library's fixed device array)."""
    out = ctypes.c_double()
    fn = lib().ginf_energy_f64_noalloc if noalloc else lib().ginf_energy_f64
    check(fn(width, height, num_labels, dtype, ctypes.c_void_p(data_ptr), ctypes.c_void_p(table_ptr),
             ctypes.c_void_p(labels_ptr), ctypes.byref(out), ctypes.c_void_p(stream)))
    return out.value


def gen_stereo(seed: int, width: int, height: int, num_labels: int, num_rects: int, dmax: int, trunc: int,
               want=("cost_f32",)) -> dict:
    """Seeded synthetic stereo instance (ginf_gen_stereo). Returns numpy arrays."""
    out = {}
    ptrs = []
    for key, shape, dt in [("left", (height, width), np.uint8), ("right", (height, width), np.uint8),
                           ("disp", (height, width), np.int32),
                           ("cost_f32", (num_labels, height, width), np.float32),
                           ("cost_i32", (num_labels, height, width), np.int32)]:
        if key in want:
            out[key] = np.empty(shape, dt)
            ptrs.append(out[key].ctypes.data_as(ctypes.c_void_p))
        else:
            ptrs.append(None)
    check(lib().ginf_gen_stereo(ctypes.c_uint64(seed), width, height, num_labels, num_rects, dmax, trunc, *ptrs))
    return out
