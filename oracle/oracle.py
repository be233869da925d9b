"""ORACLE -- test infrastructure, never imported by the product package.

ctypes front end for ``liboracle.so`` (the plain-C restatement of
GridInfCuda's CPU belief propagation, ``src/bp.cc``).  Arrays use the
reference ``Matrix<T>`` layouts with the first index fastest, which in numpy
(C order) read as:

* data costs  ``data[l, y, x]``      (``Matrix<T>(3, W, H, L)``, src/bp.cc:96)
* messages    ``msgs[l, d, y, x]``   (``Matrix<T>(4, w, h, 4, L)``, src/bp.cc:125)
* smoothness  ``table[b, a] = V(a, b)`` (``Matrix<T>(2, L, L)``, src/grid.cc:8)
* labels      ``labels[y, x]``       (``Matrix<int>(2, W, H)``)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SM_EXPLICIT, SM_POTTS, SM_TRUNC_LINEAR, SM_TRUNC_QUAD = 0, 1, 2, 3
NORM_MEAN, NORM_MIN = 0, 1

_lib = None


def build() -> str:
    """Compile liboracle.so with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L: ctypes.CDLL) -> None:
    P = ctypes.c_void_p
    i = ctypes.c_int
    for sfx in ("f32", "i32"):
        getattr(L, f"oracle_pyramid_{sfx}").argtypes = [P, i, i, i, P]
        getattr(L, f"oracle_upsample_{sfx}").argtypes = [P, i, i, P, i, i, i]
        getattr(L, f"oracle_half_sweep_{sfx}").argtypes = [P, P, P, i, i, i, i, P, i, i, i]
        getattr(L, f"oracle_labels_{sfx}").argtypes = [P, i, i, i, P]
        f = getattr(L, f"oracle_hbp_{sfx}")
        f.argtypes = [P, i, i, i, i, P, i, i, i, P, P, P, i]
        f.restype = i
        e = getattr(L, f"oracle_energy_{sfx}")
        e.argtypes = [P, P, P, i, i, i]
        e.restype = ctypes.c_float if sfx == "f32" else ctypes.c_int
        e64 = getattr(L, f"oracle_energy_f64_{sfx}")
        e64.argtypes = [P, P, P, i, i, i]
        e64.restype = ctypes.c_double
        for name in ("oracle_icm_sync", "oracle_icm_async"):
            getattr(L, f"{name}_{sfx}").argtypes = [P, P, i, i, i, P, i, P]
        getattr(L, f"oracle_max_prior_{sfx}").argtypes = [P, i, i, i, P]
    L.oracle_ab_move.argtypes = [P, P, i, i, i, P, i, i]
    L.oracle_alpha_beta.argtypes = [P, P, i, i, i, P, i, P]
    L.oracle_alpha_beta.restype = i
    L.oracle_alpha_beta_ref.argtypes = [P, P, i, i, i, P, i, P, i]
    L.oracle_alpha_beta_ref.restype = i
    L.oracle_max_threads.restype = ctypes.c_int


def _sfx(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f32"
    if dt == np.int32:
        return "i32"
    raise TypeError(f"oracle supports float32 and int32, got {dt}")


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def pyramid(fine: np.ndarray) -> np.ndarray:
    """One data-cost pyramid level: (L,h,w) -> (L,ceil(h/2),ceil(w/2))."""
    fine = np.ascontiguousarray(fine)
    L, h, w = fine.shape
    out = np.empty((L, (h + 1) // 2, (w + 1) // 2), dtype=fine.dtype)
    getattr(lib(), f"oracle_pyramid_{_sfx(fine.dtype)}")(_ptr(fine), w, h, L, _ptr(out))
    return out


def upsample(coarse_msgs: np.ndarray, h: int, w: int) -> np.ndarray:
    """Outbox-frame message upsample: (L,4,ch,cw) -> (L,4,h,w)."""
    coarse_msgs = np.ascontiguousarray(coarse_msgs)
    L, _, ch, cw = coarse_msgs.shape
    out = np.empty((L, 4, h, w), dtype=coarse_msgs.dtype)
    getattr(lib(), f"oracle_upsample_{_sfx(out.dtype)}")(_ptr(coarse_msgs), cw, ch, _ptr(out), w, h, L)
    return out


def half_sweep(data, msgs, belief, table, model, norm, it, threads=1):
    """In-place half-sweep `it` on one level (msgs, belief updated)."""
    L, h, w = data.shape
    s = _sfx(data.dtype)
    getattr(lib(), f"oracle_half_sweep_{s}")(_ptr(data), _ptr(msgs), _ptr(belief), w, h, L, model,
                                             _ptr(table), norm, it, threads)


def labels_from_belief(belief: np.ndarray, labels: np.ndarray) -> np.ndarray:
    L, h, w = belief.shape
    out = np.ascontiguousarray(labels, dtype=np.int32).copy()
    getattr(lib(), f"oracle_labels_{_sfx(belief.dtype)}")(_ptr(belief), w, h, L, _ptr(out))
    return out


def hbp(data: np.ndarray, table: np.ndarray, model: int, levels: int, iters: int,
        norm: int = NORM_MEAN, labels_init: np.ndarray | None = None, threads: int = 1,
        return_state: bool = False):
    """decodeHbp (src/bp.cc:86-180).  Returns labels [H,W] (and, with
    return_state, the level-0 outbox messages [L,4,H,W] and beliefs [L,H,W])."""
    data = np.ascontiguousarray(data)
    table = np.ascontiguousarray(table, dtype=data.dtype)
    L, H, W = data.shape
    assert table.shape == (L, L)
    labels = (np.zeros((H, W), np.int32) if labels_init is None
              else np.ascontiguousarray(labels_init, dtype=np.int32).copy())
    msgs = np.empty((L, 4, H, W), data.dtype) if return_state else None
    bel = np.empty((L, H, W), data.dtype) if return_state else None
    rc = getattr(lib(), f"oracle_hbp_{_sfx(data.dtype)}")(
        _ptr(data), W, H, L, model, _ptr(table), levels, iters, norm, _ptr(labels),
        _ptr(msgs) if msgs is not None else None, _ptr(bel) if bel is not None else None, threads)
    if rc != 0:
        raise ValueError(f"oracle_hbp rejected the arguments (rc={rc})")
    if return_state:
        return labels, msgs, bel
    return labels


def bp(data, table, model, iters, **kw):
    """decodeBp = decodeHbp with one level (src/bp.cc:182-187)."""
    return hbp(data, table, model, 1, iters, **kw)


def energy(data: np.ndarray, table: np.ndarray, labels: np.ndarray):
    """Grid::getLabelingCost accumulated in T (src/grid.cc:102-125)."""
    data = np.ascontiguousarray(data)
    table = np.ascontiguousarray(table, dtype=data.dtype)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    L, H, W = data.shape
    return getattr(lib(), f"oracle_energy_{_sfx(data.dtype)}")(_ptr(data), _ptr(table), _ptr(labels), W, H, L)


def energy_f64(data: np.ndarray, table: np.ndarray, labels: np.ndarray) -> float:
    """The same term set as getLabelingCost, accumulated in double."""
    data = np.ascontiguousarray(data)
    table = np.ascontiguousarray(table, dtype=data.dtype)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    L, H, W = data.shape
    return getattr(lib(), f"oracle_energy_f64_{_sfx(data.dtype)}")(_ptr(data), _ptr(table), _ptr(labels), W, H, L)


def icm(data: np.ndarray, table: np.ndarray, initial: np.ndarray, iters: int, sync: bool = True) -> np.ndarray:
    """decodeIcmSync / decodeIcmAsync (src/icm.cc:5-62)."""
    data = np.ascontiguousarray(data)
    table = np.ascontiguousarray(table, dtype=data.dtype)
    init = np.ascontiguousarray(initial, dtype=np.int32)
    L, H, W = data.shape
    out = np.zeros((H, W), np.int32)
    fn = getattr(lib(), f"oracle_icm_{'sync' if sync else 'async'}_{_sfx(data.dtype)}")
    fn(_ptr(data), _ptr(table), W, H, L, _ptr(init), iters, _ptr(out))
    return out


def ab_move(data: np.ndarray, table: np.ndarray, labels: np.ndarray, alpha: int, beta: int) -> np.ndarray:
    """One alpha-beta swap move (int costs), graph of src/graphcut.cc:9-46."""
    data = np.ascontiguousarray(data, dtype=np.int32)
    table = np.ascontiguousarray(table, dtype=np.int32)
    f = np.ascontiguousarray(labels, dtype=np.int32).copy()
    L, H, W = data.shape
    lib().oracle_ab_move(_ptr(data), _ptr(table), W, H, L, _ptr(f), alpha, beta)
    return f


def alpha_beta(data: np.ndarray, table: np.ndarray, initial: np.ndarray, iters: int):
    """decodeAlphaBeta (src/graphcut.cc:194-279) with fresh per-pair graphs.
    Returns (labels, accepted moves)."""
    data = np.ascontiguousarray(data, dtype=np.int32)
    table = np.ascontiguousarray(table, dtype=np.int32)
    init = np.ascontiguousarray(initial, dtype=np.int32)
    L, H, W = data.shape
    out = np.zeros((H, W), np.int32)
    n = lib().oracle_alpha_beta(_ptr(data), _ptr(table), W, H, L, _ptr(init), iters, _ptr(out))
    return out, n


def alpha_beta_ref(data: np.ndarray, table: np.ndarray, initial: np.ndarray, iters: int, max_loops: int = 1 << 20):
    """decodeAlphaBeta as the reference runs it (src/graphcut.cc:197-279):
    one graph state for the whole decode, rebuilt per pair only at the pair's
    nodes, sequential push-relabel.  Returns (labels, kept moves)."""
    data = np.ascontiguousarray(data, dtype=np.int32)
    table = np.ascontiguousarray(table, dtype=np.int32)
    init = np.ascontiguousarray(initial, dtype=np.int32)
    L, H, W = data.shape
    out = np.zeros((H, W), np.int32)
    n = lib().oracle_alpha_beta_ref(_ptr(data), _ptr(table), W, H, L, _ptr(init), iters, _ptr(out), max_loops)
    if n < 0:
        raise RuntimeError("alpha_beta_ref: a push-relabel loop did not end")
    return out, n


def max_prior(data: np.ndarray) -> np.ndarray:
    """decodeMaxPrior (src/include/estimators.h:10-25)."""
    data = np.ascontiguousarray(data)
    L, H, W = data.shape
    out = np.zeros((H, W), np.int32)
    getattr(lib(), f"oracle_max_prior_{_sfx(data.dtype)}")(_ptr(data), W, H, L, _ptr(out))
    return out


def max_threads() -> int:
    return lib().oracle_max_threads()


# --- smoothness tables, as the Grid setters fill them (src/grid.cc:70-100) ---

def table_potts(L: int, lam, dtype=np.float32) -> np.ndarray:
    a = np.arange(L)
    return (lam * (a[:, None] != a[None, :])).astype(dtype)


def table_trunc_linear(L: int, scale, d, dtype=np.float32) -> np.ndarray:
    a = np.arange(L)
    return np.minimum(scale * np.abs(a[:, None] - a[None, :]), d).astype(dtype)


def table_trunc_quad(L: int, scale, d, dtype=np.float32) -> np.ndarray:
    a = np.arange(L)
    return np.minimum(scale * (a[:, None] - a[None, :]) ** 2, d).astype(dtype)
