"""CPU oracle of MCPNet's hot paths -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.

ctypes binding of oracle/liboracle.so (the C restatement in mcp_oracle.c) and of
oracle/_ref/libref_maxmin.so (the reference's own simd_maxmin.hpp behind C entry points).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_int64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
LIB_AVX512 = HERE / "libport_avx512.so"
REF_DIR = HERE / "_ref"

_P = c_void_p
_SIGS = {
    "orc_knots_f32": (None, [c_int, c_int, _P]),
    "orc_knots_f64": (None, [c_int, c_int, _P]),
    "orc_basis_f32": (c_float, [_P, c_int, c_int, c_int, c_float]),
    "orc_basis_f64": (c_double, [_P, c_int, c_int, c_int, c_double]),
    "orc_minmax_scale_f32": (None, [_P, c_int64, c_int64, _P]),
    "orc_bspline_weights_f32": (None, [_P, c_int64, c_int64, c_int, c_int, _P]),
    "orc_bspline_weights_f64": (None, [_P, c_int64, c_int64, c_int, c_int, c_int, _P, _P]),
    "orc_bspline_mi_f64": (None, [_P, c_int64, c_int64, c_int, c_int, _P]),
    "orc_bspline_weights_x64": (None, [_P, c_int64, c_int64, c_int, c_int, c_int, _P, _P]),
    "orc_bspline_mi_x64": (None, [_P, c_int64, c_int64, c_int, c_int, _P]),
    "orc_bspline_mi_pairs_f64": (None, [_P, c_int64, c_int64, c_int, c_int, _P, _P, c_int64, _P]),
    "orc_bspline_mi_loss_f64": (c_double, [_P, c_int64, c_int64, c_int, c_int, _P, _P, _P, c_int64]),
    "orc_bspline_mi_loss_x64": (c_double, [_P, c_int64, c_int64, c_int, c_int, _P, _P, _P, c_int64]),
    "orc_bspline_mi_grad_f64": (None, [_P, c_int64, c_int64, c_int, c_int, _P, _P, _P, c_int64, _P]),
    "orc_cumhist_i32": (None, [_P, c_int64, c_int64, _P]),
    "orc_ap_mi_pair": (c_double, [_P, _P, c_int64, c_int64, c_int64, _P, c_int32, _P]),
    "orc_ap_mi": (None, [_P, c_int64, c_int64, _P]),
    "orc_ap_mi_rows": (None, [_P, c_int64, c_int64, c_int64, c_int64, _P]),
    "orc_ap_tests_rows": (c_int64, [_P, c_int64, c_int64, c_int64, c_int64]),
    "orc_ap_mi_cells": (c_double, [_P, c_int64, c_int64, c_int64, c_int64, _P, c_int32, _P]),
    "orc_maxmin_f32": (None, [_P, c_int64, _P, c_int64, _P, c_int64, c_int64, c_int64, _P, c_int64]),
    "orc_maxmin_f64": (None, [_P, c_int64, _P, c_int64, _P, c_int64, c_int64, c_int64, _P, c_int64]),
    "orc_ratio_f32": (None, [_P, _P, c_int64, _P]),
    "orc_ratio_f64": (None, [_P, _P, c_int64, _P]),
    "port_bspline_mi_rows_f32": (None, [_P, c_int64, c_int64, c_int, c_int64, c_int64, _P]),
    "orc_combo_f32": (None, [_P, _P, _P, _P, c_int64, _P, _P]),
    "orc_combo_f64": (None, [_P, _P, _P, _P, c_int64, _P, _P]),
    "orc_stouffer_f32": (None, [_P, c_int64, c_int, _P]),
    "orc_stouffer_f64": (None, [_P, c_int64, c_int, _P]),
    "orc_stouffer_tf_f32": (None, [_P, c_int64, c_int64, _P]),
    "orc_stouffer_tf_f64": (None, [_P, c_int64, c_int64, _P]),
    "orc_dpi_f32": (c_int64, [_P, c_int64, c_float, _P, _P]),
    "orc_dpi_f64": (c_int64, [_P, c_int64, c_double, _P, _P]),
}

_lib = None
_port = None
_ref = None


def build(verbose: bool = False) -> None:
    """Compiles the oracle (and oracle/_ref when the reference tree is present)."""
    res = subprocess.run(["make", "-C", str(HERE), "all"], capture_output=True, text=True)
    if verbose or res.returncode != 0:
        print(res.stdout + res.stderr)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed")


def _has_avx512() -> bool:
    try:
        return "avx512f" in Path("/proc/cpuinfo").read_text()
    except OSError:
        return False


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def port_lib():
    """The FP32 baseline port: the AVX-512 build on AVX-512 hosts, else the AVX2 one."""
    global _port
    if _port is None:
        if _has_avx512() and LIB_AVX512.exists():
            L = ctypes.CDLL(str(LIB_AVX512))
            L.port_bspline_mi_rows_f32.restype = None
            L.port_bspline_mi_rows_f32.argtypes = _SIGS["port_bspline_mi_rows_f32"][1]
            _port = L
        else:
            _port = lib()
    return _port


def ref_lib():
    """oracle/_ref/libref_maxmin*.so (the reference's max_of_pairmin), or None if not built."""
    global _ref
    if _ref is None:
        cand = REF_DIR / ("libref_maxmin_avx512.so" if _has_avx512() else "libref_maxmin.so")
        if not cand.exists():
            cand = REF_DIR / "libref_maxmin.so"
        if not cand.exists():
            return None
        L = ctypes.CDLL(str(cand))
        for name in ("ref_maxmin_f32", "ref_maxmin_f64"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [_P, c_int64, _P, c_int64, _P, c_int64, c_int64, c_int64, _P, c_int64]
        _ref = L
    return _ref


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ------------------------------------------------------------------ B-spline
def knots(bins=10, order=3, dtype=np.float32):
    out = np.zeros(bins + order, dtype=dtype)
    (lib().orc_knots_f32 if dtype == np.float32 else lib().orc_knots_f64)(bins, order, out.ctypes.data)
    return out


def basis_f32(t, bins=10, order=3):
    k = knots(bins, order, np.float32)
    return np.array([[lib().orc_basis_f32(k.ctypes.data, bins, b, order, float(x)) for b in range(bins)]
                     for x in np.atleast_1d(t)], dtype=np.float32)


def basis_f64(t, bins=10, order=3):
    k = knots(bins, order, np.float64)
    return np.array([[lib().orc_basis_f64(k.ctypes.data, bins, b, order, float(x)) for b in range(bins)]
                     for x in np.atleast_1d(t)], dtype=np.float64)


def minmax_scale_f32(X):
    X = _c(X, np.float32)
    Y = np.empty_like(X)
    lib().orc_minmax_scale_f32(X.ctypes.data, X.shape[0], X.shape[1], Y.ctypes.data)
    return Y


def bspline_weights_f32(Xs, bins=10, order=3):
    Xs = _c(Xs, np.float32)
    G, S = Xs.shape
    W = np.empty((G, bins * S + 1), dtype=np.float32)
    lib().orc_bspline_weights_f32(Xs.ctypes.data, G, S, bins, order, W.ctypes.data)
    return W


def bspline_weights_f64(X, bins=10, order=3, scale=True):
    X = _c(X, np.float32)
    G, S = X.shape
    W = np.empty((G, bins, S), dtype=np.float64)
    H = np.empty(G, dtype=np.float64)
    lib().orc_bspline_weights_f64(X.ctypes.data, G, S, bins, order, int(scale), W.ctypes.data, H.ctypes.data)
    return W, H


def bspline_mi_f64(X, bins=10, order=3):
    X = _c(X, np.float32)
    G, S = X.shape
    MI = np.empty((G, G), dtype=np.float64)
    lib().orc_bspline_mi_f64(X.ctypes.data, G, S, bins, order, MI.ctypes.data)
    return MI


def bspline_weights_x64(X, bins=10, order=3, scale=True):
    """bspline_weights_f64 on double rows (the reference's double DataType)."""
    X = _c(X, np.float64)
    G, S = X.shape
    W = np.empty((G, bins, S), dtype=np.float64)
    H = np.empty(G, dtype=np.float64)
    lib().orc_bspline_weights_x64(X.ctypes.data, G, S, bins, order, int(scale), W.ctypes.data, H.ctypes.data)
    return W, H


def bspline_mi_x64(X, bins=10, order=3):
    """bspline_mi_f64 on double rows (the reference's double DataType)."""
    X = _c(X, np.float64)
    G, S = X.shape
    MI = np.empty((G, G), dtype=np.float64)
    lib().orc_bspline_mi_x64(X.ctypes.data, G, S, bins, order, MI.ctypes.data)
    return MI


def bspline_mi_pairs_f64(X, pi, pj, bins=10, order=3):
    X = _c(X, np.float32)
    G, S = X.shape
    pi = _c(pi, np.int32)
    pj = _c(pj, np.int32)
    out = np.empty(len(pi), dtype=np.float64)
    lib().orc_bspline_mi_pairs_f64(X.ctypes.data, G, S, bins, order, pi.ctypes.data, pj.ctypes.data, len(pi),
                                   out.ctypes.data)
    return out


def bspline_mi_loss_f64(X, pi, pj, m, bins=10, order=3):
    X = _c(X, np.float32)
    pi, pj, m = _c(pi, np.int32), _c(pj, np.int32), _c(m, np.float32)
    return lib().orc_bspline_mi_loss_f64(X.ctypes.data, X.shape[0], X.shape[1], bins, order, pi.ctypes.data,
                                         pj.ctypes.data, m.ctypes.data, len(pi))


def bspline_mi_loss_x64(X, pi, pj, m, bins=10, order=3):
    X = _c(X, np.float64)
    pi, pj, m = _c(pi, np.int32), _c(pj, np.int32), _c(m, np.float32)
    return lib().orc_bspline_mi_loss_x64(X.ctypes.data, X.shape[0], X.shape[1], bins, order, pi.ctypes.data,
                                         pj.ctypes.data, m.ctypes.data, len(pi))


def bspline_mi_grad_f64(X, pi, pj, m, bins=10, order=3):
    X = _c(X, np.float32)
    pi, pj, m = _c(pi, np.int32), _c(pj, np.int32), _c(m, np.float32)
    dX = np.zeros(X.shape, dtype=np.float64)
    lib().orc_bspline_mi_grad_f64(X.ctypes.data, X.shape[0], X.shape[1], bins, order, pi.ctypes.data, pj.ctypes.data,
                                  m.ctypes.data, len(pi), dX.ctypes.data)
    return dX


def port_bspline_mi_rows_f32(W, S, bins, row0, nrows):
    W = _c(W, np.float32)
    G = W.shape[0]
    out = np.empty((nrows, G), dtype=np.float32)
    port_lib().port_bspline_mi_rows_f32(W.ctypes.data, G, S, bins, row0, nrows, out.ctypes.data)
    return out


# ------------------------------------------------------------------ adaptive partitioning
def cumhist(R):
    R = _c(R, np.int32)
    G, S = R.shape
    H = np.empty((G, S + 1), dtype=np.int32)
    lib().orc_cumhist_i32(R.ctypes.data, G, S, H.ctypes.data)
    return H


def ap_mi(R):
    R = _c(R, np.int32)
    G, S = R.shape
    MI = np.empty((G, G), dtype=np.float32)
    lib().orc_ap_mi(R.ctypes.data, G, S, MI.ctypes.data)
    return MI


def ap_tests_rows(R, row0, nrows):
    """Point comparisons the reference's quadrant scans make for rows [row0, row0 + nrows) x all genes."""
    R = _c(R, np.int32)
    return int(lib().orc_ap_tests_rows(R.ctypes.data, R.shape[0], R.shape[1], row0, nrows))


def ap_mi_rows(R, row0, nrows):
    """Rows row0..row0+nrows-1 of the AP-MI matrix (float32, nrows x G)."""
    R = _c(R, np.int32)
    G, S = R.shape
    out = np.empty((nrows, G), dtype=np.float32)
    lib().orc_ap_mi_rows(R.ctypes.data, G, S, row0, nrows, out.ctypes.data)
    return out


def ap_mi_cells(R, x, y, max_leaves=1 << 16):
    """(leaf array (n, 3) of (nxy, nx, ny) in summation order, FP64 MI) for pair (x, y)."""
    R = _c(R, np.int32)
    G, S = R.shape
    cells = np.zeros((max_leaves, 3), dtype=np.int32)
    nl = c_int32(0)
    v = lib().orc_ap_mi_cells(R.ctypes.data, G, S, int(x), int(y), cells.ctypes.data, max_leaves, ctypes.byref(nl))
    return cells[: nl.value].copy(), v


# ------------------------------------------------------------------ max-min / ratio
def maxmin(A, B=None, kmask=None):
    dt = np.float64 if np.asarray(A).dtype == np.float64 else np.float32
    A = _c(A, dt)
    B = A if B is None else _c(B, dt)
    km = None if kmask is None else _c(kmask, dt)
    M, K = A.shape
    N = B.shape[0]
    C = np.empty((M, N), dtype=dt)
    fn = lib().orc_maxmin_f64 if dt == np.float64 else lib().orc_maxmin_f32
    fn(A.ctypes.data, K, B.ctypes.data, K, km.ctypes.data if km is not None else None, M, N, K, C.ctypes.data, N)
    return C


def ref_maxmin(A, B=None, kmask=None):
    """The reference's own max_of_pairmin over all (i, j); None when oracle/_ref is not built."""
    L = ref_lib()
    if L is None:
        return None
    dt = np.float64 if np.asarray(A).dtype == np.float64 else np.float32
    A = _c(A, dt)
    B = A if B is None else _c(B, dt)
    km = None if kmask is None else _c(kmask, dt)
    M, K = A.shape
    N = B.shape[0]
    C = np.empty((M, N), dtype=dt)
    fn = L.ref_maxmin_f64 if dt == np.float64 else L.ref_maxmin_f32
    fn(A.ctypes.data, K, B.ctypes.data, K, km.ctypes.data if km is not None else None, M, N, K, C.ctypes.data, N)
    return C


def ratio(A, Mm):
    dt = np.float64 if np.asarray(A).dtype == np.float64 else np.float32
    A = _c(A, dt)
    Mm = _c(Mm, dt)
    out = np.empty_like(A)
    fn = lib().orc_ratio_f64 if dt == np.float64 else lib().orc_ratio_f32
    fn(A.ctypes.data, Mm.ctypes.data, A.size, out.ctypes.data)
    return out


def mcp(A, k):
    """MCP_k on the CPU: (ratio(A, A^(x)k), A^(x)k) following mcpnet_helper.hpp:188-364."""
    A = np.ascontiguousarray(A)
    M1 = maxmin(A, A)
    if k == 2:
        Mk = M1
    elif k == 3:
        Mk = maxmin(A, M1)
    elif k == 4:
        Mk = maxmin(M1, M1)
    else:
        raise ValueError("k must be 2, 3 or 4")
    return ratio(A, Mk), Mk


# ------------------------------------------------------------------ ensemble and filters
def maxmins(A):
    """compute_maxmins (mcpnet_helper.hpp:436-460): (A(x)A, A(x)M1, M1(x)M1)."""
    M1 = maxmin(A, A)
    return M1, maxmin(A, M1), maxmin(M1, M1)


def combo(MI, M1, M2, M3, coeffs):
    dt = np.float64 if np.asarray(MI).dtype == np.float64 else np.float32
    mats = [_c(x if x is not None else np.zeros_like(MI), dt) for x in (MI, M1, M2, M3)]
    cf = _c(coeffs, dt)
    out = np.empty_like(mats[0])
    fn = lib().orc_combo_f64 if dt == np.float64 else lib().orc_combo_f32
    fn(mats[0].ctypes.data, mats[1].ctypes.data, mats[2].ctypes.data, mats[3].ctypes.data, mats[0].size, cf.ctypes.data,
       out.ctypes.data)
    return out


def stouffer(X, exclude_ones=True):
    dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
    X = _c(X, dt)
    out = np.empty_like(X)
    fn = lib().orc_stouffer_f64 if dt == np.float64 else lib().orc_stouffer_f32
    fn(X.ctypes.data, X.shape[0], int(exclude_ones), out.ctypes.data)
    return out


def stouffer_tf(X):
    """compute_stouffer_tf on a TFs x genes matrix: (row z + column z) / sqrt(2)."""
    dt = np.float64 if np.asarray(X).dtype == np.float64 else np.float32
    X = _c(X, dt)
    out = np.empty_like(X)
    fn = lib().orc_stouffer_tf_f64 if dt == np.float64 else lib().orc_stouffer_tf_f32
    fn(X.ctypes.data, X.shape[0], X.shape[1], out.ctypes.data)
    return out


def dpi(A, tolerance=0.1, tfs=None):
    """(DPI-filtered A, number of changed entries) following dpi.hpp's per-pair scan."""
    dt = np.float64 if np.asarray(A).dtype == np.float64 else np.float32
    A = _c(A, dt)
    tf = None if tfs is None else _c(tfs, dt)
    out = np.empty_like(A)
    fn = lib().orc_dpi_f64 if dt == np.float64 else lib().orc_dpi_f32
    n = fn(A.ctypes.data, A.shape[0], tolerance, tf.ctypes.data if tf is not None else None, out.ctypes.data)
    return out, int(n)
