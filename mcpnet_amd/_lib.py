"""ctypes binding of libmcpgpu.so (the C ABI declared in include/mcpgpu/mcpgpu.h).

Arrays are accepted either as numpy arrays (host memory: the library copies in and out) or as
torch tensors on a ROCm device (device memory: pointers are passed straight through).  There is
no CPU fallback: if the library is not built, or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libmcpgpu.so"

MCPG_HOST, MCPG_DEVICE = 0, 1
PRECISIONS = {"fp32": 0, "bf16": 1, "bf16x3": 2, "i8x3": 3, "fp64": 4}
STATUS = {
    0: "ok", -1: "invalid argument", -2: "shape mismatch", -3: "non-finite input", -4: "out of device memory",
    -5: "HIP runtime error", -6: "no usable HIP device", -7: "unsupported size",
}


class MCPGError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        msg = f"{where}: {STATUS.get(status, 'unknown status')} ({status})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


_lib = None

# name -> argtypes (restype is c_int for every entry point except the string getters)
_P = c_void_p
_SIGS = {
    "mcpg_create": [POINTER(c_void_p), c_int],
    "mcpg_destroy": [_P],
    "mcpg_synchronize": [_P],
    "mcpg_get_stream": [_P, POINTER(c_void_p)],
    "mcpg_last_launch_info": [_P, POINTER(c_int), POINTER(c_char_p)],
    "mcpg_set_profiling": [_P, c_int],
    "mcpg_last_timings": [_P, c_int, POINTER(c_char_p), POINTER(c_float), POINTER(c_int)],
    "mcpg_minmax_scale_f32": [_P, c_int, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_bspline_weights_f32": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, c_int, _P, c_int64],
    "mcpg_bspline_mi_weights_f32": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, c_int, _P, c_int64],
    "mcpg_bspline_mi_f32": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, c_int, c_int, _P, c_int64],
    "mcpg_minmax_scale_f64": [_P, c_int, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_bspline_weights_f64": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, c_int, _P, c_int64],
    "mcpg_bspline_mi_weights_f64": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, _P, c_int64],
    "mcpg_bspline_mi_f64": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, c_int, _P, c_int64],
    "mcpg_ap_mi_f64": [_P, c_int, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_ap_mi_block_f64": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, _P,
                             c_int64],
    "mcpg_bspline_mi_block_f64": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, c_int, c_int64, c_int64, c_int64,
                                  c_int64, _P, c_int64],
    "mcpg_bspline_mi_block_f32": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, c_int, c_int, c_int64, c_int64, c_int64,
                                  c_int64, _P, c_int64],
    "mcpg_bspline_mi_backward_f32": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, c_int, _P, _P, _P, c_int64,
                                     _P, c_int64, POINTER(c_double)],
    "mcpg_add_noise_f32": [_P, c_int, _P, c_int64, c_int64, c_int64, c_uint64, _P, c_int64],
    "mcpg_rank_f32": [_P, c_int, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_cumhist_i32": [_P, c_int, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_ap_mi_i32": [_P, c_int, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_ap_mi_block_i32": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_ap_mi_cells_i32": [_P, c_int, _P, c_int64, c_int64, c_int64, _P, _P, c_int64, c_int32, _P, _P, _P],
    "mcpg_maxmin_f32": [_P, c_int, _P, c_int64, _P, c_int64, _P, c_int64, c_int64, c_int64, c_int, _P, c_int64],
    "mcpg_maxmin_f64": [_P, c_int, _P, c_int64, _P, c_int64, _P, c_int64, c_int64, c_int64, c_int, _P, c_int64],
    "mcpg_ratio_f32": [_P, c_int, _P, c_int64, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_ratio_f64": [_P, c_int, _P, c_int64, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_elementwise_max_f32": [_P, c_int, _P, c_int64, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_elementwise_max_f64": [_P, c_int, _P, c_int64, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_mcp_f32": [_P, c_int, _P, c_int64, c_int64, c_int, _P, c_int64, _P, c_int64],
    "mcpg_mcp_f64": [_P, c_int, _P, c_int64, c_int64, c_int, _P, c_int64, _P, c_int64],
    "mcpg_mcp_f32_noalloc": [_P, _P, c_int64, c_int64, c_int, _P, c_int64, _P, c_int64],
    "mcpg_maxmins_f32": [_P, c_int, _P, c_int64, c_int64, _P, c_int64, _P, c_int64, _P, c_int64],
    "mcpg_maxmins_f64": [_P, c_int, _P, c_int64, c_int64, _P, c_int64, _P, c_int64, _P, c_int64],
    "mcpg_combo_f32": [_P, c_int, _P, c_int64, _P, c_int64, _P, c_int64, _P, c_int64, c_int64, c_int64, _P, _P, c_int64],
    "mcpg_combo_f64": [_P, c_int, _P, c_int64, _P, c_int64, _P, c_int64, _P, c_int64, c_int64, c_int64, _P, _P, c_int64],
    "mcpg_stouffer_f32": [_P, c_int, _P, c_int64, c_int64, c_int, _P, c_int64],
    "mcpg_stouffer_f64": [_P, c_int, _P, c_int64, c_int64, c_int, _P, c_int64],
    "mcpg_mcpnet_f32": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, c_int, c_int, c_int, c_uint64, c_double, _P,
                        _P, c_int64, _P, c_int64, _P, c_int64, _P, c_int64],
    "mcpg_mcpnet_f64": [_P, c_int, _P, c_int64, c_int64, c_int64, c_int, c_int, c_int, c_int, c_uint64, c_double, _P,
                        _P, c_int64, _P, c_int64, _P, c_int64, _P, c_int64],
    "mcpg_stouffer_tf_f32": [_P, c_int, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_stouffer_tf_f64": [_P, c_int, _P, c_int64, c_int64, c_int64, _P, c_int64],
    "mcpg_dpi_f32": [_P, c_int, _P, c_int64, c_int64, c_float, _P, _P, c_int64, _P, c_int64, POINTER(c_int64)],
    "mcpg_dpi_f64": [_P, c_int, _P, c_int64, c_int64, c_double, _P, _P, c_int64, _P, c_int64, POINTER(c_int64)],
}
EXPORTS = ["mcpg_status_string", "mcpg_last_error", *_SIGS.keys()]


def load() -> ctypes.CDLL:
    """Loads libmcpgpu.so (built in-tree); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("MCPG_LIBRARY", LIB_PATH))
    # torch's ROCm build ships its own libamdhip64 and links it by the unversioned name; if this
    # library were loaded first, torch would load a second HIP runtime into the process and
    # lose the device.  Loading torch first makes both share its runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not path.exists():
        raise RuntimeError(f"{path} is not built; run `python -m mcpnet_amd.build` (HIP toolchain required)")
    lib = ctypes.CDLL(str(path))
    for name, argtypes in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = c_int
    lib.mcpg_status_string.argtypes = [c_int]
    lib.mcpg_status_string.restype = c_char_p
    lib.mcpg_last_error.argtypes = [c_void_p]
    lib.mcpg_last_error.restype = c_char_p
    _lib = lib
    return lib


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


class _Mat:
    """A 2-D (or 1-D) operand: pointer, leading dimension and where it lives."""

    def __init__(self, x, dtype, name):
        self.obj = x
        if _is_torch(x):
            import torch
            want = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32}[dtype]
            if x.dtype != want:
                raise TypeError(f"{name}: expected {want}, got {x.dtype}")
            if not x.is_cuda:
                raise TypeError(f"{name}: torch tensors must be on the GPU (use numpy for host memory)")
            if x.dim() == 2:
                if x.stride(1) != 1:
                    raise ValueError(f"{name}: rows must be contiguous")
                self.ld = x.stride(0)
            else:
                if x.dim() != 1 or x.stride(0) != 1:
                    raise ValueError(f"{name}: expected a contiguous vector")
                self.ld = x.shape[0]
            self.ptr = x.data_ptr()
            self.kind = MCPG_DEVICE
        else:
            a = np.asarray(x)
            if a.dtype != dtype:
                raise TypeError(f"{name}: expected {np.dtype(dtype)}, got {a.dtype}")
            if a.ndim == 2 and a.size == 0:
                a = np.ascontiguousarray(a)
                self.ld = max(a.shape[1], 1)
            elif a.ndim == 2:
                if a.strides[1] != a.itemsize:
                    raise ValueError(f"{name}: rows must be contiguous")
                self.ld = a.strides[0] // a.itemsize if a.shape[0] > 1 else a.shape[1]
            else:
                a = np.ascontiguousarray(a)
                self.ld = a.shape[0] if a.ndim == 1 else 0
            self.obj = a
            self.ptr = a.ctypes.data
            self.kind = MCPG_HOST
        self.shape = tuple(self.obj.shape)


def _empty_like_kind(kind, shape, dtype, ref):
    if kind == MCPG_DEVICE:
        import torch
        tdt = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32}[dtype]
        return torch.empty(shape, dtype=tdt, device=ref.device)
    return np.empty(shape, dtype=dtype)


def _out_mat(out, kind, shape, dtype, ref, name):
    """An output operand: allocated like the inputs when not given, else checked to be of the
    inputs' kind (host or device), the expected shape and dtype, with contiguous rows."""
    if out is None:
        out = _empty_like_kind(kind, shape, dtype, ref)
    m = _Mat(out, dtype, name)
    if m.kind != kind:
        raise TypeError(f"{name}: the inputs are {'device' if kind == MCPG_DEVICE else 'host'} arrays, so the output "
                        f"must be one too")
    if m.shape != tuple(shape):
        raise ValueError(f"{name}: expected shape {tuple(shape)}, got {m.shape}")
    return out, m


def _check_block(G, row0, nrows, col0, ncols):
    for name, v in (("row0", row0), ("nrows", nrows), ("col0", col0), ("ncols", ncols)):
        if not isinstance(v, (int, np.integer)) or v < 0:
            raise ValueError(f"{name} must be a non-negative integer, got {v!r}")
    if row0 + nrows > G or col0 + ncols > G:
        raise ValueError(f"block rows {row0}+{nrows}, columns {col0}+{ncols} exceed the {G} x {G} matrix")


def _dtype_of(x):
    return np.float64 if str(getattr(x, "dtype", "")).endswith("float64") else np.float32


def _same_kind(*mats):
    kinds = {m.kind for m in mats if m is not None}
    if len(kinds) > 1:
        raise TypeError("mixing host (numpy) and device (torch) operands in one call is not supported")
    return kinds.pop()


class Handle:
    """One libmcpgpu handle (device, HIP stream, workspace).  Not thread-safe."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = c_void_p()
        st = self.lib.mcpg_create(byref(h), int(device))
        if st != 0:
            raise MCPGError(st, "mcpg_create")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.mcpg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ helpers
    def _check(self, st: int, where: str):
        if st != 0:
            detail = self.lib.mcpg_last_error(self.h)
            raise MCPGError(st, where, detail.decode() if detail else "")

    def _pre(self, kind):
        if kind == MCPG_DEVICE:
            import torch
            torch.cuda.current_stream().synchronize()

    @property
    def stream(self) -> int:
        s = c_void_p()
        self._check(self.lib.mcpg_get_stream(self.h, byref(s)), "mcpg_get_stream")
        return s.value or 0

    def last_launch_info(self):
        n = c_int()
        name = c_char_p()
        self._check(self.lib.mcpg_last_launch_info(self.h, byref(n), byref(name)), "mcpg_last_launch_info")
        return n.value, (name.value or b"").decode()

    def synchronize(self):
        self._check(self.lib.mcpg_synchronize(self.h), "mcpg_synchronize")

    def set_profiling(self, enable: bool = True):
        self._check(self.lib.mcpg_set_profiling(self.h, int(enable)), "mcpg_set_profiling")

    def last_timings(self):
        """[(kernel name, ms)] for the launches of the most recent call (profiling on)."""
        n = c_int()
        self._check(self.lib.mcpg_last_timings(self.h, 0, None, None, byref(n)), "mcpg_last_timings")
        names = (c_char_p * max(n.value, 1))()
        ms = (c_float * max(n.value, 1))()
        self._check(self.lib.mcpg_last_timings(self.h, n.value, names, ms, byref(n)), "mcpg_last_timings")
        return [(names[i].decode(), float(ms[i])) for i in range(n.value)]

    # ------------------------------------------------------------------ B-spline MI
    def minmax_scale(self, X, out=None):
        """MinMaxScale per row; float64 rows run in double (the reference's double DataType)."""
        dt = _dtype_of(X)
        x = _Mat(X, dt, "X")
        G, S = x.shape
        out, y = _out_mat(out, x.kind, (G, S), dt, X, "out")
        fn, name = ((self.lib.mcpg_minmax_scale_f64, "mcpg_minmax_scale_f64") if dt == np.float64 else
                    (self.lib.mcpg_minmax_scale_f32, "mcpg_minmax_scale_f32"))
        self._pre(x.kind)
        self._check(fn(self.h, x.kind, x.ptr, G, S, x.ld, y.ptr, y.ld), name)
        return out

    def bspline_weights(self, Xs, bins=10, order=3, out=None):
        """Reference-layout weight rows (bins*S weights, then H); float64 rows run in double."""
        dt = _dtype_of(Xs)
        x = _Mat(Xs, dt, "Xs")
        G, S = x.shape
        out, w = _out_mat(out, x.kind, (G, bins * S + 1), dt, Xs, "out")
        fn, name = ((self.lib.mcpg_bspline_weights_f64, "mcpg_bspline_weights_f64") if dt == np.float64 else
                    (self.lib.mcpg_bspline_weights_f32, "mcpg_bspline_weights_f32"))
        self._pre(x.kind)
        self._check(fn(self.h, x.kind, x.ptr, G, S, x.ld, bins, order, w.ptr, w.ld), name)
        return out

    def bspline_mi_weights(self, W, S, bins=10, precision="fp32", out=None):
        """MI from reference-layout weight rows; precision "fp64" takes float64 rows."""
        dt = np.float64 if precision == "fp64" else np.float32
        w = _Mat(W, dt, "W")
        G = w.shape[0]
        out, m = _out_mat(out, w.kind, (G, G), dt, W, "out")
        self._pre(w.kind)
        if dt == np.float64:
            self._check(self.lib.mcpg_bspline_mi_weights_f64(self.h, w.kind, w.ptr, G, S, w.ld, bins, m.ptr, m.ld),
                        "mcpg_bspline_mi_weights_f64")
        else:
            self._check(self.lib.mcpg_bspline_mi_weights_f32(self.h, w.kind, w.ptr, G, S, w.ld, bins,
                                                             PRECISIONS[precision], m.ptr, m.ld),
                        "mcpg_bspline_mi_weights_f32")
        return out

    def bspline_mi(self, X, bins=10, order=3, precision="fp32", out=None):
        """All-pairs B-spline MI.  precision "fp64" takes float64 X and returns float64 MI (the
        reference's double DataType); the others take and return float32."""
        dt = np.float64 if precision == "fp64" else np.float32
        x = _Mat(X, dt, "X")
        G, S = x.shape
        out, m = _out_mat(out, x.kind, (G, G), dt, X, "out")
        _same_kind(x, m)
        self._pre(x.kind)
        if dt == np.float64:
            self._check(self.lib.mcpg_bspline_mi_f64(self.h, x.kind, x.ptr, G, S, x.ld, bins, order, m.ptr, m.ld),
                        "mcpg_bspline_mi_f64")
        else:
            self._check(self.lib.mcpg_bspline_mi_f32(self.h, x.kind, x.ptr, G, S, x.ld, bins, order,
                                                     PRECISIONS[precision], m.ptr, m.ld), "mcpg_bspline_mi_f32")
        return out

    def bspline_mi_block(self, X, row0, nrows, col0, ncols, bins=10, order=3, precision="fp32", out=None):
        """Rows row0 .. row0+nrows-1, columns col0 .. col0+ncols-1 of bspline_mi(X)."""
        dt = np.float64 if precision == "fp64" else np.float32
        x = _Mat(X, dt, "X")
        G, S = x.shape
        _check_block(G, row0, nrows, col0, ncols)
        out, m = _out_mat(out, x.kind, (nrows, ncols), dt, X, "out")
        _same_kind(x, m)
        self._pre(x.kind)
        if dt == np.float64:
            self._check(self.lib.mcpg_bspline_mi_block_f64(self.h, x.kind, x.ptr, G, S, x.ld, bins, order, row0, nrows,
                                                           col0, ncols, m.ptr, max(m.ld, 1)), "mcpg_bspline_mi_block_f64")
        else:
            self._check(self.lib.mcpg_bspline_mi_block_f32(self.h, x.kind, x.ptr, G, S, x.ld, bins, order,
                                                           PRECISIONS[precision], row0, nrows, col0, ncols, m.ptr,
                                                           max(m.ld, 1)), "mcpg_bspline_mi_block_f32")
        return out

    def bspline_mi_backward(self, X, pi, pj, m, bins=10, order=3, out=None):
        """dL/dX for L = sum_n m[n] * MI(pi[n], pj[n]); returns (dX, L)."""
        x = _Mat(X, np.float32, "X")
        G, S = x.shape
        pi = np.ascontiguousarray(pi, dtype=np.int32)
        pj = np.ascontiguousarray(pj, dtype=np.int32)
        mm = np.ascontiguousarray(m, dtype=np.float32)
        if not (pi.shape == pj.shape == mm.shape):
            raise ValueError("pair arrays must have the same length")
        out, d = _out_mat(out, x.kind, (G, S), np.float32, X, "out")
        L = c_double()
        self._pre(x.kind)
        self._check(self.lib.mcpg_bspline_mi_backward_f32(self.h, x.kind, x.ptr, G, S, x.ld, bins, order,
                                                          pi.ctypes.data, pj.ctypes.data, mm.ctypes.data, len(pi),
                                                          d.ptr, d.ld, byref(L)),
                    "mcpg_bspline_mi_backward_f32")
        return out, L.value

    # ------------------------------------------------------------------ adaptive partitioning
    def add_noise(self, X, seed, out=None):
        x = _Mat(X, np.float32, "X")
        G, S = x.shape
        out, y = _out_mat(out, x.kind, (G, S), np.float32, X, "out")
        self._pre(x.kind)
        self._check(self.lib.mcpg_add_noise_f32(self.h, x.kind, x.ptr, G, S, x.ld, c_uint64(seed), y.ptr, y.ld),
                    "mcpg_add_noise_f32")
        return out

    def rank(self, X, out=None):
        x = _Mat(X, np.float32, "X")
        G, S = x.shape
        out, r = _out_mat(out, x.kind, (G, S), np.int32, X, "out")
        self._pre(x.kind)
        self._check(self.lib.mcpg_rank_f32(self.h, x.kind, x.ptr, G, S, x.ld, r.ptr, r.ld), "mcpg_rank_f32")
        return out

    def cumhist(self, R, out=None):
        r = _Mat(R, np.int32, "R")
        G, S = r.shape
        out, hm = _out_mat(out, r.kind, (G, S + 1), np.int32, R, "out")
        self._pre(r.kind)
        self._check(self.lib.mcpg_cumhist_i32(self.h, r.kind, r.ptr, G, S, r.ld, hm.ptr, hm.ld), "mcpg_cumhist_i32")
        return out

    def ap_mi(self, R, out=None, dtype=np.float32):
        """All-pairs AP-MI of rank rows; dtype float64 keeps the MI in double (mcpg_ap_mi_f64)."""
        dt = np.float64 if np.dtype(dtype) == np.float64 else np.float32
        r = _Mat(R, np.int32, "R")
        G, S = r.shape
        out, m = _out_mat(out, r.kind, (G, G), dt, R, "out")
        self._pre(r.kind)
        fn, name = ((self.lib.mcpg_ap_mi_f64, "mcpg_ap_mi_f64") if dt == np.float64 else
                    (self.lib.mcpg_ap_mi_i32, "mcpg_ap_mi_i32"))
        self._check(fn(self.h, r.kind, r.ptr, G, S, r.ld, m.ptr, m.ld), name)
        return out

    def ap_mi_block(self, R, row0, nrows, col0, ncols, out=None, dtype=np.float32):
        """Rows row0 .. row0+nrows-1, columns col0 .. col0+ncols-1 of ap_mi(R, dtype=dtype)."""
        dt = np.float64 if np.dtype(dtype) == np.float64 else np.float32
        r = _Mat(R, np.int32, "R")
        G, S = r.shape
        _check_block(G, row0, nrows, col0, ncols)
        out, m = _out_mat(out, r.kind, (nrows, ncols), dt, R, "out")
        self._pre(r.kind)
        fn, name = ((self.lib.mcpg_ap_mi_block_f64, "mcpg_ap_mi_block_f64") if dt == np.float64 else
                    (self.lib.mcpg_ap_mi_block_i32, "mcpg_ap_mi_block_i32"))
        self._check(fn(self.h, r.kind, r.ptr, G, S, r.ld, row0, nrows, col0, ncols, m.ptr, max(m.ld, 1)), name)
        return out

    def ap_mi_cells(self, R, px, py, max_leaves=4096):
        """Leaf (nxy, nx, ny) triples per requested pair in summation order; returns
        (list of int32 arrays of shape (n_leaves, 3), float64 MI per pair)."""
        r = _Mat(R, np.int32, "R")
        G, S = r.shape
        px = np.ascontiguousarray(px, dtype=np.int32)
        py = np.ascontiguousarray(py, dtype=np.int32)
        n = len(px)
        cells = np.zeros((n, max_leaves, 3), dtype=np.int32)
        nl = np.zeros(n, dtype=np.int32)
        mi = np.zeros(n, dtype=np.float64)
        self._pre(r.kind)
        self._check(self.lib.mcpg_ap_mi_cells_i32(self.h, r.kind, r.ptr, G, S, r.ld, px.ctypes.data, py.ctypes.data, n,
                                                  max_leaves, cells.ctypes.data, nl.ctypes.data, mi.ctypes.data),
                    "mcpg_ap_mi_cells_i32")
        if (nl > max_leaves).any():
            raise ValueError(f"a pair has more than max_leaves={max_leaves} leaves")
        return [cells[k, : nl[k]].copy() for k in range(n)], mi

    # ------------------------------------------------------------------ max-min / MCP
    def maxmin(self, A, B=None, kmask=None, symmetric=False, out=None):
        dt = np.float64 if str(getattr(A, "dtype", "")).endswith("float64") else np.float32
        a = _Mat(A, dt, "A")
        b = a if B is None else _Mat(B, dt, "B")
        km = None if kmask is None else _Mat(kmask, dt, "kmask")
        M, K = a.shape
        N = b.shape[0]
        if b.shape[1] != K:
            raise ValueError("A and B must have the same number of columns (row . row product)")
        if km is not None and km.shape[0] != K:
            raise ValueError("kmask length must equal K")
        kind = _same_kind(a, b, km)
        out, c = _out_mat(out, kind, (M, N), dt, A, "out")
        fn = self.lib.mcpg_maxmin_f64 if dt == np.float64 else self.lib.mcpg_maxmin_f32
        self._pre(kind)
        self._check(fn(self.h, kind, a.ptr, a.ld, b.ptr, b.ld, km.ptr if km else None, M, N, K, int(symmetric),
                       c.ptr, c.ld), "mcpg_maxmin")
        return out

    def ratio(self, A, Mm, out=None):
        dt = np.float64 if str(getattr(A, "dtype", "")).endswith("float64") else np.float32
        a = _Mat(A, dt, "A")
        m = _Mat(Mm, dt, "M")
        if a.shape != m.shape:
            raise ValueError("ratio operands must have the same shape")
        kind = _same_kind(a, m)
        out, o = _out_mat(out, kind, a.shape, dt, A, "out")
        fn = self.lib.mcpg_ratio_f64 if dt == np.float64 else self.lib.mcpg_ratio_f32
        self._pre(kind)
        self._check(fn(self.h, kind, a.ptr, a.ld, m.ptr, m.ld, a.shape[0], a.shape[1], o.ptr, o.ld), "mcpg_ratio")
        return out

    def elementwise_max(self, A, B, out=None):
        dt = np.float64 if str(getattr(A, "dtype", "")).endswith("float64") else np.float32
        a = _Mat(A, dt, "A")
        b = _Mat(B, dt, "B")
        if a.shape != b.shape:
            raise ValueError("operands must have the same shape")
        kind = _same_kind(a, b)
        out, o = _out_mat(out, kind, a.shape, dt, A, "out")
        fn = self.lib.mcpg_elementwise_max_f64 if dt == np.float64 else self.lib.mcpg_elementwise_max_f32
        self._pre(kind)
        self._check(fn(self.h, kind, a.ptr, a.ld, b.ptr, b.ld, a.shape[0], a.shape[1], o.ptr, o.ld),
                    "mcpg_elementwise_max")
        return out

    def mcp(self, A, k, want_maxmin=False, out=None, maxmin_out=None):
        """MCP_k score: ratio(A, A^(x)k); returns (mcp, maxmin-or-None)."""
        dt = np.float64 if str(getattr(A, "dtype", "")).endswith("float64") else np.float32
        a = _Mat(A, dt, "A")
        N = a.shape[0]
        if a.shape[1] != N:
            raise ValueError("MCP needs a square matrix")
        out, o = _out_mat(out, a.kind, (N, N), dt, A, "out")
        mm = None
        if want_maxmin or maxmin_out is not None:
            maxmin_out, mm = _out_mat(maxmin_out, a.kind, (N, N), dt, A, "maxmin_out")
        fn = self.lib.mcpg_mcp_f64 if dt == np.float64 else self.lib.mcpg_mcp_f32
        self._pre(a.kind)
        self._check(fn(self.h, a.kind, a.ptr, a.ld, N, int(k), mm.ptr if mm else None, mm.ld if mm else 0, o.ptr, o.ld),
                    "mcpg_mcp")
        return out, maxmin_out

    def mcpnet(self, X, mi_method=0, orders=(2, 3, 4), want_mi=True, bins=10, order=3, precision="fp32",
               noise_seed=1296388676612, diagonal=0.0, tfs=None):
        """The mcpnet stage chain in one device-resident call: MI (mi_method 0 B-spline, 1 adaptive
        partitioning, -1 X is the MI) -> diagonal -> TF rows (tfs > 0) -> MCP of each order.
        Returns {"mi": G x G MI, 2: ..., 3: ..., 4: ...} for what was asked."""
        dt = np.float64 if str(getattr(X, "dtype", "")).endswith("float64") else np.float32
        x = _Mat(X, dt, "X")
        G, S = x.shape
        if mi_method < 0 and S != G:
            raise ValueError("an MI input must be square")
        tf = None
        R = G
        if tfs is not None:
            tf = _Mat(tfs, dt, "tfs")
            if tf.shape != (G,):
                raise ValueError("tfs needs one value per gene")
            _same_kind(x, tf)
            vals = tfs.detach().cpu().numpy() if hasattr(tfs, "detach") else np.asarray(tfs)
            R = int(np.count_nonzero(vals > 0))
        res = {}
        args = []
        for key in ("mi", 2, 3, 4):
            want = want_mi if key == "mi" else key in orders
            if want:
                res[key] = _empty_like_kind(x.kind, (G if key == "mi" else R, G), dt, X)
                m = _Mat(res[key], dt, str(key))
                args += [m.ptr, m.ld]
            else:
                args += [None, 0]
        fn = self.lib.mcpg_mcpnet_f64 if dt == np.float64 else self.lib.mcpg_mcpnet_f32
        self._pre(x.kind)
        self._check(fn(self.h, x.kind, x.ptr, G, S if mi_method >= 0 else 0, x.ld, int(mi_method), int(bins),
                       int(order), PRECISIONS[precision], c_uint64(noise_seed), c_double(diagonal),
                       tf.ptr if tf else None, *args), "mcpg_mcpnet")
        return res

    # ------------------------------------------------------------------ ensemble and filters
    def maxmins(self, A, want=(True, True, True)):
        """compute_maxmins: (M1, M2, M3) = (A(x)A, A(x)M1, M1(x)M1); entries not wanted are None."""
        dt = np.float64 if str(getattr(A, "dtype", "")).endswith("float64") else np.float32
        a = _Mat(A, dt, "A")
        N = a.shape[0]
        if a.shape[1] != N:
            raise ValueError("maxmins needs a square matrix")
        outs = [_empty_like_kind(a.kind, (N, N), dt, A) if w else None for w in want]
        ms = [_Mat(o, dt, "out") if o is not None else None for o in outs]
        args = []
        for m in ms:
            args += [m.ptr if m else None, m.ld if m else 0]
        fn = self.lib.mcpg_maxmins_f64 if dt == np.float64 else self.lib.mcpg_maxmins_f32
        self._pre(a.kind)
        self._check(fn(self.h, a.kind, a.ptr, a.ld, N, *args), "mcpg_maxmins")
        return tuple(outs)

    def combo(self, MI, M1, M2, M3, coeffs, out=None):
        """compute_combo: MI / (a MI + b M1 + c M2 + d M3) with the reference's guards."""
        dt = np.float64 if str(getattr(MI, "dtype", "")).endswith("float64") else np.float32
        mi = _Mat(MI, dt, "MI")
        M, N = mi.shape
        cf = np.ascontiguousarray(coeffs, dtype=dt)
        if cf.shape != (4,):
            raise ValueError("coeffs must hold 4 values")
        mats = [None if (x is None or cf[q + 1] == 0) else _Mat(x, dt, f"M{q + 1}") for q, x in enumerate((M1, M2, M3))]
        for m in mats:
            if m is not None and m.shape != (M, N):
                raise ValueError("operands must have the same shape")
        kind = _same_kind(mi, *mats)
        out, o = _out_mat(out, kind, (M, N), dt, MI, "out")
        args = []
        for m in mats:
            args += [m.ptr if m else None, m.ld if m else 0]
        fn = self.lib.mcpg_combo_f64 if dt == np.float64 else self.lib.mcpg_combo_f32
        self._pre(kind)
        self._check(fn(self.h, kind, mi.ptr, mi.ld, *args, M, N, cf.ctypes.data, o.ptr, o.ld), "mcpg_combo")
        return out

    def stouffer(self, X, exclude_ones=True, out=None):
        """compute_stouffer on a square score matrix."""
        dt = np.float64 if str(getattr(X, "dtype", "")).endswith("float64") else np.float32
        x = _Mat(X, dt, "X")
        N = x.shape[0]
        if x.shape[1] != N:
            raise ValueError("Stouffer needs a square matrix")
        out, o = _out_mat(out, x.kind, (N, N), dt, X, "out")
        fn = self.lib.mcpg_stouffer_f64 if dt == np.float64 else self.lib.mcpg_stouffer_f32
        self._pre(x.kind)
        self._check(fn(self.h, x.kind, x.ptr, x.ld, N, int(exclude_ones), o.ptr, o.ld), "mcpg_stouffer")
        return out

    def stouffer_tf(self, X, out=None):
        """compute_stouffer_tf on a TFs x genes score matrix (row and column z-scores)."""
        dt = np.float64 if str(getattr(X, "dtype", "")).endswith("float64") else np.float32
        x = _Mat(X, dt, "X")
        M, N = x.shape
        out, o = _out_mat(out, x.kind, (M, N), dt, X, "out")
        fn = self.lib.mcpg_stouffer_tf_f64 if dt == np.float64 else self.lib.mcpg_stouffer_tf_f32
        self._pre(x.kind)
        self._check(fn(self.h, x.kind, x.ptr, x.ld, M, N, o.ptr, o.ld), "mcpg_stouffer_tf")
        return out

    def dpi(self, A, tolerance=0.1, tfs=None, out=None, want_mask=False):
        """DPI-filtered copy of the square MI matrix A; returns (out, removed count) or, with
        want_mask, (out, removed count, uint8 mask of removed edges)."""
        dt = np.float64 if str(getattr(A, "dtype", "")).endswith("float64") else np.float32
        a = _Mat(A, dt, "A")
        N = a.shape[0]
        if a.shape[1] != N:
            raise ValueError("DPI needs a square matrix")
        tf = None if tfs is None else _Mat(tfs, dt, "tfs")
        if tf is not None and tf.shape != (N,):
            raise ValueError("tfs must have one entry per gene")
        kind = _same_kind(a, tf)
        out, o = _out_mat(out, kind, (N, N), dt, A, "out")
        mask = None
        mk = None
        if want_mask:
            if kind == MCPG_DEVICE:
                import torch
                mask = torch.empty((N, N), dtype=torch.uint8, device=A.device)
                mk_ptr, mk_ld = mask.data_ptr(), N
            else:
                mask = np.empty((N, N), dtype=np.uint8)
                mk_ptr, mk_ld = mask.ctypes.data, N
            mk = (mk_ptr, mk_ld)
        removed = c_int64(0)
        if dt == np.float64:
            fn, tol = self.lib.mcpg_dpi_f64, c_double(tolerance)
        else:
            fn, tol = self.lib.mcpg_dpi_f32, c_float(tolerance)
        self._pre(kind)
        self._check(fn(self.h, kind, a.ptr, a.ld, N, tol, tf.ptr if tf else None, o.ptr, o.ld,
                       mk[0] if mk else None, mk[1] if mk else 0, byref(removed)), "mcpg_dpi")
        return (out, removed.value, mask) if want_mask else (out, removed.value)
