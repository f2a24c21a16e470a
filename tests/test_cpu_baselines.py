"""Pins of the timed CPU baselines: the FP32 port of the reference's B-spline inner product
(oracle/cpu_port.c, built with the reference's fast-math flags) and the row-block AP-MI
restatement, against the strict FP64 / full-matrix oracle on the same inputs."""
import ctypes

import numpy as np
import pytest

import oracle
from mcpnet_amd import grn


def _port_libs():
    libs = [("x86-64-v3", oracle.lib())]
    if oracle._has_avx512() and oracle.LIB_AVX512.exists():
        L = ctypes.CDLL(str(oracle.LIB_AVX512))
        L.port_bspline_mi_rows_f32.restype = None
        L.port_bspline_mi_rows_f32.argtypes = oracle._SIGS["port_bspline_mi_rows_f32"][1]
        libs.append(("x86-64-v4", L))
    return libs


@pytest.mark.parametrize("G,S", [(60, 500), (33, 1000)])
def test_port_bspline_mi_rows_vs_fp64(G, S):
    X = grn.grn_expression(G, S, 21 + G)[0]
    W = oracle.bspline_weights_f32(oracle.minmax_scale_f32(X))
    ref = oracle.bspline_mi_f64(X)
    for name, L in _port_libs():
        for row0, nrows in ((0, G), (G // 3, 5)):
            out = np.empty((nrows, G), dtype=np.float32)
            L.port_bspline_mi_rows_f32(np.ascontiguousarray(W).ctypes.data, G, S, 10, row0, nrows, out.ctypes.data)
            err = np.abs(out.astype(np.float64) - ref[row0:row0 + nrows]).max()
            assert err <= 1e-5, (name, err)


def test_port_wrapper_selects_a_working_build():
    X = grn.grn_expression(20, 300, 5)[0]
    W = oracle.bspline_weights_f32(oracle.minmax_scale_f32(X))
    got = oracle.port_bspline_mi_rows_f32(W, 300, 10, 2, 4)
    assert np.abs(got.astype(np.float64) - oracle.bspline_mi_f64(X)[2:6]).max() <= 1e-5


def test_ap_mi_rows_equal_full_matrix_rows():
    X = grn.grn_expression(50, 300, 8)[0]
    R = grn.rank_desc(grn.jitter(X, 3))
    full = oracle.ap_mi(R)
    for row0, nrows in ((0, 50), (7, 3), (49, 1)):
        rows = oracle.ap_mi_rows(R, row0, nrows)
        assert np.array_equal(rows.view(np.uint32), full[row0:row0 + nrows].view(np.uint32))
