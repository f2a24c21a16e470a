"""B-spline MI in double on the GPU (the reference's default double DataType): scaling bit-exact
against numpy's double formula, weights bit-exact against the double oracle (same basis
operations), marginal entropies and MI within 1e-12 of the oracle (only the summation order of
the dot products and sums differs)."""
import numpy as np
import pytest

import mcpnet_amd
import oracle
from mcpnet_amd import grn

pytestmark = pytest.mark.gpu

TOL = 1e-12   # |dMI| and |dH| against the double oracle


def double_rows(G, S, seed):
    """Expression rows that no float holds exactly (the double tools read them from files)."""
    rng = np.random.default_rng(seed)
    X = grn.grn_expression(G, S, seed)[0].astype(np.float64)
    return X * (1.0 + 1e-9 * rng.standard_normal((G, S)))


def test_minmax_scale_bit_exact(handle):
    X = double_rows(20, 333, 1)
    X[3] = 2.5                                           # constant row: kept (BSplineMI.hpp:102-103)
    Y = handle.minmax_scale(X)
    mn, mx = X.min(1, keepdims=True), X.max(1, keepdims=True)
    ref = np.where(mx > mn, (1.0 / np.where(mx > mn, mx - mn, 1.0)) * (X - mn), X)
    assert Y.dtype == np.float64 and np.array_equal(Y, ref)


@pytest.mark.parametrize("bins,order", [(10, 3), (7, 2), (16, 4), (5, 1), (30, 5)])
def test_weights_match_oracle(handle, bins, order):
    X = double_rows(9, 257, bins)
    Xs = handle.minmax_scale(X)
    W = handle.bspline_weights(Xs, bins=bins, order=order)
    Wo, Ho = oracle.bspline_weights_x64(Xs, bins, order, scale=False)
    S = X.shape[1]
    assert np.array_equal(W[:, :bins * S].reshape(9, bins, S), Wo)
    assert np.abs(W[:, bins * S] - Ho).max() <= TOL


@pytest.mark.parametrize("G,S", [(1, 1), (2, 5), (17, 100), (160, 301), (333, 1024)])
def test_mi_matches_oracle(handle, G, S):
    X = double_rows(G, S, G + S)
    got = handle.bspline_mi(X, precision="fp64")
    ref = oracle.bspline_mi_x64(X)
    assert got.dtype == np.float64
    assert np.abs(got - ref).max() <= TOL
    assert np.array_equal(got, got.T)


@pytest.mark.parametrize("bins,order", [(7, 2), (16, 4), (12, 3), (30, 5)])
def test_mi_other_bins(handle, bins, order):
    X = double_rows(45, 200, bins * 10 + order)
    got = handle.bspline_mi(X, bins=bins, order=order, precision="fp64")
    assert np.abs(got - oracle.bspline_mi_x64(X, bins, order)).max() <= TOL


def test_constant_gene_and_float_rows(handle):
    X = grn.grn_expression(40, 150, 9)[0].astype(np.float64)   # float-exact rows: the FP64 oracle
    X[7] = 0.25
    got = handle.bspline_mi(X, precision="fp64")
    assert np.abs(got - oracle.bspline_mi_f64(X.astype(np.float32))).max() <= TOL


@pytest.mark.parametrize("env", [{"MCPG_MI64_BUF": "2"}, {"MCPG_MI64_BUF": "4"}])
def test_kernel_forms_bit_identical(handle, monkeypatch, env):
    """Register rings of 3 K blocks (default), 2 or 4 run the same MFMA sequence per output:
    bit-identical MI, including sample counts with one and with two K blocks."""
    for G, S in ((200, 777), (35, 8), (50, 3), (40, 13)):
        X = double_rows(G, S, G * S)
        ref = handle.bspline_mi(X, precision="fp64")
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        got = handle.bspline_mi(X, precision="fp64")
        for k in env:
            monkeypatch.delenv(k)
        assert np.array_equal(got, ref), (G, S)


def test_mi_from_weights_equals_fused(handle):
    X = double_rows(50, 180, 3)
    W = handle.bspline_weights(handle.minmax_scale(X))
    assert np.array_equal(handle.bspline_mi_weights(W, 180, precision="fp64"), handle.bspline_mi(X, precision="fp64"))


def test_device_tensors(handle, torch_cuda):
    X = double_rows(70, 260, 4)
    host = handle.bspline_mi(X, precision="fp64")
    dev = handle.bspline_mi(torch_cuda.from_numpy(X).cuda(), precision="fp64")
    assert np.array_equal(dev.cpu().numpy(), host)


def test_non_finite_rejected(handle):
    X = double_rows(6, 50, 5)
    X[2, 7] = np.nan
    with pytest.raises(mcpnet_amd.MCPGError) as e:
        handle.bspline_mi(X, precision="fp64")
    assert e.value.status == -3


def test_fp32_entry_points_reject_fp64(handle):
    X = grn.grn_expression(6, 50, 5)[0]
    out = np.empty((6, 6), np.float32)
    st = handle.lib.mcpg_bspline_mi_f32(handle.h, 0, X.ctypes.data, 6, 50, 50, 10, 3, 4, out.ctypes.data, 6)
    assert st == -1


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_mcpnet_chain_fp64(handle, dt):
    """The mcpnet chain with precision fp64: the double MI (rounded to float for the f32 chain)
    with the diagonal set, and the MCP scores of that MI."""
    X = double_rows(90, 210, 8).astype(dt)
    res = handle.mcpnet(X, mi_method=0, orders=(2, 3), precision="fp64")
    MI = handle.bspline_mi(X.astype(np.float64), precision="fp64").astype(dt)
    np.fill_diagonal(MI, 0.0)
    assert res["mi"].dtype == dt and np.array_equal(res["mi"], MI)
    for k in (2, 3):
        assert np.array_equal(res[k], handle.mcp(MI, k)[0]), k


@pytest.mark.slow
def test_mi_2000_genes_1024_samples(handle, torch_cuda):
    """2000 x 1024 double rows: every pair within 1e-12 of the double oracle."""
    X = double_rows(2000, 1024, 12)
    got = handle.bspline_mi(torch_cuda.from_numpy(X).cuda(), precision="fp64").cpu().numpy()
    err = np.abs(got - oracle.bspline_mi_x64(X))
    print(f"2000 x 1024 FP64 MI: max |dMI| vs the double oracle {err.max():.3e}")
    assert err.max() <= TOL
