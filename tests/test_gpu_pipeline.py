"""mcpg_mcpnet_*: the device-resident mcpnet stage chain (MI -> diagonal -> TF rows -> MCP2/3/4)
must give exactly what the stage entry points give one by one, and, for the TF orders, what the
oracle's masked max-min compositions give (compute_mcp{2,3,4}_tfs, mcpnet_helper.hpp:208-435)."""
import numpy as np
import pytest

import oracle
from mcpnet_amd import MCPGError, grn

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "i8x3"])
def test_bspline_chain_matches_stages(handle, precision):
    X = grn.grn_expression(150, 320, 31)[0]
    res = handle.mcpnet(X, mi_method=0, precision=precision)
    MI = handle.bspline_mi(X, precision=precision)
    np.fill_diagonal(MI, 0.0)
    assert bits_equal(res["mi"], MI)
    for k in (2, 3, 4):
        assert bits_equal(res[k], handle.mcp(MI, k)[0]), k


@pytest.mark.parametrize("orders", [(2, 3), (3, 4), (2, 4), (4,), (2, 3, 4)])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_order_subsets_share_maxmin(handle, torch_cuda, orders, dt):
    """Two or more orders share A (x) A (and M1 (x) M1, A (x) M1 as asked); every subset, with host
    and device buffers, gives the per-order entry points' bits; each order's host download
    overlaps the next order."""
    X = grn.grn_expression(130, 200, 36)[0].astype(dt)
    MI = handle.bspline_mi(X.astype(np.float32), precision="bf16x3").astype(dt)
    np.fill_diagonal(MI, 0.0)
    want = {k: handle.mcp(MI, k)[0] for k in orders}
    host = handle.mcpnet(X, mi_method=0, orders=orders, precision="bf16x3")
    dev = handle.mcpnet(torch_cuda.from_numpy(X).cuda(), mi_method=0, orders=orders, precision="bf16x3")
    for k in orders:
        assert bits_equal(host[k], want[k]), k
        assert bits_equal(dev[k].cpu().numpy(), want[k]), k
    assert set(host) == {"mi", *orders}


@pytest.mark.slow
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_host_chain_with_overlapped_mi_download(handle, dt):
    """3000 genes with host buffers: the MI (36 or 72 MB) comes back on the side stream in two or
    three 32 MB chunks while the max-min products run; every output equals the stage entry
    points' bits."""
    X = grn.grn_expression(3000, 96, 34)[0].astype(dt)
    res = handle.mcpnet(X, mi_method=0, orders=(2, 3), precision="bf16x3")
    MI = handle.bspline_mi(X.astype(np.float32), precision="bf16x3").astype(dt)
    np.fill_diagonal(MI, 0.0)
    assert bits_equal(res["mi"], MI)
    for k in (2, 3):
        assert bits_equal(res[k], handle.mcp(MI, k)[0]), k


def test_ap_chain_matches_stages(handle):
    X = grn.grn_expression(70, 256, 32)[0]
    res = handle.mcpnet(X, mi_method=1, noise_seed=99, orders=(3,), diagonal=-1.0)
    MI = handle.ap_mi(handle.rank(handle.add_noise(X, 99)))   # diagonal kept (negative --diagonal)
    assert bits_equal(res["mi"], MI)
    assert bits_equal(res[3], handle.mcp(MI, 3)[0])
    assert 2 not in res and 4 not in res


def test_double_ap_chain_keeps_the_mi_in_double(handle):
    X = grn.grn_expression(50, 256, 31)[0]
    res = handle.mcpnet(X.astype(np.float64), mi_method=1, noise_seed=99, orders=(3,), diagonal=-1.0)
    R = handle.rank(handle.add_noise(X, 99))
    MI = handle.ap_mi(R, dtype=np.float64)
    assert res["mi"].dtype == np.float64 and bits_equal(res["mi"], MI)
    assert bits_equal(res[3], handle.mcp(MI, 3)[0])


@pytest.mark.parametrize("G,orders,env", [
    (4000, (2, 3, 4), {}),                                            # panels of 16, 8, 8 row blocks
    (600, (2, 3), {"MCPG_MM_GROUP": "1"}),                            # panels of 3, 1, 1 row blocks
    (600, (3, 4), {"MCPG_MM_GROUP": "1", "MCPG_MM_PANELS": "0"}),     # the last product whole
])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_last_product_in_row_panels(handle, torch_cuda, monkeypatch, G, orders, env, dt):
    """The shared chain's last triangle product runs in row panels, each panel's scores formed and
    (host buffers) sent down while the next panel runs; host and device buffers give the per-order
    entry points' bits, including the partial last row block."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    X = grn.grn_expression(G, 64, 38)[0].astype(dt)
    host = handle.mcpnet(X, mi_method=0, orders=orders, precision="bf16x3")
    dev = handle.mcpnet(torch_cuda.from_numpy(X).cuda(), mi_method=0, orders=orders, precision="bf16x3")
    MI = handle.bspline_mi(X.astype(np.float32), precision="bf16x3").astype(dt)
    np.fill_diagonal(MI, 0.0)
    assert bits_equal(host["mi"], MI)
    for k in orders:
        want = handle.mcp(MI, k)[0]
        assert bits_equal(host[k], want), k
        assert bits_equal(dev[k].cpu().numpy(), want), k


@pytest.mark.parametrize("env", [{}, {"MCPG_MM64_TILE": "1"}])
def test_double_ap_chain_shares_fp64_maxmin(handle, torch_cuda, monkeypatch, env):
    """The double AP MI is not float-exact, so the shared chain of MCP2/3/4 runs the FP64 max-min
    kernels (padded values, or the register-staged tile kernel) with each order formed after its
    product; host and device buffers give the per-order entry points' bits."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    X = grn.grn_expression(70, 256, 37)[0].astype(np.float64)
    host = handle.mcpnet(X, mi_method=1, noise_seed=5, orders=(2, 3, 4))
    dev = handle.mcpnet(torch_cuda.from_numpy(X).cuda(), mi_method=1, noise_seed=5, orders=(2, 3, 4))
    MI = host["mi"]
    assert not np.array_equal(MI.astype(np.float32).astype(np.float64), MI)
    assert bits_equal(dev["mi"].cpu().numpy(), MI)
    for k in (2, 3, 4):
        want = handle.mcp(MI, k)[0]
        assert bits_equal(host[k], want), k
        assert bits_equal(dev[k].cpu().numpy(), want), k


def test_double_chain_widens_the_mi(handle):
    X = grn.grn_expression(60, 200, 33)[0].astype(np.float64)
    res = handle.mcpnet(X, mi_method=0, diagonal=0.5)
    MI = handle.bspline_mi(X.astype(np.float32)).astype(np.float64)
    np.fill_diagonal(MI, 0.5)
    assert res["mi"].dtype == np.float64 and bits_equal(res["mi"], MI)
    for k in (2, 3, 4):
        assert bits_equal(res[k], oracle.mcp(MI, k)[0]), k


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_tf_orders_match_oracle(handle, dt):
    X = grn.grn_expression(48, 256, 34)[0]
    A = oracle.bspline_mi_f64(X).astype(dt)
    tf_idx = [1, 4, 9, 22, 40]
    with np.errstate(over="ignore"):                          # as mcpnet's load_tfs: double limits in T
        tfs = np.full(48, np.finfo(np.float64).min).astype(dt)
        tfs[tf_idx] = np.finfo(np.float64).max
    res = handle.mcpnet(A, mi_method=-1, tfs=tfs)
    Ad = A.copy()
    np.fill_diagonal(Ad, 0)
    T = np.ascontiguousarray(Ad[tf_idx])
    assert bits_equal(res["mi"], Ad)
    mm = oracle.maxmin
    G1 = mm(Ad, Ad, tfs)
    want = {2: mm(T, Ad, tfs), 3: mm(T, G1, tfs), 4: mm(mm(T, Ad, tfs), G1, tfs)}
    for k in (2, 3, 4):
        assert bits_equal(res[k], oracle.ratio(T, want[k])), k


def test_device_tensors_match_host(handle, torch_cuda):
    torch = torch_cuda
    X = grn.grn_expression(90, 300, 35)[0]
    tfs = np.full(90, -np.inf, np.float32)
    tfs[[0, 5, 17, 60]] = np.inf
    host = handle.mcpnet(X, mi_method=0, tfs=tfs)
    dev = handle.mcpnet(torch.from_numpy(X).cuda(), mi_method=0, tfs=torch.from_numpy(tfs).cuda())
    for key in ("mi", 2, 3, 4):
        assert bits_equal(dev[key].cpu().numpy(), host[key]), key


def test_errors(handle):
    A = np.random.default_rng(1).random((8, 8)).astype(np.float32)
    with pytest.raises(MCPGError):
        handle.mcpnet(A, mi_method=-1, tfs=np.full(8, -1.0, np.float32))   # no TF selected
    with pytest.raises(MCPGError):
        handle.mcpnet(A, mi_method=2)
