"""Adjoint of the B-spline MI: GPU dL/dX against central finite differences of the FP64
oracle loss (h = 1e-3) and against the FP64 analytic oracle gradient."""
import numpy as np
import pytest

import oracle
from mcpnet_amd import grn

pytestmark = pytest.mark.gpu


def setup(G, S, prob, seed):
    X = grn.grn_expression(G, S, seed)[0]
    pi, pj, m = grn.mask_pairs(G, prob, seed + 1)
    return X, pi, pj, m


def test_loss_matches_oracle(handle):
    X, pi, pj, m = setup(60, 200, 0.2, 3)
    dX, L = handle.bspline_mi_backward(X, pi, pj, m)
    Lref = oracle.bspline_mi_loss_f64(X, pi, pj, m)
    assert abs(L - Lref) <= 1e-5 * max(1.0, abs(Lref))


def test_gradient_vs_fp64_oracle(handle):
    X, pi, pj, m = setup(50, 300, 0.3, 5)
    dX, _ = handle.bspline_mi_backward(X, pi, pj, m)
    ref = oracle.bspline_mi_grad_f64(X, pi, pj, m)
    rel = np.linalg.norm(dX - ref) / np.linalg.norm(ref)
    assert rel <= 1e-3, rel


def test_valu_kernels_match_mfma_kernels(handle, monkeypatch):
    """MCPG_ADJ_MFMA=0 selects the VALU pair-table and gather kernels: same loss and gradient
    as the MFMA kernels up to FP32 summation order."""
    X, pi, pj, m = setup(70, 260, 0.2, 9)
    dX1, L1 = handle.bspline_mi_backward(X, pi, pj, m)
    monkeypatch.setenv("MCPG_ADJ_MFMA", "0")
    dX0, L0 = handle.bspline_mi_backward(X, pi, pj, m)
    assert abs(L1 - L0) <= 1e-6 * max(1.0, abs(L0))
    assert np.linalg.norm(dX1 - dX0) / np.linalg.norm(dX0) <= 1e-5


def test_gradient_vs_finite_differences(handle):
    X, pi, pj, m = setup(40, 200, 0.3, 7)
    dX, _ = handle.bspline_mi_backward(X, pi, pj, m)
    rng = np.random.default_rng(0)
    X64 = X.astype(np.float64)
    h = 1e-3
    errs = []
    scale = np.sqrt(np.mean(dX.astype(np.float64) ** 2))
    for _ in range(200):
        g, s = rng.integers(0, X.shape[0]), rng.integers(0, X.shape[1])
        Xp = X64.copy(); Xp[g, s] += h
        Xm = X64.copy(); Xm[g, s] -= h
        fd = (oracle.bspline_mi_loss_x64(Xp, pi, pj, m) - oracle.bspline_mi_loss_x64(Xm, pi, pj, m)) / (2 * h)
        errs.append(abs(fd - dX[g, s]) / max(abs(fd), scale))
    errs = np.array(errs)
    print(f"adjoint vs FD: max rel {errs.max():.2e} median {np.median(errs):.2e}")
    assert np.median(errs) <= 1e-3
    assert np.quantile(errs, 0.95) <= 1e-2


def test_autograd_wrapper(handle, torch_cuda):
    import torch
    from mcpnet_amd.autograd import masked_mi_loss
    X, pi, pj, m = setup(30, 150, 0.3, 11)
    Xt = torch.from_numpy(X).cuda().requires_grad_(True)
    L = masked_mi_loss(handle, Xt, pi, pj, m)
    L.backward()
    ref = oracle.bspline_mi_grad_f64(X, pi, pj, m)
    g = Xt.grad.cpu().numpy()
    assert np.linalg.norm(g - ref) / np.linalg.norm(ref) <= 1e-3
    assert abs(L.item() - oracle.bspline_mi_loss_f64(X, pi, pj, m)) <= 1e-4
