"""Ensemble MCP (compute_maxmins, compute_combo, compute_stouffer, compute_stouffer_tf) and the DPI
filter on the GPU,
against the CPU oracle: bit-exact for the max-min products, the combination ratio and the DPI
decisions; the Stouffer z-scores within a few float ulps (FP64 row moments, reduction order free)."""
import numpy as np
import pytest

import oracle
from mcpnet_amd import grn

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def mi_matrix(G, S, seed, dt=np.float32):
    X = grn.grn_expression(G, S, seed)[0]
    A = oracle.bspline_mi_f64(X).astype(dt)
    np.fill_diagonal(A, 0)
    return A


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("N", [1, 37, 130, 300])
def test_maxmins_bit_exact(handle, dt, N):
    rng = np.random.default_rng(N)
    A = rng.random((N, N)).astype(dt)
    A = np.maximum(A, A.T)
    np.fill_diagonal(A, 0)
    M1, M2, M3 = handle.maxmins(A)
    R1, R2, R3 = oracle.maxmins(A)
    assert bits_equal(M1, R1) and bits_equal(M2, R2) and bits_equal(M3, R3)
    # outputs are optional; an asymmetric A takes the full (non-triangle) second product
    B = rng.random((N, N)).astype(dt)
    _, P2, _ = handle.maxmins(B, want=(False, True, False))
    assert bits_equal(P2, oracle.maxmin(B, oracle.maxmin(B, B)))


def test_maxmins_device_tensors(handle, torch_cuda):
    A = mi_matrix(150, 200, 3)
    Ad = torch_cuda.from_numpy(A).cuda()
    M1, M2, M3 = handle.maxmins(Ad)
    R = oracle.maxmins(A)
    for got, want in zip((M1, M2, M3), R):
        assert bits_equal(got.cpu().numpy(), want)


COEFFS = [[0.25, 0.25, 0.25, 0.25], [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1], [0.5, 0, 0.5, 0],
          [0.1, 0.2, 0.3, 0.4], [1, 1, 1, 1], [0, 0, 0, 0]]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_combo_bit_exact(handle, dt):
    A = mi_matrix(120, 300, 5, dt)
    A[4, 9] = A[9, 4] = 0                                 # |MI| < eps branch
    M1, M2, M3 = oracle.maxmins(A)
    for cf in COEFFS:
        cf = np.array(cf, dt)
        assert bits_equal(handle.combo(A, M1, M2, M3, cf), oracle.combo(A, M1, M2, M3, cf)), cf
    # terms with a zero coefficient may be omitted
    cf = np.array([0.5, 0, 0.5, 0], dt)
    assert bits_equal(handle.combo(A, None, M2, None, cf), oracle.combo(A, M1, M2, M3, cf))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("exclude_ones", [False, True])
def test_stouffer(handle, dt, exclude_ones):
    A = mi_matrix(140, 300, 7, dt)
    M1, M2, M3 = oracle.maxmins(A)
    R = oracle.combo(A, M1, M2, M3, np.array([0, 0, 1, 0], dt))   # an MCP3 score matrix (has exact 1s)
    R[11] = 0.5                                                   # a constant row
    got = handle.stouffer(R, exclude_ones)
    want = oracle.stouffer(R, exclude_ones)
    tol = 2e-6 if dt == np.float32 else 1e-12
    assert np.allclose(got, want, rtol=tol, atol=tol)
    assert np.all(np.isfinite(got))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 5), (23, 140), (64, 1001), (130, 64)])
def test_stouffer_tf(handle, dt, shape):
    M, N = shape
    rng = np.random.default_rng(M * 7 + N)
    R = (rng.random((M, N)) * 3).astype(dt)
    if M > 3:
        R[3] = 0.75                                   # a constant TF row
    if N > 17:
        R[:, 17] = 1.25                               # a constant gene column
    got = handle.stouffer_tf(R)
    want = oracle.stouffer_tf(R)
    tol = 2e-6 if dt == np.float32 else 1e-12
    assert np.allclose(got, want, rtol=tol, atol=tol)
    assert np.all(np.isfinite(got))


def test_stouffer_tf_device(handle, torch_cuda):
    torch = torch_cuda
    R = (np.random.default_rng(5).random((40, 333)) * 2).astype(np.float32)
    Rd = torch.from_numpy(R).cuda()
    got = handle.stouffer_tf(Rd).cpu().numpy()
    assert np.allclose(got, oracle.stouffer_tf(R), rtol=2e-6, atol=2e-6)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("with_tfs", [False, True])
@pytest.mark.parametrize("tol", [0.0, 0.1])
def test_dpi_bit_exact(handle, dt, with_tfs, tol):
    A = mi_matrix(200, 300, 9, dt)
    rng = np.random.default_rng(3)
    tfs = np.where(rng.random(200) < 0.25, np.finfo(dt).max, np.finfo(dt).min).astype(dt) if with_tfs else None
    got, n = handle.dpi(A, tol, tfs)
    want, m = oracle.dpi(A, tol, tfs)
    assert bits_equal(got, want) and n == m
    assert m > 0


@pytest.mark.parametrize("with_tfs", [False, True])
@pytest.mark.parametrize("N", [2, 3, 200])
def test_dpi_f64_float_exact_on_keys(handle, monkeypatch, with_tfs, N):
    """A float-exact FP64 MI matrix (the FP32 MI widened, as the double tools hand it on): the
    DPI products run on the key kernel with widened outputs and give the FP64 kernel's and the
    oracle's bits; N = 2 leaves every triangle empty (the products stay at lowest<double>)."""
    A = mi_matrix(N, 300, 11).astype(np.float64)
    A[0, 0] = 0.1                                    # the diagonal is excluded: not tested for exactness
    rng = np.random.default_rng(N)
    tfs = np.where(rng.random(N) < 0.3, np.finfo(np.float64).max, np.finfo(np.float64).min) if with_tfs else None
    if with_tfs:
        tfs[0] = np.finfo(np.float64).max
    handle.set_profiling(True)
    try:
        got, n = handle.dpi(A, 0.1, tfs)
        names = [k for k, _ in handle.last_timings()]
    finally:
        handle.set_profiling(False)
    assert "dpi_key_kernel" in names and not any(k.startswith("maxmin_f64_kernel") for k in names)
    monkeypatch.setenv("MCPG_MM64_KEYS", "0")
    ref, r = handle.dpi(A, 0.1, tfs)
    want, m = oracle.dpi(A, 0.1, tfs)
    assert bits_equal(got, ref) and n == r
    assert bits_equal(got, want) and n == m


def test_dpi_asymmetric_and_device(handle, torch_cuda):
    rng = np.random.default_rng(4)
    A = (rng.random((257, 257)) ** 3).astype(np.float32)     # asymmetric: i_xz uses row x
    got, n = handle.dpi(torch_cuda.from_numpy(A).cuda(), 0.05)
    want, m = oracle.dpi(A, 0.05)
    assert bits_equal(got.cpu().numpy(), want) and n == m


def test_dpi_mask_output(handle):
    A = mi_matrix(90, 200, 10)
    A[2, 7] = A[7, 2] = 0                                    # removed edge that is already 0: only the mask shows it
    got, n, mask = handle.dpi(A, 0.1, want_mask=True)
    want, m = oracle.dpi(A, 0.1)
    assert bits_equal(got, want) and n == m
    assert np.array_equal(mask.astype(bool) & (A != 0), want != A)
    assert mask.diagonal().all()
