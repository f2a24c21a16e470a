"""Pins of the CPU oracle (no GPU).  Each check comes from outside the oracle's own formulas:
closed forms, an independent library (scipy's B-spline basis), brute force, the reference's
own max_of_pairmin (oracle/_ref, built from include/mcp/filter/simd_maxmin.hpp) and its golden
vectors, and finite differences for the gradient."""
from pathlib import Path

import numpy as np
import pytest
from scipy.interpolate import BSpline

import oracle
from mcpnet_amd import grn

GOLD = np.load(Path(__file__).with_name("golden") / "maxmin_ref.npz")


# ------------------------------------------------------------------ knots and basis
@pytest.mark.parametrize("bins,order", [(10, 3), (7, 2), (16, 4), (5, 1), (4, 4)])
def test_knots_closed_form(bins, order):
    k = oracle.knots(bins, order, np.float64)
    n_int = bins - order
    ref = np.concatenate([np.zeros(order), np.arange(1, n_int + 1) / (n_int + 1), np.ones(order)])
    np.testing.assert_allclose(k, ref, rtol=4.5e-16, atol=0)   # the reference multiplies by 1/(n+1)
    if (n_int + 1) & n_int == 0:                              # dyadic spacing: exact
        np.testing.assert_array_equal(k, ref)
    k32 = oracle.knots(bins, order, np.float32)
    np.testing.assert_allclose(k32, ref.astype(np.float32), rtol=1.2e-7, atol=0)


def scipy_basis(t, bins, order):
    k = oracle.knots(bins, order, np.float64)
    out = np.zeros((len(t), bins))
    for b in range(bins):
        c = np.zeros(bins)
        c[b] = 1.0
        out[:, b] = BSpline(k, c, order - 1, extrapolate=True)(t)
    return out


@pytest.mark.parametrize("bins,order", [(10, 3), (7, 2), (16, 4), (12, 3)])
def test_basis_matches_scipy(bins, order):
    t = np.concatenate([np.linspace(0, 1, 257), np.random.default_rng(0).random(200)])
    ref = scipy_basis(t, bins, order)
    b64 = oracle.basis_f64(t, bins, order)
    np.testing.assert_allclose(b64, ref, rtol=0, atol=1e-14)
    b32 = oracle.basis_f32(t.astype(np.float32), bins, order)
    np.testing.assert_allclose(b32, ref, rtol=0, atol=2e-6)
    np.testing.assert_allclose(b64.sum(axis=1), 1.0, atol=1e-14)      # partition of unity
    assert np.all((b64 > 0).sum(axis=1) <= order)                     # local support


def test_basis_right_end_rule():
    # N_{bins-1,1}(1) = 1 (BSplineMI.hpp:133-136): the last basis function is 1 at t = 1
    b = oracle.basis_f32(np.array([1.0], np.float32))
    assert b[0, -1] == 1.0 and b[0, :-1].sum() == 0.0


def test_uniform_quadratic_interior_pieces():
    # de Boor's uniform quadratic pieces on an interior span: (1-u)^2/2, (-2u^2+2u+1)/2, u^2/2
    u = np.linspace(0, 0.99, 50)
    t = (3 + u) / 8.0                     # span [3/8, 4/8): bins 3, 4, 5 are non-zero
    B = oracle.basis_f64(t)
    np.testing.assert_allclose(B[:, 3], (1 - u) ** 2 / 2, atol=1e-14)
    np.testing.assert_allclose(B[:, 4], (-2 * u ** 2 + 2 * u + 1) / 2, atol=1e-14)
    np.testing.assert_allclose(B[:, 5], u ** 2 / 2, atol=1e-14)


def test_marginal_entropy_closed_form():
    # evenly spread samples: p_b -> integral of N_b = (k[b+order] - k[b]) / order
    S = 200000
    x = ((np.arange(S) + 0.5) / S).astype(np.float32)[None, :]
    _, H = oracle.bspline_weights_f64(x, scale=False)
    k = oracle.knots(10, 3, np.float64)
    p = (k[3:13] - k[0:10]) / 3.0
    assert abs(H[0] - float(-(p * np.log2(p)).sum())) < 1e-6


def test_minmax_scale_quirks():
    X = np.array([[2.0, 4.0, 3.0], [5.0, 5.0, 5.0]], np.float32)
    Y = oracle.minmax_scale_f32(X)
    np.testing.assert_array_equal(Y[0], np.float32([0.0, 1.0, 0.5]))
    np.testing.assert_array_equal(Y[1], X[1])          # constant row kept (BSplineMI.hpp:102-103)


# ------------------------------------------------------------------ MI
def test_mi_independent_evaluation_from_scipy_bases():
    X = grn.grn_expression(12, 300, 3)[0]
    MI = oracle.bspline_mi_f64(X)
    Xs = (X - X.min(1, keepdims=True)) / (X.max(1, keepdims=True) - X.min(1, keepdims=True))
    W = [scipy_basis(Xs[g].astype(np.float64), 10, 3) for g in range(12)]      # S x B each
    S = X.shape[1]

    def H(p):
        p = p[p > 0]
        return -(p * np.log2(p)).sum()

    for i in range(12):
        for j in range(12):
            ref = H(W[i].mean(0)) + H(W[j].mean(0)) - H((W[i].T @ W[j]) / S)
            assert abs(MI[i, j] - ref) < 1e-9


def test_mi_x64_independent_evaluation_from_scipy_bases():
    """The double-row oracle (the reference's double DataType) against numpy + scipy in double, on
    rows that no float holds exactly; on float-exact rows it is the float-row oracle, bitwise."""
    rng = np.random.default_rng(5)
    X = grn.grn_expression(10, 300, 5)[0].astype(np.float64) * (1.0 + 1e-9 * rng.standard_normal((10, 300)))
    assert not np.array_equal(X.astype(np.float32).astype(np.float64), X)
    MI = oracle.bspline_mi_x64(X)
    mn, mx = X.min(1, keepdims=True), X.max(1, keepdims=True)
    Xs = (1.0 / (mx - mn)) * (X - mn)
    W = [scipy_basis(Xs[g], 10, 3) for g in range(10)]
    S = X.shape[1]

    def H(p):
        p = p[p > 0]
        return -(p * np.log2(p)).sum()

    for i in range(10):
        for j in range(10):
            ref = H(W[i].mean(0)) + H(W[j].mean(0)) - H((W[i].T @ W[j]) / S)
            assert abs(MI[i, j] - ref) < 1e-12
    Wx, Hx = oracle.bspline_weights_x64(X)
    np.testing.assert_allclose(Wx[4].T, W[4], rtol=0, atol=1e-14)
    Xf = grn.grn_expression(8, 200, 6)[0]
    assert np.array_equal(oracle.bspline_mi_x64(Xf.astype(np.float64)), oracle.bspline_mi_f64(Xf))


def test_mi_invariances():
    X = grn.grn_expression(10, 400, 4)[0]
    MI = oracle.bspline_mi_f64(X)
    assert np.array_equal(MI, MI.T)
    perm = np.random.default_rng(1).permutation(400)
    np.testing.assert_allclose(oracle.bspline_mi_f64(np.ascontiguousarray(X[:, perm])), MI, atol=1e-12)
    X2 = X.copy()
    X2[3] = 7.0 * X2[3] + 1.5
    np.testing.assert_allclose(oracle.bspline_mi_f64(X2), MI, atol=1e-5)
    assert np.all(MI >= -1e-12)
    assert np.all(np.diag(MI) >= MI.max(axis=1) - 1e-12)   # MI(x,x) = H(x) bounds MI(x, .)


def test_mi_pairs_matches_full():
    X = grn.grn_expression(15, 200, 5)[0]
    MI = oracle.bspline_mi_f64(X)
    iu, ju = np.triu_indices(15)
    np.testing.assert_array_equal(oracle.bspline_mi_pairs_f64(X, iu, ju), MI[iu, ju])


# ------------------------------------------------------------------ AP-MI
@pytest.mark.parametrize("S", [8, 32, 128, 4096])
def test_ap_identical_ranks_closed_form(S):
    R = np.tile(np.arange(S, dtype=np.int32), (2, 1))
    cells, v = oracle.ap_mi_cells(R, 0, 1)
    assert v == np.log2(S) - 2
    assert cells[:, 0].sum() == S


def ap_python(R, x, y, work=None):
    """Independent recursive restatement of AdaptivePartitionRankMIKernel2 (depth-first, children
    in order 0..3 visited last-first), returning leaves in summation order and the MI.  work
    (a one-element list) accumulates the points a reference scan compares: ny for every quadrant
    count of a non-degenerate window with nx, ny > 0 (AdaptivePartitioningMI.hpp:495-523)."""
    S = R.shape[1]
    cnt = np.zeros((R.shape[0], S + 1), np.int64)
    for g in range(R.shape[0]):
        for r in R[g]:
            cnt[g, r + 1] += 1
    cH = np.cumsum(cnt, axis=1)
    pts = list(zip(R[x].tolist(), R[y].tolist()))
    leaves = []

    def count(x1, y1, x2, y2):
        if work is not None and x1 < x2 and y1 < y2:
            nx, ny = int(cH[x, x2] - cH[x, x1]), int(cH[y, y2] - cH[y, y1])
            if nx and ny:
                work[0] += ny
        return sum(1 for a, b in pts if x1 <= a < x2 and y1 <= b < y2)

    def term(w):
        x1, y1, x2, y2, n = w
        return (n, int(cH[x, x2] - cH[x, x1]), int(cH[y, y2] - cH[y, y1]))

    stack = [(0, 0, S, S, S, True)]
    while stack:
        x1, y1, x2, y2, n, whole = stack[-1]
        mx, my = (x1 + x2) >> 1, (y1 + y2) >> 1
        q = [(x1, y1, mx, my), (mx, y1, x2, my), (x1, my, mx, y2), (mx, my, x2, y2)]
        q = [(a, b, c, d, count(a, b, c, d)) for a, b, c, d in q]
        e = n / 4.0
        chi = sum((c[4] - e) ** 2 for c in q) / e if e else float("nan")
        if chi > 7.815 or whole:
            stack.pop()
            for c in q:
                if ((c[0] + 1 < c[2]) or (c[1] + 1 < c[3])) and c[4] > 2:
                    stack.append(c + (False,))
                elif c[4] > 0:
                    leaves.append(term(c))
        else:
            leaves.append(term((x1, y1, x2, y2, n)))
            stack.pop()
    mi = 0.0
    for n, a, b in leaves:
        mi += n * np.log2(n / (a * b))
    return np.array(leaves, dtype=np.int64).reshape(-1, 3), mi / S + np.log2(S)


@pytest.mark.parametrize("seed", range(6))
def test_ap_matches_independent_restatement(seed):
    rng = np.random.default_rng(seed)
    S = int(rng.integers(5, 64))
    if seed % 2:
        R = rng.integers(0, S, (3, S)).astype(np.int32)       # ties
    else:
        R = np.stack([rng.permutation(S) for _ in range(3)]).astype(np.int32)
        R[1] = np.argsort(R[0] + rng.integers(0, 6, S)).astype(np.int32)   # dependent pair
    for x, y in [(0, 1), (1, 2), (0, 0)]:
        cells, v = oracle.ap_mi_cells(R, x, y)
        ref_cells, ref_v = ap_python(R, x, y)
        np.testing.assert_array_equal(cells, ref_cells)
        assert abs(v - ref_v) < 1e-12


def test_ap_work_counter_matches_restatement():
    rng = np.random.default_rng(17)
    S = 48
    R = np.stack([rng.permutation(S) for _ in range(5)]).astype(np.int32)
    R[3] = np.argsort(R[1] + rng.integers(0, 5, S)).astype(np.int32)
    want = 0
    for i in (1, 2):
        for j in range(5):
            w = [0]
            ap_python(R, min(i, j), max(i, j), w)
            want += w[0]
    assert oracle.ap_tests_rows(R, 1, 2) == want > 0


def test_cumhist_identity_for_permutations():
    R = np.stack([np.random.default_rng(s).permutation(50) for s in range(4)]).astype(np.int32)
    H = oracle.cumhist(R)
    assert np.array_equal(H, np.tile(np.arange(51, dtype=np.int32), (4, 1)))


# ------------------------------------------------------------------ max-min and MCP
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_maxmin_golden_vectors(dt):
    X = GOLD[f"r10c10_{dt}"]
    assert np.array_equal(oracle.maxmin(X, X), GOLD[f"mm_r10c10_{dt}"])
    assert np.array_equal(oracle.maxmin(X, X, GOLD[f"tf_mask_{dt}"]), GOLD[f"mm_r10c10_masked_{dt}"])
    R = GOLD[f"rand_{dt}"]
    assert np.array_equal(oracle.maxmin(R[:96], R[:71]), GOLD[f"mm_rand_{dt}"])


def test_maxmin_brute_force():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((13, 17)).astype(np.float32)
    B = rng.standard_normal((9, 17)).astype(np.float32)
    ref = np.minimum(A[:, None, :], B[None, :, :]).max(axis=2)
    assert np.array_equal(oracle.maxmin(A, B), ref)


def test_maxmin_matches_reference_build():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference tree at build time)")
    rng = np.random.default_rng(3)
    for dt in (np.float32, np.float64):
        A = rng.random((40, 77)).astype(dt)
        B = rng.random((33, 77)).astype(dt)
        km = np.where(rng.random(77) < 0.5, np.finfo(dt).max, np.finfo(dt).min).astype(dt)
        assert np.array_equal(oracle.maxmin(A, B), oracle.ref_maxmin(A, B))
        assert np.array_equal(oracle.maxmin(A, B, km), oracle.ref_maxmin(A, B, km))


def widest_walk(A, k):
    """max over walks i -> ... -> j of k-1 intermediate steps of the minimum edge (brute force)."""
    N = A.shape[0]
    import itertools
    out = np.full((N, N), -np.inf)
    for i in range(N):
        for j in range(N):
            for mid in itertools.product(range(N), repeat=k - 1):
                path = (i,) + mid + (j,)
                out[i, j] = max(out[i, j], min(A[path[t], path[t + 1]] for t in range(k)))
    return out


@pytest.mark.parametrize("k", [2, 3, 4])
def test_mcp_products_are_widest_walks(k):
    rng = np.random.default_rng(k)
    A = rng.random((5, 5)).astype(np.float32)
    A = np.maximum(A, A.T)
    np.fill_diagonal(A, 0.0)
    _, Mk = oracle.mcp(A, k)
    assert np.array_equal(Mk, widest_walk(A, k).astype(np.float32))
    assert np.array_equal(Mk, Mk.T)        # A^(x)k of a symmetric A is symmetric


def test_ratio_branches():
    f = np.float32
    eps = np.finfo(f).eps
    A = np.array([0, eps / 2, 0.3, 0.3, -0.2, 1], f)
    M = np.array([0, 0.5, 0, eps / 4, 0.7, 3], f)
    out = oracle.ratio(A, M)
    assert out[0] == 0 and out[1] == 0
    assert out[2] == np.finfo(f).max and out[3] == np.finfo(f).max
    assert out[4] == np.float32(-0.2 / 0.7) and out[5] == np.float32(1.0 / 3.0)


# ------------------------------------------------------------------ gradient
def test_gradient_matches_finite_differences():
    X = grn.grn_expression(8, 60, 9)[0]
    pi, pj, m = grn.mask_pairs(8, 0.5, 10)
    g = oracle.bspline_mi_grad_f64(X, pi, pj, m)
    X64 = X.astype(np.float64)
    rng = np.random.default_rng(0)
    h = 1e-6
    for _ in range(40):
        a, b = rng.integers(0, 8), rng.integers(0, 60)
        Xp = X64.copy(); Xp[a, b] += h
        Xm = X64.copy(); Xm[a, b] -= h
        fd = (oracle.bspline_mi_loss_x64(Xp, pi, pj, m) - oracle.bspline_mi_loss_x64(Xm, pi, pj, m)) / (2 * h)
        assert abs(fd - g[a, b]) <= 1e-5 * max(1.0, abs(fd)), (a, b, fd, g[a, b])


# ------------------------------------------------------------------ ensemble and filters
def _sym(rng, n, dt):
    A = rng.random((n, n)).astype(dt)
    A = np.maximum(A, A.T)
    np.fill_diagonal(A, 0)
    return A


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_combo_unit_coefficients_reduce_to_ratios(dt):
    """A coefficient row (0,1,0,0) is MCP2 = ratio(MI, M1), (0,0,1,0) MCP3, (0,0,0,1) MCP4 -- each
    through orc_ratio, an independent function -- and (1,0,0,0) gives 1 wherever |MI| >= eps."""
    rng = np.random.default_rng(11)
    A = _sym(rng, 40, dt)
    A[3, 5] = A[5, 3] = 0
    M1, M2, M3 = oracle.maxmins(A)
    for q, M in enumerate((M1, M2, M3)):
        cf = np.zeros(4, dt)
        cf[q + 1] = 1
        assert np.array_equal(oracle.combo(A, M1, M2, M3, cf), oracle.ratio(A, M))
    ones = oracle.combo(A, M1, M2, M3, np.array([1, 0, 0, 0], dt))
    assert np.array_equal(ones, np.where(np.abs(A) < np.finfo(dt).eps, 0, 1).astype(dt))
    # a general row equals the double-precision expression of compute_combo
    cf = np.array([0.25, 0.25, 0.25, 0.25], dt)
    s = sum(float(c) * M.astype(np.float64) for c, M in zip(cf, (A, M1, M2, M3)))
    want = np.where(np.abs(A) < np.finfo(dt).eps, 0, np.where(np.abs(s) < np.finfo(dt).eps, np.finfo(dt).max,
                                                                A.astype(np.float64) / s)).astype(dt)
    assert np.array_equal(oracle.combo(A, M1, M2, M3, cf), want)


@pytest.mark.parametrize("exclude_ones", [False, True])
def test_stouffer_matches_numpy_zscores(exclude_ones):
    rng = np.random.default_rng(12)
    X = rng.random((50, 50))
    X[rng.random((50, 50)) < 0.1] = 1.0
    X[7] = 0.25                                      # a constant row: sd 0 -> its z-scores are 0
    keep = (X != 1.0) if exclude_ones else np.ones_like(X, bool)
    mean = np.array([X[r][keep[r]].mean() for r in range(50)])
    sd = np.array([X[r][keep[r]].std(ddof=1) for r in range(50)])
    zi = np.where(sd[:, None] < np.finfo(np.float64).eps, 0, (X - mean[:, None]) / np.where(sd[:, None] == 0, 1, sd[:, None]))
    zj = np.where(sd[None, :] < np.finfo(np.float64).eps, 0, (X - mean[None, :]) / np.where(sd[None, :] == 0, 1, sd[None, :]))
    want = (zi + zj) * np.sqrt(0.5)
    got = oracle.stouffer(X, exclude_ones)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    got32 = oracle.stouffer(X.astype(np.float32), exclude_ones)
    assert np.allclose(got32, want, rtol=1e-5, atol=1e-5)


def test_stouffer_tf_matches_scipy_zscores():
    """compute_stouffer_tf: StandardScore<T, T, false> of the TF rows and of the gene columns.
    The `false` is the population (ddof 0) standard deviation: only with it do z-scores give the
    Pearson correlation as dot / count, which is how the reference's PearsonKernel
    (include/mcp/correlation/Pearson.hpp:34-36) uses the same StandardScore<T, T, false>."""
    from scipy import stats
    rng = np.random.default_rng(21)
    X = rng.random((12, 40))
    X[3] = 0.5                                       # constant TF row: its z-scores are 0
    X[:, 17] = 0.25                                  # constant gene column: its z-scores are 0
    with np.errstate(invalid="ignore", divide="ignore"):
        zr = np.nan_to_num(stats.zscore(X, axis=1, ddof=0))
        zc = np.nan_to_num(stats.zscore(X, axis=0, ddof=0))
    want = (zr + zc) * np.sqrt(0.5)
    assert np.allclose(oracle.stouffer_tf(X), want, rtol=1e-12, atol=1e-12)
    assert np.allclose(oracle.stouffer_tf(X.astype(np.float32)), want, rtol=1e-5, atol=1e-5)
    Y = rng.random((6, 30))
    zy = stats.zscore(Y, axis=1, ddof=0)
    assert np.allclose(zy @ zy.T / Y.shape[1], np.corrcoef(Y), rtol=1e-12, atol=1e-12)
    zy1 = stats.zscore(Y, axis=1, ddof=1)
    assert not np.allclose(zy1 @ zy1.T / Y.shape[1], np.corrcoef(Y), rtol=1e-6, atol=1e-6)


def _dpi_numpy(A, tol, tfs=None):
    """Vectorised statement of the DPI rule: edge x-z goes when some y != x, z has
    A[x][y] >= i_xz and A[z][y] >= i_xz (y restricted to TFs when exactly one endpoint is a TF)."""
    n = A.shape[0]
    dt = A.dtype
    out = A.copy()
    for x in range(n):
        for z in range(n):
            if x == z:
                out[x, z] = 0
                continue
            t = np.float64(dt.type(tol))                # the tolerance arrives as T
            if tfs is None:
                ixz = dt.type(np.float64(A[x, z]) * (1.0 / (1.0 - t)))
                ok = np.ones(n, bool)
            else:
                ixz = A[x, z] * dt.type(1.0 / (1.0 - t))
                ok = np.ones(n, bool) if (tfs[x] > 0) == (tfs[z] > 0) else (tfs > 0)
            ok[[x, z]] = False
            if np.any(ok & (A[x] >= ixz) & (A[z] >= ixz)):
                out[x, z] = 0
    return out


def test_dpi_removes_the_weakest_triangle_edge():
    A = np.array([[0, .9, .5], [.9, 0, .8], [.5, .8, 0]], np.float32)
    out, n = oracle.dpi(A, 0.0)
    assert np.array_equal(out, np.array([[0, .9, 0], [.9, 0, .8], [0, .8, 0]], np.float32)) and n == 2
    # a tolerance above the gap keeps it: 0.5 / (1 - 0.45) > 0.8
    out, n = oracle.dpi(A, 0.45)
    assert np.array_equal(out, A) and n == 0


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("with_tfs", [False, True])
def test_dpi_matches_vectorised_statement(dt, with_tfs):
    rng = np.random.default_rng(13)
    n = 45
    A = (rng.random((n, n)) ** 2).astype(dt)        # not symmetric: i_xz uses row x
    tfs = np.where(rng.random(n) < 0.3, np.finfo(dt).max, np.finfo(dt).min).astype(dt) if with_tfs else None
    for tol in (0.0, 0.1, 0.3):
        got, cnt = oracle.dpi(A, tol, tfs)
        want = _dpi_numpy(A, tol, tfs)
        assert np.array_equal(got, want)
        assert cnt == int(np.count_nonzero(got != A))
