"""Max-min products, ratio and MCP2/3/4 on the GPU: bit-exact against the reference's own
max_of_pairmin (golden vectors from oracle/_ref) and the CPU oracle."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from mcpnet_amd import grn

pytestmark = pytest.mark.gpu

GOLD = np.load(Path(__file__).with_name("golden") / "maxmin_ref.npz")


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_golden_r10c10(handle, dt):
    X = GOLD[f"r10c10_{dt}"]
    assert bits_equal(handle.maxmin(X, X), GOLD[f"mm_r10c10_{dt}"])
    assert bits_equal(handle.maxmin(X, X, symmetric=True), GOLD[f"mm_r10c10_{dt}"])
    assert bits_equal(handle.maxmin(X, X, kmask=GOLD[f"tf_mask_{dt}"]), GOLD[f"mm_r10c10_masked_{dt}"])


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_golden_random_with_infinities(handle, dt):
    R = GOLD[f"rand_{dt}"]
    got = handle.maxmin(np.ascontiguousarray(R[:96]), np.ascontiguousarray(R[:71]))
    assert bits_equal(got, GOLD[f"mm_rand_{dt}"])


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 5, 7), (127, 129, 131), (128, 128, 32), (257, 130, 1000),
                                   (200, 300, 33)])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_random_shapes_vs_oracle(handle, shape, dt):
    M, N, K = shape
    rng = np.random.default_rng(M * 1000 + N + K)
    A = (rng.random((M, K)) * 2 - 0.5).astype(dt)
    B = (rng.random((N, K)) * 2 - 0.5).astype(dt)
    got = handle.maxmin(A, B)
    assert bits_equal(got, oracle.maxmin(A, B))
    km = np.where(rng.random(K) < 0.4, np.finfo(dt).max, np.finfo(dt).min).astype(dt)
    assert bits_equal(handle.maxmin(A, B, kmask=km), oracle.maxmin(A, B, km))


def test_unaligned_device_views(handle, torch_cuda):
    rng = np.random.default_rng(5)
    big = rng.random((150, 203)).astype(np.float32)
    bd = torch_cuda.from_numpy(big).cuda()
    A = bd[:, 1:202]           # misaligned start, ld = 203 elements
    B = bd[3:120, 2:203]
    ref = oracle.maxmin(np.ascontiguousarray(big[:, 1:202]), np.ascontiguousarray(big[3:120, 2:203]))
    assert bits_equal(handle.maxmin(A, B).cpu().numpy(), ref)


def test_empty_k_follows_reference_init(handle):
    A32 = np.zeros((4, 0), np.float32)
    A64 = np.zeros((4, 0), np.float64)
    assert np.all(handle.maxmin(A32, A32) == -np.inf)
    assert np.all(handle.maxmin(A64, A64) == -np.finfo(np.float64).max)


def test_symmetric_equals_full(handle, torch_cuda):
    rng = np.random.default_rng(11)
    A = rng.random((333, 333)).astype(np.float32)
    A = np.maximum(A, A.T)
    Ad = torch_cuda.from_numpy(A).cuda()
    full = handle.maxmin(Ad, Ad).cpu().numpy()
    sym = handle.maxmin(Ad, Ad, symmetric=True).cpu().numpy()
    assert bits_equal(full, sym)
    assert bits_equal(sym, sym.T)


def test_ratio_branches(handle):
    eps32 = np.finfo(np.float32).eps
    A = np.array([[0.0, eps32 / 2, 0.3, 0.3, -0.2, 1.0]], np.float32)
    M = np.array([[0.0, 0.5, 0.0, eps32 / 4, 0.7, 3.0]], np.float32)
    got = handle.ratio(A, M)
    assert bits_equal(got, oracle.ratio(A, M))
    assert got[0, 0] == 0 and got[0, 1] == 0 and got[0, 2] == np.finfo(np.float32).max
    rng = np.random.default_rng(3)
    A = rng.random((300, 301)).astype(np.float32)
    M = rng.random((300, 301)).astype(np.float32)
    assert bits_equal(handle.ratio(A, M), oracle.ratio(A, M))
    assert bits_equal(handle.ratio(A.astype(np.float64), M.astype(np.float64)),
                      oracle.ratio(A.astype(np.float64), M.astype(np.float64)))


def mi_like(G, seed):
    X, _, _ = grn.grn_expression(G, 256, seed)
    A = oracle.bspline_mi_f64(X).astype(np.float32)
    np.fill_diagonal(A, 0.0)
    return A


@pytest.mark.parametrize("k", [2, 3, 4])
def test_mcp_vs_oracle(handle, k):
    A = mi_like(300, 99)
    out, mm = handle.mcp(A, k, want_maxmin=True)
    ref_out, ref_mm = oracle.mcp(A, k)
    assert bits_equal(mm, ref_mm)
    assert bits_equal(out, ref_out)


def test_mcp3_asymmetric_input_takes_full_path(handle):
    rng = np.random.default_rng(17)
    A = rng.random((200, 200)).astype(np.float32)   # not symmetric
    out, mm = handle.mcp(A, 3, want_maxmin=True)
    ref_out, ref_mm = oracle.mcp(A, 3)
    assert bits_equal(mm, ref_mm)
    assert bits_equal(out, ref_out)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_mcp_device_pointers(handle, torch_cuda, dt):
    A = mi_like(257, 5).astype(dt)
    Ad = torch_cuda.from_numpy(A).cuda()
    for k in (2, 3, 4):
        out, mm = handle.mcp(Ad, k, want_maxmin=True)
        ref_out, ref_mm = oracle.mcp(A, k)
        assert bits_equal(mm.cpu().numpy(), ref_mm)
        assert bits_equal(out.cpu().numpy(), ref_out)


@pytest.mark.slow
def test_mcp3_1000_genes_full(handle, torch_cuda):
    """MCP3 on a 1000-gene FP32 MI matrix (1000 samples), bit-exact against the CPU oracle."""
    X = grn.expression(1)
    with_gpu = handle.bspline_mi(torch_cuda.from_numpy(X).cuda(), precision="fp32")
    A = with_gpu.cpu().numpy()
    np.fill_diagonal(A, 0.0)
    out, mm = handle.mcp(torch_cuda.from_numpy(A).cuda(), 3, want_maxmin=True)
    ref_out, ref_mm = oracle.mcp(A, 3)
    assert bits_equal(mm.cpu().numpy(), ref_mm)
    assert bits_equal(out.cpu().numpy(), ref_out)


@pytest.mark.parametrize("tri", [True, False])
def test_split_tail_matches_unsplit(handle, monkeypatch, tri):
    """Grids larger than one round of resident workgroups run their last tiles split over K with
    atomicMax on the keys: bit-identical to the unsplit launch and to the oracle."""
    rng = np.random.default_rng(7 if tri else 8)
    N, K = (5200, 96) if tri else (3700, 96)
    A = rng.standard_normal((N, K)).astype(np.float32)
    A[rng.random((N, K)) < 0.01] = -np.inf
    B = A if tri else rng.standard_normal((N, K)).astype(np.float32)
    monkeypatch.setenv("MCPG_MM_SPLIT", "0")
    ref = handle.maxmin(A, B, symmetric=tri)
    monkeypatch.setenv("MCPG_MM_SPLIT", "1")
    handle.set_profiling(True)
    got = handle.maxmin(A, B, symmetric=tri)
    names = [n for n, _ in handle.last_timings()]
    handle.set_profiling(False)
    assert "mm_finish_tail_kernel" in names          # the split path ran
    assert bits_equal(got, ref)
    for rows in (slice(0, 40), slice(N - 40, N)):   # the tail tiles sit at the end of the order
        assert bits_equal(got[rows], oracle.maxmin(np.ascontiguousarray(A[rows]), B))
    if tri:
        M1 = handle.maxmin(A, A, symmetric=True)
        assert bits_equal(M1, M1.T)


def test_maxmin_rejects_bad_arguments(handle):
    import mcpnet_amd
    A = np.ones((4, 5), np.float32)
    B = np.ones((4, 5), np.float32)
    with pytest.raises(mcpnet_amd.MCPGError):
        handle.maxmin(A, B, symmetric=True)   # symmetric needs the same operand
    with pytest.raises(ValueError):
        handle.maxmin(A, np.ones((4, 6), np.float32))


@pytest.mark.parametrize("env", [{"MCPG_MM_KC": "32"}, {"MCPG_MM_GROUP": "3", "MCPG_MM_XCD": "1"}])
def test_tuning_switches_keep_bits(handle, monkeypatch, env):
    """K chunk, tile-group side and XCD deal change the schedule only: bit-identical results,
    for a rectangular product and a symmetric (triangle) one."""
    rng = np.random.default_rng(11)
    A = rng.standard_normal((700, 300)).astype(np.float32)
    B = rng.standard_normal((450, 300)).astype(np.float32)
    Y = rng.standard_normal((500, 500)).astype(np.float32)
    Y = np.ascontiguousarray(np.maximum(Y, Y.T))
    ref_rect = handle.maxmin(A, B)
    ref_tri = handle.maxmin(Y, Y, symmetric=True)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    assert bits_equal(handle.maxmin(A, B), ref_rect)
    assert bits_equal(handle.maxmin(Y, Y, symmetric=True), ref_tri)
    assert bits_equal(ref_rect[:30], oracle.maxmin(np.ascontiguousarray(A[:30]), B))


def _names(handle, fn):
    handle.set_profiling(True)
    try:
        out = fn()
        names = [n for n, _ in handle.last_timings()]
    finally:
        handle.set_profiling(False)
    return out, names


@pytest.mark.parametrize("shape,tri", [((1500, 1500, 1500), True), ((700, 900, 400), False),
                                       ((4224, 4224, 96), True), ((2900, 2900, 64), False)])
def test_f64_split_tail_matches_unsplit(handle, monkeypatch, shape, tri):
    """FP64 tiles beyond the last full round of resident workgroups (every tile of a grid smaller
    than one round) run split over K, their partial maxima folded with a 64-bit atomicMax on
    order keys: bit-identical to the unsplit launch and to the oracle."""
    M, N, K = shape
    rng = np.random.default_rng(M + K)
    A = rng.standard_normal((M, K))
    A[rng.random((M, K)) < 0.01] = -np.inf
    B = A if tri else rng.standard_normal((N, K))
    monkeypatch.setenv("MCPG_MM_SPLIT", "0")
    ref = handle.maxmin(A, B, symmetric=tri)
    monkeypatch.setenv("MCPG_MM_SPLIT", "1")
    got, names = _names(handle, lambda: handle.maxmin(A, B, symmetric=tri))
    if M <= 1500:                        # far below one round of resident workgroups: always split
        assert "mv_finish_tail_kernel" in names
    assert bits_equal(got, ref)
    for rows in (slice(0, 24), slice(M - 24, M)):
        assert bits_equal(got[rows], oracle.maxmin(np.ascontiguousarray(A[rows]), B))


@pytest.mark.parametrize("env", [{"MCPG_MM64_KD": "8"}, {"MCPG_MM64_TILE": "1"}])
def test_f64_kernels_agree(handle, monkeypatch, env):
    """The direct-load FP64 kernel at both K chunks and the register-staged tile kernel give the
    same bits, for a masked rectangular product, a triangle product and MCP3."""
    rng = np.random.default_rng(21)
    A = rng.standard_normal((500, 333))
    B = rng.standard_normal((260, 333))
    km = np.where(rng.random(333) < 0.5, np.finfo(np.float64).max, np.finfo(np.float64).min)
    Y = mi_like(300, 8).astype(np.float64)
    monkeypatch.setenv("MCPG_MM64_KEYS", "0")         # Y is float-exact: keep it on the FP64 kernels
    ref = (handle.maxmin(A, B, kmask=km), handle.maxmin(Y, Y, symmetric=True), handle.mcp(Y, 3, want_maxmin=True))
    assert bits_equal(ref[0], oracle.maxmin(A, B, km))
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    got = (handle.maxmin(A, B, kmask=km), handle.maxmin(Y, Y, symmetric=True), handle.mcp(Y, 3, want_maxmin=True))
    assert bits_equal(got[0], ref[0])
    assert bits_equal(got[1], ref[1])
    assert bits_equal(got[2][0], ref[2][0]) and bits_equal(got[2][1], ref[2][1])


# ---- FP64 operands that every float holds exactly run on the key kernel

F64_MAX = np.finfo(np.float64).max
F64_LOW = np.finfo(np.float64).min


def float_exact(rng, shape):
    return rng.standard_normal(shape).astype(np.float32).astype(np.float64)


def _on_fp64_kernel(handle, monkeypatch, fn):
    monkeypatch.setenv("MCPG_MM64_KEYS", "0")
    try:
        return fn()
    finally:
        monkeypatch.delenv("MCPG_MM64_KEYS")


def _route(names):
    keys = "key_prep_f64_kernel" in names and any(n.startswith("maxmin_key_kernel") for n in names)
    vals = any(n.startswith("maxmin_f64_kernel") for n in names)
    assert keys != vals, names
    return "keys" if keys else "fp64"


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 5, 7), (129, 127, 131), (257, 130, 1000), (400, 400, 300)])
def test_f64_key_route_matches_fp64_kernel(handle, monkeypatch, shape):
    """Float-exact FP64 operands: the key kernel with widened outputs gives the FP64 kernel's and
    the oracle's bits, unmasked, with max/lowest<double> masks (the mcpnet TF mask), with every K
    masked (lowest<double> everywhere), with finite float-exact mask values, and for a symmetric
    triangle."""
    M, N, K = shape
    rng = np.random.default_rng(M + 7 * K)
    A, B = float_exact(rng, (M, K)), float_exact(rng, (N, K))
    S = float_exact(rng, (M, M))
    S = np.maximum(S, S.T)
    cases = [(A, B, None, False), (A, B, np.where(rng.random(K) < 0.4, F64_MAX, F64_LOW), False),
             (A, B, np.full(K, F64_LOW), False), (A, B, np.full(K, -np.inf), False),
             (A, B, float_exact(rng, K), False), (S, S, None, True)]
    for a, b, km, sym in cases:
        got, names = _names(handle, lambda: handle.maxmin(a, b, kmask=km, symmetric=sym))
        assert _route(names) == "keys"
        ref = _on_fp64_kernel(handle, monkeypatch, lambda: handle.maxmin(a, b, kmask=km, symmetric=sym))
        assert bits_equal(got, ref)
        rows = slice(0, min(M, 40))
        assert bits_equal(got[rows], oracle.maxmin(np.ascontiguousarray(a[rows]), b, km))
    assert np.all(handle.maxmin(A, B, kmask=np.full(K, F64_LOW)) == F64_LOW)


@pytest.mark.parametrize("case", ["inexact", "minus_zero", "inf", "nan", "mask_inexact", "mask_huge_negative"])
def test_f64_key_route_falls_back(handle, monkeypatch, case):
    """One entry a float does not hold exactly (or -0, inf, NaN; or a mask value that is neither
    float-exact nor beyond the float range in a way the key kernel reproduces) sends the product
    to the FP64 kernel, whose bits it returns."""
    rng = np.random.default_rng(5)
    A, B = float_exact(rng, (70, 90)), float_exact(rng, (60, 90))
    km = None
    if case == "inexact":
        A[33, 17] = 0.1
    elif case == "minus_zero":
        B[59, 89] = -0.0
    elif case == "inf":
        A[0, 0] = np.inf
    elif case == "nan":
        B[5, 5] = np.nan
    elif case == "mask_inexact":
        km = np.full(90, F64_MAX)
        km[3] = 0.1
    else:
        km = np.full(90, F64_MAX)
        km[3] = -1e300        # all-masked results would be -1e300, not lowest<double>
    got, names = _names(handle, lambda: handle.maxmin(A, B, kmask=km))
    assert _route(names) == "fp64"
    assert bits_equal(got, _on_fp64_kernel(handle, monkeypatch, lambda: handle.maxmin(A, B, kmask=km)))
    if case != "nan":
        assert bits_equal(got, oracle.maxmin(A, B, km))


def test_f64_key_route_mask_beyond_float_range(handle, monkeypatch):
    """Mask values above FLT_MAX act as +inf on the key kernel: min(a, m) = a for every finite
    float a, as in FP64."""
    rng = np.random.default_rng(6)
    A, B = float_exact(rng, (80, 64)), float_exact(rng, (75, 64))
    km = np.where(rng.random(64) < 0.5, 1e300, F64_LOW)
    km[:2] = [np.inf, 3.5e38]
    got, names = _names(handle, lambda: handle.maxmin(A, B, kmask=km))
    assert _route(names) == "keys"
    assert bits_equal(got, _on_fp64_kernel(handle, monkeypatch, lambda: handle.maxmin(A, B, kmask=km)))
    assert bits_equal(got, oracle.maxmin(A, B, km))


@pytest.mark.parametrize("where", ["host", "device"])
def test_f64_key_route_mcp_and_maxmins(handle, torch_cuda, monkeypatch, where):
    """MCP2/3/4 (with and without the max-min output) and the three max-min chains of a
    float-exact FP64 MI matrix run on the key kernel, the ratio taken in double: the FP64
    kernels' and the oracle's bits.  An asymmetric float-exact input takes MCP3's full product."""
    Y = mi_like(300, 12).astype(np.float64)
    rng = np.random.default_rng(13)
    Z = float_exact(rng, (200, 200))
    put = (lambda x: torch_cuda.from_numpy(x).cuda()) if where == "device" else (lambda x: x)
    get = (lambda x: x.cpu().numpy()) if where == "device" else (lambda x: x)
    for A, ks in ((Y, (2, 3, 4)), (Z, (3,))):
        for k in ks:
            (out, mm), names = _names(handle, lambda: handle.mcp(put(A), k, want_maxmin=True))
            assert _route(names) == "keys"
            ref_out, ref_mm = oracle.mcp(A, k)
            assert bits_equal(get(mm), ref_mm), k
            assert bits_equal(get(out), ref_out), k
            only = handle.mcp(put(A), k)[0]
            assert bits_equal(get(only), ref_out), k
    (m1, m2, m3), names = _names(handle, lambda: handle.maxmins(put(Y)))
    assert _route(names) == "keys"
    ref = _on_fp64_kernel(handle, monkeypatch, lambda: handle.maxmins(put(Y)))
    for g, r in zip((m1, m2, m3), ref):
        assert bits_equal(get(g), get(r))
    assert bits_equal(get(m1), oracle.maxmin(Y, Y))


def test_f64_key_route_tf_chain(handle):
    """mcpnet on a float-exact FP64 MI with a TF mask of max/lowest<double> (mcpnet's load_tfs):
    every masked product runs on the key kernel; the scores equal the oracle's."""
    X = grn.grn_expression(48, 256, 34)[0]
    A = oracle.bspline_mi_f64(X).astype(np.float32).astype(np.float64)
    tf_idx = [1, 4, 9, 22, 40]
    tfs = np.full(48, F64_LOW)
    tfs[tf_idx] = F64_MAX
    res, names = _names(handle, lambda: handle.mcpnet(A, mi_method=-1, tfs=tfs))
    assert "maxmin_f64_kernel" not in names and "key_prep_f64_kernel" in names
    Ad = A.copy()
    np.fill_diagonal(Ad, 0)
    T = np.ascontiguousarray(Ad[tf_idx])
    mm = oracle.maxmin
    G1 = mm(Ad, Ad, tfs)
    want = {2: mm(T, Ad, tfs), 3: mm(T, G1, tfs), 4: mm(mm(T, Ad, tfs), G1, tfs)}
    for k in (2, 3, 4):
        assert bits_equal(res[k], oracle.ratio(T, want[k])), k


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("k,sym", [(2, True), (3, True), (3, False), (4, True)])
def test_mcp_host_row_panels(handle, torch_cuda, dt, k, sym):
    """Host buffers at 4000 genes: a triangle last product runs in three row panels (16, 8 and 8
    row blocks, the last one partial), the scores and the max-min matrix of each panel downloaded
    while the next runs; an asymmetric A makes MCP3's last product a full grid, downloaded whole.
    Both equal the bits of the device-buffer call, whose products run whole."""
    X = grn.grn_expression(4000, 64, 39)[0]
    A = handle.bspline_mi(X, precision="bf16x3")
    np.fill_diagonal(A, 0.0)
    if not sym:
        A[5, 9] = np.nextafter(A[5, 9], np.float32(1))
    A = A.astype(dt)
    got, got_mm = handle.mcp(A, k, want_maxmin=True)
    dev, dev_mm = handle.mcp(torch_cuda.from_numpy(A).cuda(), k, want_maxmin=True)
    assert bits_equal(got, dev.cpu().numpy())
    assert bits_equal(got_mm, dev_mm.cpu().numpy())
    only = handle.mcp(A, k)[0]                      # scores alone: one download per panel
    assert bits_equal(only, got)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("want", [(True, True, True), (False, True, True), (True, False, True), (True, True, False),
                                  (True, False, False)])
def test_maxmins_host_downloads_overlap(handle, torch_cuda, dt, want):
    """Host buffers at 4000 genes: each wanted matrix goes down while the next product runs and
    the last product runs in row panels; every matrix equals the device-buffer call's bits."""
    X = grn.grn_expression(4000, 64, 40)[0]
    A = handle.bspline_mi(X, precision="bf16x3")
    np.fill_diagonal(A, 0.0)
    A = A.astype(dt)
    got = handle.maxmins(A, want)
    dev = handle.maxmins(torch_cuda.from_numpy(A).cuda(), want)
    for g, d, w in zip(got, dev, want):
        assert (g is None) == (not w) and (d is None) == (not w)
        if w:
            assert bits_equal(g, d.cpu().numpy())
