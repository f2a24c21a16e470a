"""Blocks of the G x G MI output (the row-partition unit of the reference's distributed
InnerProduct, mcpnet_helper.hpp:90-98 and :156-162): every block must equal, bit for bit, the
same entries of the whole-matrix result, for blocks on, off and across the diagonal, ragged
against the 16-gene MI tiles, down to single entries."""
import numpy as np
import pytest
import torch

from mcpnet_amd import grn

pytestmark = pytest.mark.gpu

G = 150
BLOCKS = [
    (0, G, 0, G),        # whole matrix
    (0, 37, 0, G),       # row band (one rank's partition)
    (37, 41, 0, G),
    (0, G, 101, 49),     # column band
    (16, 32, 48, 80),    # tile-aligned, above the diagonal
    (96, 54, 0, 40),     # below the diagonal
    (5, 60, 3, 70),      # ragged, across the diagonal
    (77, 1, 77, 1),      # one diagonal entry
    (149, 1, 0, 1),      # the last row's first entry
    (20, 0, 0, 5),       # empty
]


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                                  np.ascontiguousarray(b).view(np.uint32))


@pytest.mark.parametrize("precision", ["fp32", "bf16", "bf16x3", "i8x3"])
def test_bspline_mi_blocks_equal_whole_matrix(handle, precision):
    X = grn.grn_expression(G, 300, 4242)[0]
    whole = handle.bspline_mi(X, precision=precision)
    Xd = torch.from_numpy(X).cuda()
    for r0, nr, c0, nc in BLOCKS:
        got = handle.bspline_mi_block(X, r0, nr, c0, nc, precision=precision)
        assert same_bits(got, whole[r0:r0 + nr, c0:c0 + nc]), (precision, r0, nr, c0, nc)
        gd = handle.bspline_mi_block(Xd, r0, nr, c0, nc, precision=precision)
        torch.cuda.synchronize()
        assert same_bits(gd.cpu().numpy(), whole[r0:r0 + nr, c0:c0 + nc]), (precision, "device", r0, nr, c0, nc)


def test_bspline_mi_block_into_strided_output(handle):
    X = grn.grn_expression(G, 200, 99)[0]
    whole = handle.bspline_mi(X, precision="bf16x3")
    buf = np.full((40, 64), -7.0, np.float32)
    handle.bspline_mi_block(X, 30, 40, 90, 50, precision="bf16x3", out=buf[:, :50])
    assert same_bits(buf[:, :50], whole[30:70, 90:140])
    assert np.all(buf[:, 50:] == -7.0)


def test_ap_mi_blocks_equal_whole_matrix(handle):
    X = grn.grn_expression(120, 250, 777)[0]
    R = grn.rank_desc(grn.jitter(X, 5))
    whole = handle.ap_mi(R)
    Rd = torch.from_numpy(R).cuda()
    for r0, nr, c0, nc in [(0, 120, 0, 120), (0, 30, 0, 120), (30, 31, 0, 120), (0, 120, 90, 30), (10, 50, 5, 70),
                           (100, 20, 0, 30), (64, 1, 64, 1), (119, 1, 0, 1), (7, 0, 0, 3)]:
        got = handle.ap_mi_block(R, r0, nr, c0, nc)
        assert same_bits(got, whole[r0:r0 + nr, c0:c0 + nc]), (r0, nr, c0, nc)
        gd = handle.ap_mi_block(Rd, r0, nr, c0, nc)
        torch.cuda.synchronize()
        assert same_bits(gd.cpu().numpy(), whole[r0:r0 + nr, c0:c0 + nc]), ("device", r0, nr, c0, nc)


@pytest.mark.parametrize("bins,precision", [(7, "fp32"), (7, "bf16x3"), (16, "bf16"), (7, "i8x3")])
def test_bspline_mi_blocks_one_row_order(handle, bins, precision):
    # bins other than 10 take the one-row-order kernels and their LDS epilogue
    X = grn.grn_expression(70, 150, 31)[0]
    whole = handle.bspline_mi(X, bins=bins, precision=precision)
    for r0, nr, c0, nc in [(0, 70, 0, 70), (3, 20, 40, 30), (50, 20, 0, 25), (10, 40, 10, 40), (69, 1, 69, 1)]:
        got = handle.bspline_mi_block(X, r0, nr, c0, nc, bins=bins, precision=precision)
        assert same_bits(got, whole[r0:r0 + nr, c0:c0 + nc]), (bins, precision, r0, nr, c0, nc)


def test_ap_mi_blocks_with_tied_ranks(handle):
    # ranks with ties (not a permutation): the histograms are not the identity
    rng = np.random.default_rng(12)
    R = rng.integers(0, 90, size=(40, 90)).astype(np.int32)
    whole = handle.ap_mi(R)
    for r0, nr, c0, nc in [(0, 40, 0, 40), (5, 10, 20, 20), (30, 10, 0, 35)]:
        got = handle.ap_mi_block(R, r0, nr, c0, nc)
        assert same_bits(got, whole[r0:r0 + nr, c0:c0 + nc]), (r0, nr, c0, nc)


def same_bits64(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                                                  np.ascontiguousarray(b).view(np.uint64))


def test_bspline_mi_blocks_fp64(handle):
    """The double DataType's row partitions: FP64 MI blocks equal the whole FP64 matrix bitwise."""
    rng = np.random.default_rng(3)
    X = grn.grn_expression(G, 240, 4343)[0].astype(np.float64) * (1.0 + 1e-9 * rng.standard_normal((G, 240)))
    whole = handle.bspline_mi(X, precision="fp64")
    Xd = torch.from_numpy(X).cuda()
    for r0, nr, c0, nc in BLOCKS:
        got = handle.bspline_mi_block(X, r0, nr, c0, nc, precision="fp64")
        assert got.dtype == np.float64 and same_bits64(got, whole[r0:r0 + nr, c0:c0 + nc]), (r0, nr, c0, nc)
        gd = handle.bspline_mi_block(Xd, r0, nr, c0, nc, precision="fp64")
        torch.cuda.synchronize()
        assert same_bits64(gd.cpu().numpy(), whole[r0:r0 + nr, c0:c0 + nc]), ("device", r0, nr, c0, nc)


def test_ap_mi_blocks_fp64(handle):
    X = grn.grn_expression(90, 250, 778)[0]
    R = grn.rank_desc(grn.jitter(X, 6))
    whole = handle.ap_mi(R, dtype=np.float64)
    for r0, nr, c0, nc in [(0, 90, 0, 90), (0, 30, 0, 90), (10, 50, 5, 70), (64, 1, 64, 1), (7, 0, 0, 3)]:
        got = handle.ap_mi_block(R, r0, nr, c0, nc, dtype=np.float64)
        assert got.dtype == np.float64 and same_bits64(got, whole[r0:r0 + nr, c0:c0 + nc]), (r0, nr, c0, nc)
