"""The allocation-free MCP entry point, mcpg_mcp_f32_noalloc, called through the C ABI: its
max-min and ratio outputs equal mcpg_mcp_f32's bit for bit (which the oracle pins elsewhere), on
consecutive calls whose sizes shrink, so each call runs on what the previous one left in the
shared static scratch; sizes beyond the scratch are rejected before any GPU work."""
import numpy as np
import pytest

import oracle
from mcpnet_amd import grn

pytestmark = pytest.mark.gpu

MCPG_ERR_UNSUPPORTED = -7
MAX_N = {2: 4064, 3: 2848, 4: 2848}


def bits(t):
    return t.contiguous().view(t.new_empty(0, dtype=__import__("torch").int32).dtype).cpu().numpy()


def noalloc(handle, A, k, want_mm=True):
    import torch
    N = A.shape[0]
    out = torch.full((N, N), float("nan"), dtype=torch.float32, device=A.device)
    mm = torch.full((N, N), float("nan"), dtype=torch.float32, device=A.device) if want_mm else None
    st = handle.lib.mcpg_mcp_f32_noalloc(handle.h, A.data_ptr(), A.stride(0), N, k,
                                         mm.data_ptr() if mm is not None else None, mm.stride(0) if mm is not None else 0,
                                         out.data_ptr(), out.stride(0))
    return st, out, mm


def mi_matrix(handle, torch, G, S, seed):
    X = grn.grn_expression(G, S, seed)[0]
    A = handle.bspline_mi(torch.from_numpy(X).cuda(), precision="bf16x3")
    A.fill_diagonal_(0.0)
    return A


def test_noalloc_equals_mcp_on_shrinking_calls(handle, torch_cuda):
    torch = torch_cuda
    mats = [mi_matrix(handle, torch, 2000, 256, 31), mi_matrix(handle, torch, 700, 256, 32)]
    refs = {}
    for i, A in enumerate(mats):
        for k in (2, 3, 4):
            refs[i, k] = handle.mcp(A, k, want_maxmin=True)
    for k in (3, 2, 4):                      # two calls in a row, the second one smaller
        for i, A in enumerate(mats):
            print(f"\nnoalloc MCP{k} on {A.shape[0]} genes", flush=True)
            st, out, mm = noalloc(handle, A, k)
            assert st == 0, handle.lib.mcpg_last_error(handle.h)
            ref_out, ref_mm = refs[i, k]
            assert np.array_equal(bits(mm), bits(ref_mm))
            assert np.array_equal(bits(out), bits(ref_out))
    # without the max-min output, after a larger call
    st, out, _ = noalloc(handle, mats[1], 3, want_mm=False)
    assert st == 0
    assert np.array_equal(bits(out), bits(refs[1, 3][0]))
    # the 700-gene MCP3 also against the oracle
    A = mats[1].cpu().numpy()
    ref_out, ref_mm = oracle.mcp(A, 3)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref_out.view(np.uint32))


def test_noalloc_asymmetric_and_largest_inputs(handle, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(9)
    R = rng.random((333, 333)).astype(np.float32)             # not symmetric: full A (x) M1 product
    Ad = torch.from_numpy(R).cuda()
    ref_out, ref_mm = handle.mcp(Ad, 3, want_maxmin=True)
    st, out, mm = noalloc(handle, Ad, 3)
    assert st == 0
    assert np.array_equal(bits(mm), bits(ref_mm)) and np.array_equal(bits(out), bits(ref_out))
    N = MAX_N[4]
    Y = rng.random((N, N)).astype(np.float32)
    Y = np.maximum(Y, Y.T)
    np.fill_diagonal(Y, 0.0)
    Yd = torch.from_numpy(Y).cuda()
    print(f"\nnoalloc MCP4 on {N} genes (largest)", flush=True)
    ref_out, ref_mm = handle.mcp(Yd, 4, want_maxmin=True)
    st, out, mm = noalloc(handle, Yd, 4)
    assert st == 0
    assert np.array_equal(bits(mm), bits(ref_mm)) and np.array_equal(bits(out), bits(ref_out))


@pytest.mark.parametrize("k", [2, 3, 4])
def test_noalloc_rejects_sizes_beyond_the_scratch(handle, torch_cuda, k):
    torch = torch_cuda
    N = MAX_N[k] + 1
    A = torch.zeros((N, N), dtype=torch.float32, device="cuda")
    st, out, mm = noalloc(handle, A, k)
    assert st == MCPG_ERR_UNSUPPORTED
    assert b"scratch" in handle.lib.mcpg_last_error(handle.h)
    assert bool(torch.isnan(out).all())                        # nothing ran
