"""The call-site program (tests/dropin/callsites.cpp: the reference's compute_mi, compute_ap_mi and
compute_mcp2/3/4 template arguments on the drop-in headers) run on the GPU, every output
checked against the oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from mcpnet_amd import grn

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "dropin" / "callsites"
DEFAULT_NOISE_SEED = 1296388676612   # splash::kernel::random_number_generator<>'s default seed


def bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                                                 np.ascontiguousarray(b).view(np.uint8))


def test_reference_call_sites_on_gpu(tmp_path):
    assert EXE.exists(), "tests/dropin/callsites is not built (run __graft_entry__.build())"
    G, S = 300, 256
    X = grn.grn_expression(G, S, 0x5EED)[0]
    X.tofile(tmp_path / "X.f32")
    print(f"\ncallsites: {G} x {S}", flush=True)
    res = subprocess.run([str(EXE), str(tmp_path / "X.f32"), str(G), str(S), str(tmp_path)], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr

    def load(name, dt):
        return np.fromfile(tmp_path / name, dtype=dt).reshape(G, G)

    mi = load("mi.f32", np.float32)
    assert np.abs(mi.astype(np.float64) - oracle.bspline_mi_f64(X)).max() <= 1e-5
    assert bits_equal(mi, mi.T)
    ap = load("ap.f32", np.float32)
    R = grn.rank_desc(grn.jitter(X, DEFAULT_NOISE_SEED))
    assert np.abs(ap.astype(np.float64) - oracle.ap_mi(R)).max() <= 1e-6
    A = mi.copy()
    np.fill_diagonal(A, 0.0)
    for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
        Ad = A.astype(dt)
        for k in (2, 3, 4):
            ref_out, ref_mm = oracle.mcp(Ad, k)
            assert bits_equal(load(f"mm{k}.{tag}", dt), ref_mm), f"MCP{k} max-min {tag}"
            assert bits_equal(load(f"mcp{k}.{tag}", dt), ref_out), f"MCP{k} ratio {tag}"
