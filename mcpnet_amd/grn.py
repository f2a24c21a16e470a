"""Seeded synthetic expression data: a sparse linear-Gaussian gene-regulatory network.

Generator (deterministic on every machine; numpy only):
  * random DAG over genes 0..G-1: gene g >= 1 takes min(3, g) distinct parents drawn uniformly
    from the genes below it;
  * edge weights U[-1, 1]; additive noise N(0, sd=0.1);
  * x_g = sum_p w_gp * x_p + eps_g per sample, then exp() (log-normal), then per-gene
    min-max normalisation to [0, 1], stored as float32.
Random numbers come from a counter-based SplitMix64 stream (Steele, Lea & Flood, OOPSLA 2014):
draw c of stream s under seed k is splitmix64(k, (s << 40) + c); uniforms take the top 53 bits.
The same generator, with stream index g*S+s, makes the U[0, 1e-5) rank jitter, so
`jitter()` here and the library's mcpg_add_noise_f32 produce identical float32 values.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

CONFIG_SEED_BASE = 0x4D43504E00

# Workloads (genes, samples) of the five benchmark configurations.
WORKLOADS = {
    1: dict(G=1000, S=1000, what="FP32 B-spline MI + MCP3"),
    2: dict(G=2000, S=1024, what="BF16 / BF16x3 B-spline MI + MCP2/3/4"),
    3: dict(G=20000, S=2000, what="BF16x3 B-spline MI + MCP3 + ratio"),
    4: dict(G=2000, S=1000, what="adaptive-partitioning MI"),
    5: dict(G=5000, S=500, what="B-spline MI adjoint"),
}


def config_seed(cid: int) -> int:
    return CONFIG_SEED_BASE + int(cid)


def splitmix64(seed: int, ctr) -> np.ndarray:
    """Counter-based SplitMix64: one 64-bit draw per counter value."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (np.asarray(ctr, dtype=np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, start: int, n: int) -> np.ndarray:
    """n uniforms in [0, 1) (float64) from counters start..start+n-1 of `stream`."""
    ctr = (np.uint64(stream) << np.uint64(40)) + np.arange(start, start + n, dtype=np.uint64)
    return (splitmix64(seed, ctr) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, stream: int, start: int, n: int) -> np.ndarray:
    """n standard normals by Box-Muller from counter pairs (2i, 2i+1)."""
    u = uniform(seed, stream, 2 * start, 2 * n)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def grn_expression(G: int, S: int, seed: int, parents: int = 3, noise_sd: float = 0.1):
    """Returns (X float32 G x S in [0,1], parent lists, weight lists)."""
    x = np.empty((G, S), dtype=np.float64)
    par = []
    wts = []
    pick = uniform(seed, 1, 0, G * 16).reshape(G, 16)
    wdraw = uniform(seed, 2, 0, G * parents).reshape(G, parents)
    for g in range(G):
        eps = noise_sd * normal(seed, 3, g * S, S)
        if g == 0:
            p = np.zeros(0, dtype=np.int64)
        elif g <= parents:
            p = np.arange(g, dtype=np.int64)[:parents]
        else:
            cand = np.floor(pick[g] * g).astype(np.int64)
            _, first = np.unique(cand, return_index=True)
            uniq = cand[np.sort(first)]
            if len(uniq) < parents:  # vanishingly rare: fill deterministically
                rest = np.setdiff1d(np.arange(g), uniq)[: parents - len(uniq)]
                uniq = np.concatenate([uniq, rest])
            p = uniq[:parents]
        w = 2.0 * wdraw[g, : len(p)] - 1.0
        x[g] = (w[:, None] * x[p]).sum(axis=0) + eps if len(p) else eps
        par.append(p)
        wts.append(w)
    x = np.exp(x)
    mn = x.min(axis=1, keepdims=True)
    mx = x.max(axis=1, keepdims=True)
    rng = np.where(mx > mn, mx - mn, 1.0)
    X = ((x - mn) / rng).astype(np.float32)
    return X, par, wts


def expression(cid: int) -> np.ndarray:
    """Expression matrix of benchmark configuration `cid` (1..5)."""
    c = WORKLOADS[cid]
    return grn_expression(c["G"], c["S"], config_seed(cid))[0]


def jitter(X: np.ndarray, seed: int) -> np.ndarray:
    """X + U[0, 1e-5) as float32, identical to mcpg_add_noise_f32 (element (g,s) uses counter g*S+s)."""
    G, S = X.shape
    ctr = np.arange(G * S, dtype=np.uint64)
    u = (splitmix64(seed, ctr) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    j = (u * 1e-5).astype(np.float32).reshape(G, S)
    return (X.astype(np.float32) + j).astype(np.float32)


def rank_desc(X: np.ndarray) -> np.ndarray:
    """Descending rank per row (largest -> 0), ties by ascending column index (int32)."""
    G, S = X.shape
    order = np.argsort(-X.astype(np.float64), axis=1, kind="stable")
    R = np.empty((G, S), dtype=np.int32)
    rows = np.arange(G)[:, None]
    R[rows, order] = np.arange(S, dtype=np.int32)[None, :]
    return R


def mask_pairs(G: int, prob: float, seed: int):
    """Bernoulli(prob) mask over unordered pairs i < j, with weights m = 1; returns (pi, pj, m)."""
    iu, ju = np.triu_indices(G, k=1)
    keep = uniform(seed, 4, 0, len(iu)) < prob
    pi = iu[keep].astype(np.int32)
    pj = ju[keep].astype(np.int32)
    return pi, pj, np.ones(len(pi), dtype=np.float32)
