"""Writes tests/golden/maxmin_ref.npz: outputs of the reference's own max_of_pairmin
(include/mcp/filter/simd_maxmin.hpp, compiled into oracle/_ref by oracle/Makefile) on
  * the reference's data/r10c10.exp expression matrix (10 genes x 10 samples), plain and
    TF-masked (mask = FLT_MAX for genes 0, 3, 7 and lowest otherwise, as mcp.cpp builds it),
  * a seeded 96 x 130 matrix with negative values and +-inf entries (f32 and f64).
Run in the build container (needs the reference tree):  python tests/golden/make_golden.py
"""
from pathlib import Path
import sys

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle  # noqa: E402

REF_DATA = Path("/root/reference/data/r10c10.exp")


def read_exp(path):
    rows = [ln.rstrip("\n").split("\t") for ln in path.read_text().splitlines() if ln.strip()]
    return np.array([[float(v) for v in r[2:]] for r in rows[1:]], dtype=np.float64)


def main():
    if oracle.ref_lib() is None:
        oracle.build()
    X64 = read_exp(REF_DATA)
    X32 = X64.astype(np.float32)
    tf = np.zeros(X32.shape[0], dtype=bool)
    tf[[0, 3, 7]] = True
    m32 = np.where(tf, np.finfo(np.float32).max, np.finfo(np.float32).min).astype(np.float32)
    m64 = np.where(tf, np.finfo(np.float64).max, np.finfo(np.float64).min)
    rng = np.random.default_rng(20240611)
    R = rng.standard_normal((96, 130)).astype(np.float32)
    R[rng.random(R.shape) < 0.01] = np.inf
    R[rng.random(R.shape) < 0.01] = -np.inf
    out = dict(
        r10c10_f32=X32, r10c10_f64=X64, tf_mask_f32=m32, tf_mask_f64=m64,
        mm_r10c10_f32=oracle.ref_maxmin(X32, X32), mm_r10c10_f64=oracle.ref_maxmin(X64, X64),
        mm_r10c10_masked_f32=oracle.ref_maxmin(X32, X32, m32), mm_r10c10_masked_f64=oracle.ref_maxmin(X64, X64, m64),
        rand_f32=R, mm_rand_f32=oracle.ref_maxmin(R[:96], R[:71]),
        rand_f64=R.astype(np.float64), mm_rand_f64=oracle.ref_maxmin(R.astype(np.float64)[:96], R.astype(np.float64)[:71]),
    )
    dst = Path(__file__).with_name("maxmin_ref.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, dst.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
