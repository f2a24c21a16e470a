#!/usr/bin/env python3
"""Benchmark of the MCPNet hot path on one MI355X.

Workload (default): 20000 genes x 2000 samples of the seeded GRN generator.  One step is the
pipeline the reference runs in mcpnet (apps/pipeline/mcpnet.cpp:286-515): all-pairs B-spline MI
(10 bins, order 3) with split-BF16 MFMA contraction, the MI diagonal set to 0, then MCP3 =
ratio(MI, MI (x) (MI (x) MI)) with two max-min products.  Inputs are resident in HBM before the
timed region; stage times come from HIP events on the library's own stream, per-kernel launch
durations from the library's event bracketing of each launch.

Prints ONE JSON line.  value = MI gene-pairs/s of the MI stage; mcp3_maxmin_gops = 1.6e13
algorithmic max-min ops / MCP3 stage time (BASELINE's unit: one (min, max) per (i, k, j) triple,
M*N*K per product); mcp3_maxmin_gops_executed counts only the triples the triangle launches
compute (about half).  The default run then times the other workloads briefly (1000 x 1000 FP32
MI + MCP3, 2000 x 1024 BF16X3 MI + MCP2/3/4 + ensemble + DPI, 2000 x 1000 AP-MI, 5000 x 500 MI
adjoint; three steps each, no CPU legs) and reports them under "secondary_configs".
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# MI355X peaks (vendor figures): HBM3E 8 TB/s; dense MFMA BF16 2.5 PFLOP/s, FP32 157.3 TFLOP/s.
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP32_TFLOPS = 157.3
PEAK_INT8_TOPS = 5000.0
PEAK_FP64_TFLOPS = 78.6
# VALU issue: 256 CUs x 64 lanes x 2.4 GHz; the max-min inner loop spends 1.5 instructions per
# (i, k, j) triple (v_min_i32 per triple + v_max3_i32 per two triples).
CUS, LANES, CLOCK_GHZ = 256, 64, 2.4
PEAK_MAXMIN_GOPS = CUS * LANES * CLOCK_GHZ / 1.5
# AP-MI point tests: one compare, one ballot and one popcount per 64 points, i.e. 3 lane
# instructions per test of the reference's quadrant scans
PEAK_AP_GTESTS = CUS * LANES * CLOCK_GHZ / 3.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def omp_threads():
    try:
        return int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    except ValueError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


SIMD_FLAGS = ("sse4_2", "avx", "avx2", "fma", "avx512f", "avx512bw", "avx512vl", "avx512dq", "avx512_bf16",
              "avx512_vnni", "amx_bf16")


def cpu_simd_flags():
    """The SIMD extensions of the host CPU the baselines may use (from /proc/cpuinfo)."""
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("flags"):
                have = set(ln.split(":", 1)[1].split())
                return [f for f in SIMD_FLAGS if f in have]
    except OSError:
        pass
    return []


def cpu_host():
    return {"model": cpu_model(), "simd": cpu_simd_flags(), "threads": omp_threads()}


def cpu_baselines(X, S, bins, mi_host, budget_s):
    """Bounded CPU samples: the FP32 B-spline MI port (reference inner product) and the
    reference's own max_of_pairmin (oracle/_ref) on the same data."""
    import oracle

    G = X.shape[0]
    out = {}
    # ---- B-spline MI: rows [0, R) against genes [0, C) of the same workload (same S, bins)
    C = min(G, 2000)
    W = oracle.bspline_weights_f32(oracle.minmax_scale_f32(X[:C]), bins=bins, order=3)
    rows, t = 4, 0.0
    while True:   # grow the sample until it takes about the budget
        t0 = time.perf_counter()
        oracle.port_bspline_mi_rows_f32(W, S, bins, 0, rows)
        t = time.perf_counter() - t0
        if t >= 0.5 * budget_s or rows >= C:
            break
        rows = int(min(C, rows * max(2.0, 0.9 * budget_s / max(t, 1e-3))))
    out["mi"] = {
        "value": rows * C / t, "unit": "gene-pairs/s", "cores": omp_threads(), "kind": "port", "host": cpu_host(),
        "sample": f"FP32 B-spline MI inner product (100 dots of S={S} + log2 per pair, OpenMP) on "
                  f"{rows} x {C} gene pairs of the same data, {t:.1f} s; host {cpu_model()}",
    }
    # ---- max-min: rows [0, R) of MI (x) MI with the reference's max_of_pairmin
    if mi_host is not None and oracle.ref_lib() is not None:
        A = mi_host
        N = A.shape[0]
        r = 4
        while True:
            t0 = time.perf_counter()
            oracle.ref_maxmin(A[:r], A)
            t = time.perf_counter() - t0
            if t >= 0.5 * budget_s or r >= N:
                break
            r = int(min(N, r * max(2.0, 0.9 * budget_s / max(t, 1e-3))))
        out["maxmin"] = {
            "value": r * N * N / t / 1e9, "unit": "Gop/s", "cores": omp_threads(), "kind": "reference",
            "host": cpu_host(),
            "sample": f"reference max_of_pairmin (simd_maxmin.hpp, float = scalar path) over rows 0..{r - 1} "
                      f"of MI (x) MI, N = K = {N}, {t:.1f} s",
        }
    return out


def _timed_steps(args, fn, stream, torch):
    for i in range(args.warmup):
        log(f"bench: warmup step {i + 1}/{args.warmup}")
        fn()
    torch.cuda.synchronize()
    ms = []
    for i in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
        log(f"bench: step {i + 1}/{args.steps} done")
    return float(np.mean(ms))


def bench_ap(args, G, S, seed):
    """Config 4: jitter + descending rank + adaptive-partitioning MI on the GPU."""
    import torch

    import mcpnet_amd
    from mcpnet_amd import grn

    log(f"bench: generating {G} x {S} GRN expression (seed {seed:#x})")
    X = grn.grn_expression(G, S, seed)[0]
    h = mcpnet_amd.Handle(0)
    h.set_profiling(True)
    stream = torch.cuda.ExternalStream(h.stream)
    Xd = torch.from_numpy(X).cuda()
    J = torch.empty_like(Xd)
    R = torch.empty((G, S), dtype=torch.int32, device="cuda")
    MI = torch.empty((G, G), dtype=torch.float32, device="cuda")
    kern = {}

    def step():
        for call in (lambda: h.add_noise(Xd, seed, out=J), lambda: h.rank(J, out=R), lambda: h.ap_mi(R, out=MI)):
            call()
            for name, t in h.last_timings():
                kern.setdefault(name, []).append(t)

    ms = _timed_steps(args, step, stream, torch)
    P = G * (G - 1) / 2
    per = {n: float(np.mean(v)) for n, v in kern.items()}
    result = {
        "metric": f"adaptive-partitioning MI gene-pairs/s @{G} genes", "value": P / (ms / 1e3),
        "unit": "gene-pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32 counts, f64 sums",
        "data": "synthetic: seeded GRN expression + U[0,1e-5) jitter, descending int32 ranks",
        "config": {"workload": f"config 4: {G} genes x {S} samples: jitter -> rank -> AP-MI (chi-square 7.815)",
                   "genes": G, "samples": S, "parallelism": "single GPU"},
        "kernel_mean_ms": per,
    }
    if not args.no_cpu:
        import oracle

        Rh = R.cpu().numpy()
        ok = bool(np.array_equal(Rh, grn.rank_desc(grn.jitter(X, seed))))
        rows = 2
        t0 = time.perf_counter()
        ref = oracle.ap_mi_rows(Rh, 0, rows)
        tp = time.perf_counter() - t0
        rows = int(min(G, max(2, rows * args.cpu_seconds / max(tp, 1e-3))))
        t0 = time.perf_counter()
        ref = oracle.ap_mi_rows(Rh, 0, rows)
        t = time.perf_counter() - t0
        got = MI[:rows].cpu().numpy()
        result["cpu_baseline"] = {"value": rows * G / t, "unit": "gene-pairs/s", "cores": omp_threads(), "kind": "port",
                                  "host": cpu_host(),
                                  "sample": f"AdaptivePartitionRankMIKernel2 restatement (OpenMP) on rows 0..{rows - 1} "
                                            f"x {G} genes, {t:.1f} s; host {cpu_model()}"}
        result["ranks_bit_exact"] = ok
        result["mi_max_abs_diff_vs_oracle"] = float(np.abs(got.astype(np.float64) - ref).max())
        result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
        # algorithmic work: the points the reference's quadrant scans compare, counted by the
        # oracle on the sampled rows and scaled to all G (G + 1) / 2 evaluated pairs
        tests = oracle.ap_tests_rows(Rh, 0, rows)
        work = tests / (rows * G) * G * (G + 1) / 2
        ap_ms = per.get("ap_mi_kernel")
        if ap_ms:
            ach = work / (ap_ms / 1e3) / 1e9
            result["roofline"] = {"bound": "valu", "achieved": ach, "peak": PEAK_AP_GTESTS, "unit": "Gtests/s",
                                  "frac": ach / PEAK_AP_GTESTS, "traffic": None, "mean_ms": ap_ms,
                                  "work_per_launch": work, "kernel": "ap_mi_kernel",
                                  "work_unit": f"point comparisons of the reference's quadrant scans "
                                               f"({tests / (rows * G):.0f} per pair on rows 0..{rows - 1})"}
    h.close()
    return result


def bench_adjoint(args, G, S, seed):
    """Config 5: dL/dX of L = sum over a 1% Bernoulli pair mask of MI, with W recomputed."""
    import torch

    import mcpnet_amd
    from mcpnet_amd import grn

    log(f"bench: generating {G} x {S} GRN expression (seed {seed:#x})")
    X = grn.grn_expression(G, S, seed)[0]
    pi, pj, m = grn.mask_pairs(G, 0.01, seed + 1)
    h = mcpnet_amd.Handle(0)
    h.set_profiling(True)
    stream = torch.cuda.ExternalStream(h.stream)
    Xd = torch.from_numpy(X).cuda()
    dX = torch.empty_like(Xd)
    kern = {}
    holder = {}

    def step():
        holder["L"] = h.bspline_mi_backward(Xd, pi, pj, m, out=dX)[1]
        for name, t in h.last_timings():
            kern.setdefault(name, []).append(t)

    ms = _timed_steps(args, step, stream, torch)
    result = {
        "metric": f"B-spline MI adjoint: masked gene-pairs/s @{G} genes", "value": len(pi) / (ms / 1e3),
        "unit": "masked gene-pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 marginals)",
        "data": "synthetic: seeded GRN expression, seeded 1% Bernoulli pair mask",
        "config": {"workload": f"config 5: {G} genes x {S} samples, {len(pi)} masked pairs: forward recompute + "
                               f"pair tables + gather GEMM through spline derivatives",
                   "genes": G, "samples": S, "pairs": int(len(pi)), "parallelism": "single GPU"},
        "kernel_mean_ms": {n: float(np.mean(v)) for n, v in kern.items()}, "loss": holder["L"],
        "cpu_baseline": None,
        "cpu_baseline_note": "the reference has no adjoint",
    }
    if not args.no_cpu:
        import oracle

        g = dX.cpu().numpy()
        rng = np.random.default_rng(0)
        errs = []
        X64 = X.astype(np.float64)
        scale = float(np.sqrt(np.mean(g.astype(np.float64) ** 2)))
        for _ in range(50):
            a, b = int(rng.integers(0, G)), int(rng.integers(0, S))
            sel = (pi == a) | (pj == a)
            if not sel.any():
                continue
            qi, qj, qm = pi[sel], pj[sel], m[sel]
            Xp = X64.copy(); Xp[a, b] += 1e-3
            Xm = X64.copy(); Xm[a, b] -= 1e-3
            fd = (oracle.bspline_mi_loss_x64(Xp, qi, qj, qm) - oracle.bspline_mi_loss_x64(Xm, qi, qj, qm)) / 2e-3
            errs.append(abs(fd - g[a, b]) / max(abs(fd), scale))
        errs = np.array(errs)
        result["fd_check"] = {"entries": int(len(errs)), "median_rel_err": float(np.median(errs)),
                              "max_rel_err": float(errs.max()), "h": 1e-3}
    h.close()
    return result


def bench_ensemble_dpi(args, h, MI, stream, torch):
    """The ensemble MCP (compute_maxmins, compute_combo at the default coefficients, compute_stouffer)
    and the DPI filter (tolerance 0.1) on the MI matrix of the workload (diagonal 0), timed with
    HIP events on the library's stream; checked against the oracle unless --no-cpu."""
    G = MI.shape[0]
    cf = np.full(4, 0.25, np.float32)
    out = {}

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            r = fn()
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        return float(np.mean(ms)), r

    t_mm, (M1, M2, M3) = timed(lambda: h.maxmins(MI))
    t_cb, R = timed(lambda: h.combo(MI, M1, M2, M3, cf))
    t_st, Z = timed(lambda: h.stouffer(R))
    t_dpi, (D, removed) = timed(lambda: h.dpi(MI, 0.1))
    ops = 3.0 * G ** 3
    out["ensemble"] = {"maxmins_ms": t_mm, "combo_ms": t_cb, "stouffer_ms": t_st,
                       "maxmins_gops": ops / (t_mm / 1e3) / 1e9,
                       "work": "M1 = A(x)A, M2 = A(x)M1, M3 = M1(x)M1 (3 G^3 algorithmic max-min ops), combination "
                               "ratio at coefficients 0.25 x 4, Stouffer of the ratio"}
    out["dpi"] = {"ms": t_dpi, "removed": int(removed), "tolerance": 0.1,
                  "work": "one max-min triangle product with the diagonal excluded + mask application"}
    if not args.no_cpu:
        import oracle

        A = MI.cpu().numpy()
        r1, r2, r3 = oracle.maxmins(A)
        out["ensemble"]["maxmins_bit_exact"] = bool(all(np.array_equal(g.cpu().numpy(), w)
                                                        for g, w in zip((M1, M2, M3), (r1, r2, r3))))
        rw = oracle.combo(A, r1, r2, r3, cf)
        out["ensemble"]["combo_bit_exact"] = bool(np.array_equal(R.cpu().numpy(), rw))
        out["ensemble"]["stouffer_max_abs_diff"] = float(np.abs(Z.cpu().numpy().astype(np.float64)
                                                                - oracle.stouffer(rw)).max())
        t0 = time.perf_counter()
        dw, rm = oracle.dpi(A, 0.1)
        tc = time.perf_counter() - t0
        out["dpi"]["bit_exact"] = bool(np.array_equal(D.cpu().numpy(), dw) and rm == removed)
        out["dpi"]["cpu_baseline"] = {"value_ms": tc * 1e3, "cores": omp_threads(), "kind": "port",
                                      "sample": f"the reference's per-pair early-exit DPI scan (oracle restatement, "
                                                f"OpenMP) on the same {G} x {G} matrix; host {cpu_model()}"}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3, choices=[1, 2, 3, 4, 5],
                    help="workload: 3 = 20000 x 2000 B-spline MI + MCP3 (the headline, default); 1 = 1000 x 1000 "
                         "FP32 MI + MCP3; 2 = 2000 x 1024 MI + MCP2/3/4; 4 = 2000 x 1000 adaptive-partitioning MI; "
                         "5 = 5000 x 500 B-spline MI adjoint")
    ap.add_argument("--genes", type=int, default=None)
    ap.add_argument("--samples", type=int, default=None)
    ap.add_argument("--precision", default=None, choices=["fp32", "bf16", "bf16x3", "i8x3"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of each bounded CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 MI comparison pass")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 (double DataType) MI pass")
    ap.add_argument("--extras", action="store_true", help="also time the ensemble MCP and the DPI filter "
                                                          "(always on for config 2)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="config 3 only: skip the short timed runs of configs 1, 2, 4 and 5 that are reported under "
                         "'secondary_configs'")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path (the products would "
              "shard by row panels with an all-gather of M1's keys); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    result = run_config(args, args.config)
    if args.config == 3 and not args.no_secondary and args.genes is None and args.samples is None:
        # the other workloads, timed briefly (no CPU legs or oracle checks: their parity is in the
        # GPU tests), so that one default run records every config's throughput
        sec = {}
        for cid in (1, 2, 4, 5):
            sub = argparse.Namespace(**vars(args))
            sub.steps, sub.warmup, sub.precision = 3, 1, None
            sub.no_cpu, sub.no_fp32, sub.no_fp64, sub.extras = True, True, True, False
            log(f"bench: secondary config {cid}")
            r = run_config(sub, cid)
            sec[f"config{cid}"] = {k: r[k] for k in ("metric", "value", "unit", "ms_per_step", "dtype", "config",
                                                     "stage_ms", "kernel_mean_ms", "roofline_kernels",
                                                     "mcp3_maxmin_gops", "ensemble", "dpi") if k in r}
        result["secondary_configs"] = sec
    print(json.dumps(result))


def mi_kernel_name(precision):
    return "bspline_mi_i8_kernel" if precision == "i8x3" else f"bspline_mi_kernel<{precision}>"


def mi_roofline(precision, flops, ms):
    """Algorithmic work (2*B^2*S per unordered pair; the split products are not credited) per
    launch over the launch's mean duration, against the dense MFMA peak of the arithmetic used."""
    peak, unit = {"fp32": (PEAK_FP32_TFLOPS, "TFLOP/s"), "i8x3": (PEAK_INT8_TOPS, "TOP/s"),
                  "fp64": (PEAK_FP64_TFLOPS, "TFLOP/s")}.get(precision, (PEAK_BF16_TFLOPS, "TFLOP/s"))
    ach = flops / (ms / 1e3) / 1e12
    return {"bound": "mfma", "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak, "traffic": None,
            "mean_ms": ms, "work_per_launch": flops,
            "work_unit": "algorithmic multiply-adds x 2 (2*B^2*S per unordered gene pair; split products not credited)"}


def run_config(args, cid):
    """Runs one workload and returns its JSON record."""
    import torch

    import mcpnet_amd
    from mcpnet_amd import grn

    G = args.genes or grn.WORKLOADS[cid]["G"]
    S = args.samples or grn.WORKLOADS[cid]["S"]
    bins = 10
    if args.precision is None:
        args.precision = "fp32" if cid == 1 else "bf16x3"
    seed = grn.config_seed(cid)
    if cid == 4:
        return bench_ap(args, G, S, seed)
    if cid == 5:
        return bench_adjoint(args, G, S, seed)
    mcp_orders = [2, 3, 4] if cid == 2 else [3]
    log(f"bench: generating {G} x {S} GRN expression (seed {seed:#x})")
    t0 = time.perf_counter()
    X = grn.grn_expression(G, S, seed)[0]
    log(f"bench: data ready in {time.perf_counter() - t0:.1f} s")

    dev = torch.device("cuda:0")
    h = mcpnet_amd.Handle(0)
    h.set_profiling(True)
    stream = torch.cuda.ExternalStream(h.stream)
    Xd = torch.from_numpy(X).to(dev)
    MI = torch.empty((G, G), dtype=torch.float32, device=dev)
    MCP = torch.empty((G, G), dtype=torch.float32, device=dev)
    MM = torch.empty((G, G), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    kern = {}

    def step(record):
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record(stream)
        h.bspline_mi(Xd, bins=bins, precision=args.precision, out=MI)
        e1.record(stream)
        if record:
            for name, ms in h.last_timings():
                kern.setdefault(name, []).append(ms)
        MI.fill_diagonal_(0.0)   # mcpnet.cpp:299-308, default diagonal 0
        for k in mcp_orders:
            h.mcp(MI, k, out=MCP, maxmin_out=MM)
            if record:
                for name, ms in h.last_timings():
                    kern.setdefault(name, []).append(ms)
        e2.record(stream)
        return e0, e1, e2

    for i in range(args.warmup):
        log(f"bench: warmup step {i + 1}/{args.warmup}")
        step(False)
    torch.cuda.synchronize()
    h.synchronize()
    events = []
    t_wall = time.perf_counter()
    for i in range(args.steps):
        events.append(step(True))
        log(f"bench: step {i + 1}/{args.steps} done")
    torch.cuda.synchronize()
    h.synchronize()
    t_wall = time.perf_counter() - t_wall
    t_mi = np.mean([a.elapsed_time(b) for a, b, _ in events]) / 1e3
    t_mcp = np.mean([b.elapsed_time(c) for _, b, c in events]) / 1e3
    t_step = np.mean([a.elapsed_time(c) for a, _, c in events]) / 1e3

    P = G * (G - 1) / 2
    mi_rate = P / t_mi
    # algorithmic max-min ops (M*N*K per product): MCP2 one product, MCP3 and MCP4 two
    mcp3_ops = sum({2: 1.0, 3: 2.0, 4: 2.0}[k] for k in mcp_orders) * G * G * G
    mcp3_gops = mcp3_ops / t_mcp / 1e9

    # ---- per-kernel rooflines (algorithmic work per launch / mean launch duration)
    def mean_ms(name):
        v = kern.get(name)
        return float(np.mean(v)) if v else None

    mi_name = mi_kernel_name(args.precision)
    mi_ms = mean_ms(mi_name)
    mi_flops = 2.0 * bins * bins * S * P          # 2*B^2*S per unordered pair
    roof = {}
    if mi_ms:
        roof[mi_name] = mi_roofline(args.precision, mi_flops, mi_ms)
    for name, tri in (("maxmin_key_kernel<f32,tri>", True), ("maxmin_key_kernel<f32>", False)):
        ms = mean_ms(name)
        if not ms:
            continue
        work = (G * (G + 1) / 2 if tri else G * G) * G
        ach = work / (ms / 1e3) / 1e9
        roof[name] = {"bound": "valu", "achieved": ach, "peak": PEAK_MAXMIN_GOPS, "unit": "Gop/s",
                      "frac": ach / PEAK_MAXMIN_GOPS, "traffic": None, "mean_ms": ms, "work_per_launch": work,
                      "work_unit": "(min,max) triples over the output the launch must produce"
                                   + (" (upper triangle, mirrored)" if tri else "")}
    # dominant kernel = largest time per step
    per_step = {n: float(np.sum(v)) / args.steps for n, v in kern.items()}
    dominant = max(per_step, key=per_step.get) if per_step else None
    # HBM bytes per launch: FETCH_SIZE and WRITE_SIZE of the same kernel on the same workload
    # (rocprofv3 --pmc runs, copy-calibrated), kept in profiles/traffic.json
    traffic_file = ROOT / "profiles" / "traffic.json"
    tr = {}
    if traffic_file.exists():
        try:
            tr = json.loads(traffic_file.read_text()).get(f"config{cid}", {})
        except (OSError, ValueError):
            tr = {}

    def attach_traffic(n, r):
        t = tr.get(n, tr.get(n.replace(",tri", "")))   # the rocprof names carry no ",tri" launch mode
        if t is not None:
            r["traffic"] = t
            r["traffic_source"] = "profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"
        r["hbm_gbs"] = r["traffic"] / (r["mean_ms"] / 1e3) / 1e9 if r.get("traffic") else None
        r["hbm_frac"] = r["hbm_gbs"] / PEAK_HBM_GBS if r["hbm_gbs"] else None

    for n, r in roof.items():
        attach_traffic(n, r)
    # max-min work the launches executed: a triangle launch produces G (G + 1) / 2 outputs and
    # mirrors them, a full launch G^2, each over K = G
    mm_launches = {n: len(v) / args.steps for n, v in kern.items() if n.startswith("maxmin_key_kernel")}
    mcp_ops_executed = sum(c * (G * (G + 1) / 2 if n.endswith(",tri>") else G * G) * G
                           for n, c in mm_launches.items())

    # ---- FP32 MI pass for the BF16-split / FP32 speed ratio and accuracy on sampled pairs
    extra = {}
    if not args.no_fp32 and args.precision != "fp32":
        MI32 = torch.empty((G, G), dtype=torch.float32, device=dev)
        h.bspline_mi(Xd, bins=bins, precision="fp32", out=MI32)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        h.bspline_mi(Xd, bins=bins, precision="fp32", out=MI32)
        e1.record(stream)
        torch.cuda.synchronize()
        t32 = e0.elapsed_time(e1) / 1e3
        extra["mi_fp32_pairs_per_s"] = P / t32
        extra["mi_speedup_vs_fp32"] = t32 / t_mi
        del MI32
    i8 = None
    if args.precision == "bf16x3" and bins == 10:
        # the int8 two-digit split on the same matrix: kernel time, pairs/s and (CPU leg) its error
        MI8 = torch.empty((G, G), dtype=torch.float32, device=dev)
        h.bspline_mi(Xd, bins=bins, precision="i8x3", out=MI8)
        ms8, t8 = [], []
        for _ in range(max(2, min(args.steps, 5))):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            h.bspline_mi(Xd, bins=bins, precision="i8x3", out=MI8)
            e1.record(stream)
            ms8 += [ms for n, ms in h.last_timings() if n == mi_kernel_name("i8x3")]
            torch.cuda.synchronize()
            t8.append(e0.elapsed_time(e1) / 1e3)
        t8 = float(np.mean(t8))
        i8 = {"pairs_per_s": P / t8, "stage_ms": t8 * 1e3, "speedup_vs_bf16x3": t_mi / t8,
              "dtype": "two int8 digits per weight, w ~= (h + l/254)/127; h*h and h*l + l*h on the int8 MFMA, "
                       "exact int32 sums",
              "roofline": mi_roofline("i8x3", mi_flops, float(np.mean(ms8))) if ms8 else None}
        if i8["roofline"]:
            attach_traffic(mi_kernel_name("i8x3"), i8["roofline"])
        extra["mi_i8x3"] = i8
        del MI8
    f64 = None
    if args.precision == "bf16x3" and not args.no_fp64:
        # the reference's double DataType on the same matrix (its values as doubles): scaling,
        # basis, marginals and joint tables in double on the FP64 MFMA
        X64 = Xd.double()
        MI64 = torch.empty((G, G), dtype=torch.float64, device=dev)
        h.bspline_mi(X64, bins=bins, precision="fp64", out=MI64)
        ms64, t64 = [], []
        for _ in range(1 if G > 5000 else 3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            h.bspline_mi(X64, bins=bins, precision="fp64", out=MI64)
            e1.record(stream)
            ms64 += [ms for n, ms in h.last_timings() if n == "bspline_mi64_kernel"]
            torch.cuda.synchronize()
            t64.append(e0.elapsed_time(e1) / 1e3)
        t64 = float(np.mean(t64))
        f64 = {"pairs_per_s": P / t64, "stage_ms": t64 * 1e3, "bf16x3_speedup": t64 / t_mi,
               "dtype": "double throughout (the reference's default double DataType): v_mfma_f64_16x16x4_f64",
               "roofline": mi_roofline("fp64", mi_flops, float(np.mean(ms64))) if ms64 else None}
        if f64["roofline"]:
            attach_traffic("bspline_mi64_kernel", f64["roofline"])
        extra["mi_fp64"] = f64

    cpu = {}
    if not args.no_cpu:
        # CPU leg: the oracle is the checker (MI error vs FP64 on sampled pairs) and the timed
        # CPU baselines; nothing here feeds the GPU numbers above.
        import oracle

        log("bench: accuracy check against the FP64 oracle")
        rng = np.random.default_rng(7)
        genes = np.sort(rng.choice(G, min(G, 200), replace=False))
        li, lj = np.triu_indices(len(genes), k=1)
        h.bspline_mi(Xd, bins=bins, precision=args.precision, out=MI)
        got = MI[torch.from_numpy(genes[li]).long().to(dev), torch.from_numpy(genes[lj]).long().to(dev)].cpu().numpy()
        ref = oracle.bspline_mi_pairs_f64(X[genes], li.astype(np.int32), lj.astype(np.int32))
        err = np.abs(got.astype(np.float64) - ref)
        extra["mi_err_vs_fp64"] = {"max": float(err.max()), "mean": float(err.mean()),
                                   "pairs": int(len(li)), "sample": "all pairs of 200 seeded random genes"}
        if i8 is not None:
            h.bspline_mi(Xd, bins=bins, precision="i8x3", out=MI)
            got8 = MI[torch.from_numpy(genes[li]).long().to(dev),
                      torch.from_numpy(genes[lj]).long().to(dev)].cpu().numpy()
            e8 = np.abs(got8.astype(np.float64) - ref)
            i8["err_vs_fp64"] = {"max": float(e8.max()), "mean": float(e8.mean()), "pairs": int(len(li)),
                                 "sample": "the same pairs"}
        if f64 is not None:
            got64 = MI64[torch.from_numpy(genes[li]).long().to(dev),
                         torch.from_numpy(genes[lj]).long().to(dev)].cpu().numpy()
            e64 = np.abs(got64 - ref)
            f64["err_vs_fp64"] = {"max": float(e64.max()), "mean": float(e64.mean()), "pairs": int(len(li)),
                                  "sample": "the same pairs"}
        if G <= 2000:   # small configs: every pair, every precision
            full = oracle.bspline_mi_f64(X)
            errs = {}
            for prec in ("fp32", "bf16", "bf16x3") + (("i8x3",) if bins == 10 else ()):
                h.bspline_mi(Xd, bins=bins, precision=prec, out=MI)
                e = np.abs(MI.cpu().numpy().astype(np.float64) - full)
                errs[prec] = {"max": float(e.max()), "mean": float(e.mean())}
            extra["mi_err_vs_fp64_all_pairs"] = errs
            MI.fill_diagonal_(0.0)
            A = MI.cpu().numpy()
            for k in mcp_orders:
                out, mm = h.mcp(MI, k, want_maxmin=True)
                ro, rm = oracle.mcp(A, k)
                extra[f"mcp{k}_bit_exact"] = bool(np.array_equal(out.cpu().numpy(), ro)
                                                  and np.array_equal(mm.cpu().numpy(), rm))
        log("bench: CPU baselines")
        MI.fill_diagonal_(0.0)
        cpu = cpu_baselines(X, S, bins, MI.cpu().numpy(), args.cpu_seconds)
    if cid == 2 or args.extras:
        log("bench: ensemble MCP and DPI")
        MI.fill_diagonal_(0.0)
        extra.update(bench_ensemble_dpi(args, h, MI, stream, torch))

    result = {
        "metric": ("MI gene-pairs/s (split-BF16 B-spline MI) & MCP3 max-min Gop/s @20k genes" if cid == 3 else
                   f"MI gene-pairs/s ({args.precision} B-spline MI) & MCP{'/'.join(map(str, mcp_orders))} "
                   f"max-min Gop/s @{G} genes"),
        "value": mi_rate,
        "unit": "gene-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (f"{args.precision} MFMA (fp32 accumulate) MI; f32 max-min" if args.precision != "i8x3" else
                  "i8x3 int8 MFMA (exact int32 sums) MI; f32 max-min"),
        "data": "synthetic: seeded linear-Gaussian GRN (3 parents/gene), exp, per-gene min-max",
        "config": {"workload": f"config {cid}: {G} genes x {S} samples: B-spline MI (10 bins, order 3, "
                               f"{args.precision}) -> diagonal 0 -> "
                               + ", ".join(f"MCP{k}" for k in mcp_orders) + " (max-min products + ratio)",
                   "genes": G, "samples": S, "bins": bins, "spline_order": 3, "parallelism": "single GPU"},
        "mcp3_maxmin_gops": mcp3_gops,
        "mcp3_maxmin_gops_executed": mcp_ops_executed / t_mcp / 1e9,
        "skipped_work": (f"mcp3_maxmin_gops credits G^3 = {G ** 3:.3g} (min, max) triples per product (BASELINE's unit). "
                         f"A, A(x)A and A(x)(A(x)A) are symmetric, so a triangle launch (',tri') computes only the upper "
                         f"triangle of its output and mirrors it: {mcp_ops_executed:.3g} executed triples per step "
                         f"({', '.join(f'{c:g} x {n}' for n, c in mm_launches.items())}); "
                         f"mcp3_maxmin_gops_executed is that count / MCP stage time"),
        "stage_ms": {"mi": t_mi * 1e3, "mcp3": t_mcp * 1e3},
        "wall_s_timed_region": t_wall,
        "roofline": dict(roof.get(dominant, {}), kernel=dominant) if dominant in roof else None,
        "roofline_kernels": roof,
        "kernel_ms_per_step": per_step,
        "cpu_baseline": cpu.get("mi"),
        "cpu_baseline_maxmin": cpu.get("maxmin"),
        **extra,
    }
    if cpu.get("mi"):
        result["speedup_vs_cpu_mi"] = mi_rate / cpu["mi"]["value"]
    if cpu.get("maxmin"):
        result["speedup_vs_cpu_maxmin"] = mcp3_gops / cpu["maxmin"]["value"]
    h.close()
    return result


if __name__ == "__main__":
    main()
