"""Profiling helpers (used with rocprofv3; nothing here is on the compute path).

  python -m mcpnet_amd.profiling copy [--mib N]
      a plain device-to-device copy of N MiB (torch), run 3 times: the reference workload for
      scaling FETCH_SIZE / WRITE_SIZE counters to bytes.
  python -m mcpnet_amd.profiling summarize --trace DIR [--pmc-fetch DIR] [--pmc-write DIR]
                                           [--copy-mib N] [--title T] [--command C]
      condenses rocprofv3 output (kernel_stats.csv, counter_collection.csv) into a short
      markdown table: calls and mean duration per kernel, and HBM KiB per dispatch.
  python -m mcpnet_amd.profiling traffic --pmc-fetch DIR --pmc-write DIR [--pmc-*-copy DIR]
                                         [--config N --update FILE]
      scaled HBM bytes per dispatch (FETCH + WRITE) per kernel, printed as JSON or merged into
      FILE under the key config<N> (bench.py reads profiles/traffic.json that way).
  python -m mcpnet_amd.profiling maxmin --genes N [--dtype f32|f64] [--reps R] [--mcp K] [--float-exact]
      max-min products of a seeded symmetric N x N matrix with diagonal 0 on the GPU (the
      triangle launch, R times, or MCP_K): the workload for kernel traces of the product alone;
      prints each launch's duration from the library's own event timing.
  python -m mcpnet_amd.profiling mi --genes N --samples S [--precision P] [--reps R]
      B-spline MI (10 bins, order 3) of seeded synthetic expression on the GPU, R times: the
      workload for kernel traces of the MI kernel alone; prints each launch's duration.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sys
from collections import defaultdict


def copy_main(mib: int) -> None:
    import torch

    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x 3", flush=True)


def maxmin_main(genes: int, dtype: str, reps: int, mcp: int, float_exact: bool = False) -> None:
    import numpy as np
    import torch

    import mcpnet_amd

    dt = np.float64 if dtype == "f64" else np.float32
    rng = np.random.default_rng(0x4D4D)
    A = rng.random((genes, genes), dtype=np.float64)
    if float_exact:   # FP32 values held in doubles, as mcpnet's double chain hands its MI on
        A = A.astype(np.float32)
    A = A.astype(dt)
    A = np.maximum(A, A.T)
    np.fill_diagonal(A, 0.0)
    h = mcpnet_amd.Handle(0)
    h.set_profiling(True)
    Ad = torch.from_numpy(A).cuda()
    C = torch.empty_like(Ad)
    for r in range(reps):
        if mcp:
            h.mcp(Ad, mcp, out=C)
        else:
            h.maxmin(Ad, Ad, symmetric=True, out=C)
        for name, ms in h.last_timings():
            print(f"{genes} genes {dtype} rep {r}: {name} {ms:.3f} ms", flush=True)
    h.close()


def mi_main(genes: int, samples: int, precision: str, reps: int, seed: int = 0x4D49) -> None:
    import torch

    import mcpnet_amd
    from mcpnet_amd import grn

    X = grn.grn_expression(genes, samples, seed)[0]
    h = mcpnet_amd.Handle(0)
    h.set_profiling(True)
    Xd = torch.from_numpy(X).cuda()
    MI = torch.empty((genes, genes), dtype=torch.float32, device="cuda")
    for prec in precision.split(","):
        x, out = Xd, MI
        if prec == "fp64":   # double rows and MI (the values of the same expression matrix)
            x, out = Xd.double(), torch.empty((genes, genes), dtype=torch.float64, device="cuda")
        for r in range(reps):
            h.bspline_mi(x, precision=prec, out=out)
            for name, ms in h.last_timings():
                print(f"{genes} x {samples} {prec} rep {r}: {name} {ms:.3f} ms", flush=True)
    h.close()


def _db(d: str):
    import sqlite3

    hits = glob.glob(os.path.join(d, "**", "*results.db"), recursive=True) if d else []
    return sqlite3.connect(hits[0]) if hits else None


def _short(name: str) -> str:
    for key in ("bspline_mi_kernel", "bspline_mi64_kernel", "bspline_prep64_kernel", "weights_import64_kernel",
                "maxmin_key_kernel", "maxmin_tile_kernel", "key_prep_kernel",
                "key_prep_f64_kernel", "key_pad_kernel",
                "bspline_prep_kernel", "ap_mi_kernel", "ratio_kernel", "symmetry_check_kernel", "rank_kernel",
                "cumhist_kernel", "noise_kernel", "weights_import_kernel", "bspline_pair_adjoint_kernel",
                "bspline_gene_adjoint_kernel", "emax_kernel", "combo_kernel", "stouffer_kernel", "row_stats_kernel",
                "zrow_moments_kernel", "zcol_moments_kernel", "stouffer_tf_kernel",
                "dpi_kernel", "copyBuffer", "fillBuffer"):
        if key in name:
            tmpl = name[name.find(key) + len(key):]
            if key == "maxmin_tile_kernel":
                return key + ("<f64>" if "TrF64" in tmpl else "<f32>")
            if key == "maxmin_key_kernel":   # <KC, WIDE>: float or widened double outputs
                return key + ("<f64 out>" if tmpl.split("(")[0].rstrip(">").endswith("true") else "<f32>")
            if key == "bspline_mi_kernel":
                return key + {"0": "<fp32>", "1": "<bf16>", "2": "<bf16x3>"}.get(tmpl[1:2], "")
            if key == "bspline_prep_kernel":   # <NB, NO, TWO, I8>: BF16 two-order or int8 planes
                flags = tmpl.split("(")[0]
                return key + ("<int8>" if flags.rstrip(">").endswith("true") else
                              "<bf16 two orders>" if "true" in flags else "")
            return key
    if name.startswith("void at::") or name.startswith("at::"):
        return name[:60]
    # other kernels of this library: drop the return type, namespaces and the argument list
    base = name[5:] if name.startswith("void ") else name
    base = base.replace("(anonymous namespace)::", "").split("(")[0].split("<")[0]
    return base.rsplit("::", 1)[-1][:60]


def _counter(d: str, counter: str):
    db = _db(d)
    if db is None:
        return {}
    acc = defaultdict(list)
    for kn, cn, v in db.execute("select kernel_name, counter_name, value from counters_collection"):
        if cn == counter:
            acc[_short(kn)].append(float(v))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def _busy(d: str):
    """Per kernel: mean duration (ms), clock (GHz) and MFMA-busy fraction from a rocprofv3 --pmc
    GRBM_GUI_ACTIVE SQ_VALU_MFMA_BUSY_CYCLES run.  On MI355X GRBM_GUI_ACTIVE is summed over the 8
    XCDs and SQ_VALU_MFMA_BUSY_CYCLES over the 1024 SIMDs (256 CUs x 4)."""
    db = _db(d)
    if db is None:
        return {}
    per = defaultdict(dict)
    names = {}
    for disp, kn, cn, v, dur in db.execute(
            "select dispatch_id, kernel_name, counter_name, value, duration from counters_collection"):
        per[disp][cn] = per[disp].get(cn, 0.0) + float(v)
        per[disp]["dur"] = float(dur)
        names[disp] = _short(kn)
    acc = defaultdict(list)
    for disp, r in per.items():
        if "GRBM_GUI_ACTIVE" not in r or r["dur"] <= 0:
            continue
        cyc = r["GRBM_GUI_ACTIVE"] / 8.0
        busy = r.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / 1024.0 / cyc if cyc else 0.0
        acc[names[disp]].append((r["dur"] / 1e6, cyc / (r["dur"] / 1e9) / 1e9, busy))
    return {k: tuple(sum(x[i] for x in v) / len(v) for i in range(3)) for k, v in acc.items()}


def summarize(trace, fetch, write, fetch_copy, write_copy, copy_mib, title, command, busy=None) -> str:
    """Markdown summary of rocprofv3 databases (rocpd SQLite output)."""
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", ""]
    db = _db(trace)
    if db is not None:
        lines += ["## Kernel trace (rocprofv3 --kernel-trace --stats; durations from the rocpd database, in ms)", "",
                  "| kernel | calls | mean ms | total ms | share |", "|---|---|---|---|---|"]
        for name, calls, total, avg, pct in db.execute(
                "select name, total_calls, total_duration, average, percentage from top_kernels"):
            lines.append(f"| {_short(name)} | {calls} | {avg / 1e3:.3f} | {total / 1e3:.1f} | {pct:.1f}% |")
        lines.append("")
    scale = {}
    for label, d, dc in (("FETCH_SIZE", fetch, fetch_copy), ("WRITE_SIZE", write, write_copy)):
        if not d:
            continue
        vals = _counter(d, label)
        s = 1.0
        if dc and copy_mib:
            cv = _counter(dc, label).get("copyBuffer")
            if cv:
                s = copy_mib * 1024.0 / cv
        scale[label] = s
        lines += [f"## {label} (rocprofv3 --pmc {label}), KiB per dispatch", "",
                  f"Scale factor from a {copy_mib} MiB device copy (`python -m mcpnet_amd.profiling copy`): "
                  f"{s:.3f} (bytes moved / counter).", "",
                  "| kernel | mean KiB | scaled GB per dispatch |", "|---|---|---|"]
        for k, v in sorted(vals.items()):
            lines.append(f"| {k} | {v:.0f} | {v * 1024 * s / 1e9:.3f} |")
        lines.append("")
    if busy:
        vals = _busy(busy)
        lines += ["## Clock and MFMA pipe (rocprofv3 --pmc GRBM_GUI_ACTIVE SQ_VALU_MFMA_BUSY_CYCLES)", "",
                  "Clock = GRBM_GUI_ACTIVE / 8 XCDs / duration; MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES / 1024 SIMDs / "
                  "(GRBM_GUI_ACTIVE / 8).", "",
                  "| kernel | mean ms (counter run) | clock GHz | MFMA busy |", "|---|---|---|---|"]
        for k, (ms, ghz, b) in sorted(vals.items()):
            if b > 0:
                lines.append(f"| {k} | {ms:.3f} | {ghz:.2f} | {b:.3f} |")
        lines.append("")
    return "\n".join(lines)


def traffic(fetch, write, fetch_copy, write_copy, copy_mib) -> dict:
    """Scaled HBM bytes per dispatch (FETCH + WRITE) per kernel."""
    out = defaultdict(float)
    for label, d, dc in (("FETCH_SIZE", fetch, fetch_copy), ("WRITE_SIZE", write, write_copy)):
        vals = _counter(d, label)
        s = 1.0
        cv = _counter(dc, label).get("copyBuffer") if dc else None
        if cv and copy_mib:
            s = copy_mib * 1024.0 / cv
        for k, v in vals.items():
            out[k] += v * 1024 * s
    return dict(out)


def main(argv=None) -> None:
    import json

    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    m = sub.add_parser("maxmin")
    m.add_argument("--genes", type=int, default=2000)
    m.add_argument("--dtype", choices=["f32", "f64"], default="f64")
    m.add_argument("--reps", type=int, default=3)
    m.add_argument("--mcp", type=int, default=0, choices=[0, 2, 3, 4])
    m.add_argument("--float-exact", action="store_true", help="f64: values a float holds exactly (the key route)")
    b = sub.add_parser("mi")
    b.add_argument("--genes", type=int, default=20000)
    b.add_argument("--samples", type=int, default=2000)
    b.add_argument("--precision", default="bf16x3", help="fp32, bf16, bf16x3, i8x3 or fp64, or a comma-separated list")
    b.add_argument("--seed", type=lambda v: int(v, 0), default=0x4D49, help="GRN generator seed")
    b.add_argument("--reps", type=int, default=3)
    for name in ("summarize", "traffic"):
        s = sub.add_parser(name)
        s.add_argument("--trace")
        s.add_argument("--pmc-fetch")
        s.add_argument("--pmc-write")
        s.add_argument("--pmc-fetch-copy")
        s.add_argument("--pmc-write-copy")
        s.add_argument("--copy-mib", type=int, default=1024)
        s.add_argument("--pmc-busy", help="summarize: run with GRBM_GUI_ACTIVE and SQ_VALU_MFMA_BUSY_CYCLES")
        s.add_argument("--title", default="rocprofv3 summary")
        s.add_argument("--command", default="")
        s.add_argument("--config", type=int, help="traffic: workload number the values are filed under")
        s.add_argument("--update", help="traffic: JSON file to merge the values into, keyed config<N>")
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_main(a.mib)
    elif a.cmd == "maxmin":
        maxmin_main(a.genes, a.dtype, a.reps, a.mcp, a.float_exact)
    elif a.cmd == "mi":
        mi_main(a.genes, a.samples, a.precision, a.reps, a.seed)
    elif a.cmd == "summarize":
        sys.stdout.write(summarize(a.trace, a.pmc_fetch, a.pmc_write, a.pmc_fetch_copy, a.pmc_write_copy, a.copy_mib,
                                   a.title, a.command, a.pmc_busy) + "\n")
    else:
        t = traffic(a.pmc_fetch, a.pmc_write, a.pmc_fetch_copy, a.pmc_write_copy, a.copy_mib)
        if a.update and a.config is not None:
            try:
                with open(a.update) as f:
                    doc = json.load(f)
            except (OSError, ValueError):
                doc = {}
            doc[f"config{a.config}"] = {**doc.get(f"config{a.config}", {}), **t}
            with open(a.update, "w") as f:
                json.dump(doc, f, indent=1)
                f.write("\n")
        else:
            print(json.dumps(t, indent=1))


if __name__ == "__main__":
    main()
