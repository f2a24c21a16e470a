"""Profiling helpers (used with rocprofv3; nothing here is on the compute path).

  python -m mcpnet_amd.profiling copy [--mib N]
      a plain device-to-device copy of N MiB (torch), run 3 times: the reference workload for
      scaling FETCH_SIZE / WRITE_SIZE counters to bytes.
  python -m mcpnet_amd.profiling summarize --trace DIR [--pmc-fetch DIR] [--pmc-write DIR]
                                           [--copy-mib N] [--title T] [--command C]
      condenses rocprofv3 output (kernel_stats.csv, counter_collection.csv) into a short
      markdown table: calls and mean duration per kernel, and HBM KiB per dispatch.
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


def _find(d: str, name: str):
    hits = glob.glob(os.path.join(d, "**", f"*{name}"), recursive=True)
    return hits[0] if hits else None


def _short(name: str) -> str:
    for key in ("bspline_mi_kernel", "maxmin_tile_kernel", "bspline_prep_kernel", "ap_mi_kernel", "ratio_kernel",
                "symmetry_check_kernel", "rank_kernel", "cumhist_kernel", "noise_kernel", "weights_import_kernel",
                "bspline_pair_adjoint_kernel", "bspline_gene_adjoint_kernel", "emax_kernel"):
        if key in name:
            tmpl = name[name.find(key) + len(key):]
            return key + (tmpl.split("(")[0] if tmpl.startswith("<") else "")
    return name[:60]


def summarize(trace: str, fetch: str | None, write: str | None, copy_mib: int | None, title: str, command: str) -> str:
    lines = [f"# {title}", "", f"Command: `{command}`", ""]
    stats = _find(trace, "kernel_stats.csv") if trace else None
    if stats:
        lines += ["| kernel | calls | mean ms | total ms |", "|---|---|---|---|"]
        with open(stats) as f:
            for r in csv.DictReader(f):
                calls = int(r["Calls"])
                avg_ms = float(r["AverageNs"]) / 1e6
                lines.append(f"| {_short(r['Name'])} | {calls} | {avg_ms:.3f} | {float(r['TotalDurationNs']) / 1e6:.1f} |")
        lines.append("")
    for label, d, counter in (("FETCH_SIZE", fetch, "FETCH_SIZE"), ("WRITE_SIZE", write, "WRITE_SIZE")):
        if not d:
            continue
        path = _find(d, "counter_collection.csv")
        if not path:
            continue
        acc = defaultdict(list)
        with open(path) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name", counter) != counter:
                    continue
                acc[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        scale = None
        if copy_mib:
            copy_vals = [v for k, vs in acc.items() if "copy" in k.lower() or "elementwise" in k.lower() for v in vs]
            if copy_vals:
                scale = (copy_mib * 1024.0) / (sum(copy_vals) / len(copy_vals))
        lines += [f"## {label} (KiB per dispatch" + (f"; copy-kernel scale factor {scale:.3f})" if scale else ")"), "",
                  "| kernel | dispatches | mean KiB | scaled GB |", "|---|---|---|---|"]
        for k, vs in sorted(acc.items()):
            m = sum(vs) / len(vs)
            gb = m * 1024 * (scale or 1.0) / 1e9
            lines.append(f"| {k} | {len(vs)} | {m:.0f} | {gb:.3f} |")
        lines.append("")
    return "\n".join(lines)


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    s = sub.add_parser("summarize")
    s.add_argument("--trace")
    s.add_argument("--pmc-fetch")
    s.add_argument("--pmc-write")
    s.add_argument("--copy-mib", type=int)
    s.add_argument("--title", default="rocprofv3 summary")
    s.add_argument("--command", default="")
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_main(a.mib)
    else:
        sys.stdout.write(summarize(a.trace, a.pmc_fetch, a.pmc_write, a.copy_mib, a.title, a.command) + "\n")


if __name__ == "__main__":
    main()
