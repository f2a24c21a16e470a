"""Build recipe for libmcpgpu.so (HIP, gfx950) and the C++ command-line tools.

The library is compiled in-tree with hipcc; device code is built without fast-math (the
ratio and the spline arithmetic rely on IEEE division) and with packed-FP32 instruction
selection disabled.  Run `python -m mcpnet_amd.build` or call `build()`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
BINDIR = PKG / "bin"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["context.hip", "maxmin.hip", "bspline.hip", "apmi.hip", "ensemble.hip", "pipeline.hip"]
ARCH = os.environ.get("MCPG_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-pass-failed",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP toolchain is required to build libmcpgpu")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd[:3])} ...")
    noise = [ln for ln in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if verbose and noise:
        print("\n".join(noise))


def build_library(verbose: bool = False, jobs: int = 4) -> Path:
    """Compiles each HIP source to an object (in parallel) and links libmcpgpu.so."""
    hipcc = _hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [hipcc, *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((cmd, out))
    if failed:
        for cmd, out in failed:
            sys.stderr.write(out)
        raise RuntimeError("HIP compilation failed: " + ", ".join(Path(c[-3]).name for c, _ in failed))
    lib = LIBDIR / "libmcpgpu.so"
    tmp = LIBDIR / "libmcpgpu.so.tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs)], verbose)
    os.replace(tmp, lib)
    return lib


TOOLS = {"mi": "mi.cpp", "mcp": "mcp.cpp", "mcpnet": "mcpnet.cpp", "dpi": "dpi.cpp", "convert": "convert.cpp"}
HOST_ONLY_TOOLS = {"convert"}   # no GPU work: not linked against libmcpgpu

# HDF5 header for the tools' .h5 reader/writer; libhdf5 itself is opened at run time (dlopen)
HDF5_INCLUDE_DIRS = [os.environ.get("MCPG_HDF5_INCLUDE", ""), "/usr/include/hdf5/serial", "/usr/include",
                     "/opt/conda/include"]


def _hdf5_flags() -> list[str]:
    for d in HDF5_INCLUDE_DIRS:
        if d and (Path(d) / "hdf5.h").exists():
            # only the HDF5 headers are wanted from a prefix that holds other libraries' headers too
            return ["-idirafter", d]
    return ["-DMIO_NO_HDF5"]


def build_tools(verbose: bool = False) -> list[Path]:
    """Builds the C++ command-line tools (host code only, linked against libmcpgpu.so)."""
    tooldir = PKG / "tools"
    if not tooldir.exists():
        return []
    BINDIR.mkdir(exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    rocm_inc = "/opt/rocm/include"
    h5 = _hdf5_flags()
    out = []
    for name, src in TOOLS.items():
        if not (tooldir / src).exists():
            continue
        exe = BINDIR / name
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{tooldir}", f"-I{rocm_inc}", *h5,
               str(tooldir / src), "-o", str(exe)]
        if name not in HOST_ONLY_TOOLS:
            cmd += [f"-L{LIBDIR}", "-lmcpgpu", "-Wl,-rpath,$ORIGIN/../lib"]
        cmd += ["-ldl"]
        _run(cmd, verbose)
        out.append(exe)
    return out


def build(verbose: bool = False) -> Path:
    lib = build_library(verbose=verbose)
    build_tools(verbose=verbose)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
