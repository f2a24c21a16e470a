"""Matrix files of the command-line tools: the reference's EXP text format and its HDF5 files
(pandas fixed layout, data/r10c10.h5 and data/r128c1000.h5 copied to tests/golden/), read and
written by the host-only `convert` tool (apps/transform/convert.cpp).

Pins: data/r10c10.h5 and data/r10c10.exp hold the same matrix, so reading either must give the same
doubles; h5dump (an independent HDF5 reader, when installed) must print the values convert reads
from r128c1000.h5, and must print for a file convert writes exactly what it prints for the
reference's own file."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "mcpnet_amd" / "bin"
GOLDEN = Path(__file__).with_name("golden")
R10_EXP = GOLDEN / "r10c10.exp"
R10_H5 = GOLDEN / "r10c10.h5"
R128_H5 = GOLDEN / "r128c1000.h5"
H5DUMP = shutil.which("h5dump") or next((p for p in ("/opt/conda/bin/h5dump",) if Path(p).exists()), None)


def convert(*args, ok=True):
    exe = BIN / "convert"
    assert exe.exists(), f"{exe} is not built (python -m mcpnet_amd.build)"
    res = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=60)
    if ok:
        assert res.returncode == 0, res.stdout + res.stderr
    return res


def read_text(path, sep="\t", skip=2):
    rows = [ln.rstrip("\n").split(sep) for ln in Path(path).read_text().splitlines() if ln.strip()]
    return [r[0] for r in rows[1:]], rows[0][skip:], np.array([[float(v) for v in r[skip:]] for r in rows[1:]])


def h5dump_values(path, dataset):
    out = subprocess.run([H5DUMP, "-d", dataset, "-m", "%.17g", "-y", "-w", "0", str(path)], capture_output=True,
                         text=True, check=True).stdout
    body = out[out.index("DATA {") + len("DATA {"):]
    body = body[:body.index("}")]
    return np.array([float(v) for v in body.replace(",", " ").split()])


def h5dump_text(path):
    return subprocess.run([H5DUMP, str(path)], capture_output=True, text=True, check=True).stdout.split("\n", 1)[1]


def test_h5_reads_the_same_matrix_as_exp(tmp_path):
    out = tmp_path / "r10.exp"
    convert("-i", R10_H5, "-o", out)
    g, s, X = read_text(out)
    g0, s0, X0 = read_text(R10_EXP)
    assert g == g0 == [str(i) for i in range(10)]
    assert s == s0
    assert np.array_equal(X, X0)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump is not installed")
def test_h5_values_match_h5dump(tmp_path):
    out = tmp_path / "r128.exp"
    convert("-i", R128_H5, "-o", out)
    genes, samples, X = read_text(out)
    assert X.shape == (128, 1000)
    assert genes[:3] == ["0", "1", "2"] and samples[-1] == "999"
    assert np.array_equal(X.ravel(), h5dump_values(R128_H5, "/array/block0_values"))


@pytest.mark.skipif(H5DUMP is None, reason="h5dump is not installed")
@pytest.mark.parametrize("src", ["r10c10.exp", "r10c10.h5", "r128c1000.h5"])
def test_written_h5_matches_reference_file_layout(tmp_path, src):
    """Every dataset, attribute, type and value of a written file is what pandas wrote."""
    out = tmp_path / "w.h5"
    convert("-i", GOLDEN / src, "-o", out)
    ref = R128_H5 if src.startswith("r128") else R10_H5
    assert h5dump_text(out) == h5dump_text(ref)


@pytest.mark.parametrize("fmt", ["h5", "exp", "csv", "tsv"])
def test_round_trip_is_exact(tmp_path, fmt):
    a = tmp_path / f"a.{fmt}"
    convert("-i", R128_H5, "-o", a)
    b = tmp_path / "b.exp"
    convert("-i", a, "-o", b)
    c = tmp_path / "c.exp"
    convert("-i", R128_H5, "-o", c)
    assert read_text(b)[2].tobytes() == read_text(c)[2].tobytes()


def test_single_precision_h5(tmp_path):
    a = tmp_path / "a.h5"
    convert("-i", R128_H5, "-o", a, "--single")
    if H5DUMP:
        assert "H5T_IEEE_F32LE" in subprocess.run([H5DUMP, "-H", "-d", "/array/block0_values", str(a)],
                                                  capture_output=True, text=True, check=True).stdout
    b = tmp_path / "b.exp"
    convert("-i", a, "-o", b)
    c = tmp_path / "c.exp"
    convert("-i", R128_H5, "-o", c)
    assert np.array_equal(read_text(b)[2], read_text(c)[2].astype(np.float32).astype(np.float64))


def test_h5_key_options(tmp_path):
    # names default to 0, 1, ... when the name datasets are not there
    out = tmp_path / "o.csv"
    convert("-i", R10_H5, "-o", out, "--h5-gene-key", "nope", "--h5-samples-key", "nope2")
    g, s, X = read_text(out, sep=",", skip=1)
    assert g == [str(i) for i in range(10)] and s == [str(i) for i in range(10)]
    assert np.array_equal(X, read_text(R10_EXP)[2])
    # the values dataset must exist
    res = convert("-i", R10_H5, "-o", tmp_path / "x.exp", "--h5-matrix-key", "missing", ok=False)
    assert res.returncode != 0 and "no dataset /array/missing" in res.stderr
    res = convert("-i", R10_H5, "-o", tmp_path / "x.exp", "--h5-group", "other", ok=False)
    assert res.returncode != 0 and "no dataset /other/block0_values" in res.stderr
    res = convert("-i", tmp_path / "absent.h5", "-o", tmp_path / "x.exp", ok=False)
    assert res.returncode != 0 and "cannot open" in res.stderr
