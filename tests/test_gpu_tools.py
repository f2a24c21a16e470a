"""The mi / mcp / mcpnet command-line tools (C++, through the splash pattern shim and the C ABI),
run on the reference's data/r10c10.exp (copied to tests/golden/) and on seeded GRN data, checked
against the CPU oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from mcpnet_amd import grn

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "mcpnet_amd" / "bin"
R10 = Path(__file__).with_name("golden") / "r10c10.exp"
R10_H5 = Path(__file__).with_name("golden") / "r10c10.h5"
R128_H5 = Path(__file__).with_name("golden") / "r128c1000.h5"
R127_H5 = Path(__file__).with_name("golden") / "r127c999.h5"


def read_exp(path):
    rows = [ln.rstrip("\n").split("\t") for ln in Path(path).read_text().splitlines() if ln.strip()]
    return [r[0] for r in rows[1:]], np.array([[float(v) for v in r[2:]] for r in rows[1:]])


def write_exp(path, X, names=None):
    G, S = X.shape
    names = names or [f"g{i}" for i in range(G)]
    with open(path, "w") as f:
        f.write("Id\tAlias\t" + "\t".join(f"s{j}" for j in range(S)) + "\n")
        for i in range(G):
            f.write(names[i] + "\t---\t" + "\t".join(repr(float(v)) for v in X[i]) + "\n")


def run(tool, *args):
    exe = BIN / tool
    assert exe.exists(), f"{exe} is not built (python -m mcpnet_amd.build)"
    res = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


def h5_to_exp(path):
    out = Path(str(path) + ".exp")
    run("convert", "-i", path, "-o", out)
    return read_exp(out)


@pytest.mark.parametrize("src", [R10, R10_H5])
def test_mi_bspline_on_reference_data(tmp_path, src):
    out = tmp_path / "r10c10.mi.exp"
    run("mi", "-i", src, "-o", out, "-m", "0", "--single")
    _, mi = read_exp(out)
    _, X = read_exp(R10)
    # 10 genes x 10 samples: the shape also matches (samples x genes), so mi.cpp:95-96 transposes
    ref = oracle.bspline_mi_f64(np.ascontiguousarray(X.T).astype(np.float32))
    assert np.abs(mi - ref).max() <= 1e-5


@pytest.mark.parametrize("src", [R10, R10_H5])
def test_mi_bspline_double_default(tmp_path, src):
    """mi -m 0 without --single runs the reference's double DataType: scaling, weights and MI in
    double (the *_f64 entry points), within 1e-12 of the double-row oracle on the reference's data;
    --precision bf16x3 in the double run contracts FP32 images of the double weights."""
    out = tmp_path / "r10c10.mi.exp"
    run("mi", "-i", src, "-o", out, "-m", "0")
    _, mi = read_exp(out)
    _, X = read_exp(R10)
    ref = oracle.bspline_mi_x64(np.ascontiguousarray(X.T))
    assert np.abs(mi - ref).max() <= 1e-12
    run("mi", "-i", src, "-o", out, "-m", "0", "--precision", "bf16x3")
    assert np.abs(read_exp(out)[1] - ref).max() <= 1e-4


def test_mi_tool_rejects_fp64_with_single(tmp_path):
    exe = BIN / "mi"
    res = subprocess.run([str(exe), "-i", str(R10), "-o", str(tmp_path / "o.exp"), "-m", "0", "--single",
                          "--precision", "fp64"], capture_output=True, text=True, timeout=120)
    assert res.returncode != 0 and "double" in res.stderr


@pytest.mark.parametrize("precision", ["bf16x3", "i8x3"])
def test_mi_tool_split_precisions(tmp_path, precision):
    """mi --precision bf16x3 / i8x3 on 60 genes x 500 samples of GRN data, against the FP64 oracle."""
    X = grn.grn_expression(60, 500, seed=99)[0]
    src, out = tmp_path / "x.exp", tmp_path / "x.mi.exp"
    write_exp(src, X)
    run("mi", "-i", src, "-o", out, "-m", "0", "--single", "--precision", precision)
    _, mi = read_exp(out)
    assert np.abs(mi - oracle.bspline_mi_f64(X)).max() <= 1e-4


def test_mcpnet_on_reference_h5(tmp_path):
    """mcpnet -i data/r128c1000.h5 --mi-method 0 -m 2: the .h5 MI and MCP3 outputs against the oracle."""
    prefix = tmp_path / "r128"
    run("mcpnet", "-i", R128_H5, "-o", prefix, "--mi-method", "0", "-m", "2", "--single")
    X = h5_to_exp(R128_H5)[1].astype(np.float32)
    names, mi = h5_to_exp(str(prefix) + "_bs.mi.h5")
    assert names == [str(i) for i in range(128)]
    ref = oracle.bspline_mi_f64(X)
    np.fill_diagonal(ref, 0.0)                        # written after --diagonal 0 (mcpnet.cpp:299-356)
    assert np.abs(mi - ref).max() <= 1e-5
    A = mi.astype(np.float32)
    got = h5_to_exp(str(prefix) + "_bs.mcp3.h5")[1].astype(np.float32)
    assert np.array_equal(got, oracle.mcp(A, 3)[0])


def test_mcpnet_on_odd_sized_reference_h5(tmp_path):
    """mcpnet -i data/r127c999.h5 (127 genes x 999 samples: neither a tile nor a K-step multiple)
    --mi-method 0 -m 1 2 3: MI against the FP64 oracle, MCP2/3/4 bit-exact on the written MI."""
    prefix = tmp_path / "r127"
    run("mcpnet", "-i", R127_H5, "-o", prefix, "--mi-method", "0", "-m", "1", "2", "3", "--single")
    X = h5_to_exp(R127_H5)[1].astype(np.float32)
    assert X.shape == (127, 999)
    names, mi = h5_to_exp(str(prefix) + "_bs.mi.h5")
    assert len(names) == 127
    ref = oracle.bspline_mi_f64(X)
    np.fill_diagonal(ref, 0.0)
    assert np.abs(mi - ref).max() <= 1e-5
    A = mi.astype(np.float32)
    for k in (2, 3, 4):
        got = h5_to_exp(str(prefix) + f"_bs.mcp{k}.h5")[1].astype(np.float32)
        assert np.array_equal(got, oracle.mcp(A, k)[0]), k


def test_mi_adaptive_partitioning_double(tmp_path):
    X = grn.grn_expression(24, 300, 9)[0]
    src = tmp_path / "x.exp"
    write_exp(src, X.astype(np.float64))
    out = tmp_path / "ap.exp"
    run("mi", "-i", src, "-o", out, "-m", "1", "--noise-seed", "77")
    _, mi = read_exp(out)
    Xin = read_exp(src)[1].astype(np.float32)
    R = grn.rank_desc(grn.jitter(Xin, 77))
    assert np.abs(mi - oracle.ap_mi(R)).max() <= 1e-6
    # the double run keeps each pair's double (AdaptivePartitionRankMIKernel2<IT, double>)
    for i in range(24):
        for j in range(i, 24):
            assert abs(mi[i, j] - oracle.ap_mi_cells(R, i, j)[1]) <= 1e-12


@pytest.mark.parametrize("method,k", [(1, 2), (2, 3), (3, 4)])
def test_mcp_matches_oracle_bit_exact(tmp_path, method, k):
    X = grn.grn_expression(40, 200, 4)[0]
    A = oracle.bspline_mi_f64(X)           # double MI, as the reference's default DataType
    names = [f"g{i}" for i in range(40)]
    src = tmp_path / "mi.exp"
    write_exp(src, A, names)
    out = tmp_path / f"mcp{k}.exp"
    mm = tmp_path / f"mm{k}.exp"
    run("mcp", "-i", src, "-o", out, "-m", method, "--diagonal", "0", "-x", mm)
    Ain = read_exp(src)[1]
    np.fill_diagonal(Ain, 0.0)
    ref_out, ref_mm = oracle.mcp(Ain, k)
    assert np.array_equal(read_exp(out)[1], ref_out)
    assert np.array_equal(read_exp(mm)[1], ref_mm)


def test_mcp_from_partials(tmp_path):
    """-m 5 (mcp.cpp:618-655): max-min of the --first and --second partials, row . row, then the
    ratio against the MI input.  The second partial is read from --second (the reference re-reads
    --first, mcp.cpp:192); the partials differ here, so reading --first twice would fail."""
    X = grn.grn_expression(40, 200, 12)[0]
    A = oracle.bspline_mi_f64(X)
    np.fill_diagonal(A, 0.0)
    src = tmp_path / "mi.exp"
    write_exp(src, A)
    m1 = tmp_path / "m1.exp"
    run("mcp", "-i", src, "-o", tmp_path / "mcp2.exp", "-m", 1, "-x", m1)     # first partial: A (x) A
    out = tmp_path / "mcp_p.exp"
    mm = tmp_path / "mm_p.exp"
    run("mcp", "-i", src, "-o", out, "-m", 5, "--first", m1, "--second", src, "-x", mm)
    Ain = read_exp(src)[1]
    M1 = read_exp(m1)[1]
    ref_mm = oracle.maxmin(M1, Ain)
    assert not np.array_equal(ref_mm, oracle.maxmin(M1, M1))
    assert np.array_equal(read_exp(mm)[1], ref_mm)
    assert np.array_equal(read_exp(out)[1], oracle.ratio(Ain, ref_mm))
    tf_file = tmp_path / "tfs.txt"
    tf_file.write_text("g0\ng3\n")
    res = subprocess.run([str(BIN / "mcp"), "-i", str(src), "-o", str(out), "-m", "5", "--first", str(m1),
                          "--second", str(src), "-f", str(tf_file)], capture_output=True, text=True, timeout=120)
    assert res.returncode != 0 and "partials" in res.stdout + res.stderr


def test_mcp_tf_masked(tmp_path):
    X = grn.grn_expression(30, 200, 6)[0]
    A = oracle.bspline_mi_f64(X)
    np.fill_diagonal(A, 0.0)
    names = [f"g{i}" for i in range(30)]
    src = tmp_path / "mi.exp"
    write_exp(src, A, names)
    tfs = [0, 4, 7, 19]
    tf_file = tmp_path / "tfs.txt"
    tf_file.write_text("\n".join(names[t] for t in tfs) + "\nNOT_A_GENE\n")
    out = tmp_path / "mcp2tf.exp"
    run("mcp", "-i", src, "-o", out, "-m", "1", "-f", tf_file, "--tf-gene-transition", "1")
    got_names, got = read_exp(out)
    assert got_names == [names[t] for t in tfs]
    mask = np.full(30, np.finfo(np.float64).min)
    mask[tfs] = np.finfo(np.float64).max
    Ain = read_exp(src)[1]
    mm = oracle.maxmin(np.ascontiguousarray(Ain[tfs]), Ain, mask)      # tfs-tfs-genes (masked)
    assert np.array_equal(got, oracle.ratio(np.ascontiguousarray(Ain[tfs]), mm))


def test_mcpnet_pipeline(tmp_path):
    X = grn.grn_expression(32, 256, 12)[0]
    src = tmp_path / "x.exp"
    write_exp(src, X.astype(np.float64))
    prefix = tmp_path / "net"
    run("mcpnet", "-i", src, "-o", prefix, "--mi-method", "0", "-m", "1", "2", "--single", "--format", "exp")
    _, mi = read_exp(str(prefix) + "_bs.mi.exp")
    Xin = read_exp(src)[1].astype(np.float32)
    ref = oracle.bspline_mi_f64(Xin)
    np.fill_diagonal(ref, 0.0)
    assert np.abs(mi - ref).max() <= 1e-5
    A = mi.astype(np.float32)
    for k in (2, 3):
        got = read_exp(str(prefix) + f"_bs.mcp{k}.exp")[1].astype(np.float32)
        assert np.array_equal(got, oracle.mcp(A, k)[0])


def test_mcpnet_pipeline_double(tmp_path):
    """mcpnet --mi-method 0 in the default double run: the MI in double (within 1e-12 of the
    double-row oracle, diagonal 0) and the MCP scores of that MI, bit-exact against the oracle."""
    rng = np.random.default_rng(13)
    X = grn.grn_expression(40, 300, 13)[0].astype(np.float64) * (1.0 + 1e-9 * rng.standard_normal((40, 300)))
    src = tmp_path / "x.exp"
    write_exp(src, X)
    prefix = tmp_path / "net"
    run("mcpnet", "-i", src, "-o", prefix, "--mi-method", "0", "-m", "1", "2", "--format", "exp")
    _, mi = read_exp(str(prefix) + "_bs.mi.exp")
    ref = oracle.bspline_mi_x64(read_exp(src)[1])
    np.fill_diagonal(ref, 0.0)
    assert np.abs(mi - ref).max() <= 1e-12
    for k in (2, 3):
        got = read_exp(str(prefix) + f"_bs.mcp{k}.exp")[1]
        assert np.array_equal(got, oracle.mcp(mi, k)[0])


def test_tools_report_errors(tmp_path):
    exe = BIN / "mcp"
    res = subprocess.run([str(exe), "-i", str(tmp_path / "missing.exp"), "-o", str(tmp_path / "o.exp")],
                         capture_output=True, text=True, timeout=60)
    assert res.returncode != 0 and "cannot open" in res.stderr
    res = subprocess.run([str(BIN / "mi"), "-i", str(R10_H5), "-o", str(tmp_path / "o.h5"), "--h5-matrix-key", "x"],
                         capture_output=True, text=True, timeout=60)
    assert res.returncode != 0 and "no dataset /array/x" in res.stderr


@pytest.mark.parametrize("with_tfs", [False, True])
def test_dpi_tool(tmp_path, with_tfs):
    X = grn.grn_expression(36, 200, 15)[0]
    A = oracle.bspline_mi_f64(X)
    names = [f"g{i}" for i in range(36)]
    src = tmp_path / "mi.exp"
    write_exp(src, A, names)
    out = tmp_path / "dpi.exp"
    args = ["-i", src, "-o", out, "-l", "0.1"]
    tf_idx = [1, 5, 9, 30]
    if with_tfs:
        tf_file = tmp_path / "tfs.txt"
        tf_file.write_text("\n".join(names[t] for t in tf_idx) + "\n")
        args += ["-f", tf_file]
    stdout = run("dpi", *args)
    Ain = np.ascontiguousarray(read_exp(src)[1].T)   # square input: dpi.cpp:123-125 transposes it
    np.fill_diagonal(Ain, 0.0)                    # --diagonal defaults to 0
    tfs = None
    if with_tfs:
        tfs = np.full(36, np.finfo(np.float64).min)
        tfs[tf_idx] = np.finfo(np.float64).max
    want, removed = oracle.dpi(Ain, 0.1, tfs)
    got_names, got = read_exp(out)
    if with_tfs:
        assert got_names == [names[t] for t in tf_idx]
        want = want[tf_idx]
    assert np.array_equal(got, want)
    assert f"Mask removing {removed} elements" in stdout


def test_mcpnet_ensemble(tmp_path):
    X = grn.grn_expression(48, 256, 16)[0]
    A = oracle.bspline_mi_f64(X)
    names = [f"g{i}" for i in range(48)]
    src = tmp_path / "mi.exp"
    write_exp(src, A, names)
    cf = np.array([[0.25, 0.25, 0.25, 0.25], [1, 0, 0, 0], [0.5, 0, 0.5, 0]])
    cfile = tmp_path / "coeffs.exp"
    write_exp(cfile, cf, ["r0", "r1", "r2"])
    prefix = tmp_path / "ens"
    run("mcpnet", "-i", src, "-o", prefix, "--mi-input", "-m", "4", "-f", cfile, "--format", "exp")
    prefix = f"{prefix}_inmi"
    Ain = read_exp(src)[1]
    np.fill_diagonal(Ain, 0.0)
    M1, M2, M3 = oracle.maxmins(Ain)
    for row in (cf[0], cf[2]):
        tag = "_".join(f"{v:.3f}" for v in row)
        got = read_exp(f"{prefix}.mcp_{tag}.exp")[1]
        want = oracle.combo(Ain, M1, M2, M3, row)
        assert np.array_equal(got, want)
        gst = read_exp(f"{prefix}.mcpst_{tag}.exp")[1]
        assert np.allclose(gst, oracle.stouffer(want, True), rtol=1e-12, atol=1e-12)
    assert not Path(f"{prefix}.mcp_1.000_0.000_0.000_0.000.exp").exists()   # (1,0,0,0) is skipped


def test_mcpnet_tf_mode_on_reference_gene_and_tf_lists(tmp_path):
    """mcpnet --tf-input with the reference's own lists: the yeast TF names of
    data/infergs_yeast_tfs.txt against the first 600 gene names of data/gnw2000_names.txt (both
    copied to tests/golden/; most TF names are absent there), on a seeded symmetric MI-like
    matrix: TF rows and MCP2/3/4 bit-exact against the oracle's masked products."""
    gold = Path(__file__).with_name("golden")
    names = (gold / "gnw2000_names.txt").read_text().split()[:600]
    tf_list = (gold / "infergs_yeast_tfs.txt").read_text().split()
    tf_set = set(tf_list)
    tf_idx = [i for i, g in enumerate(names) if g in tf_set]
    rng = np.random.default_rng(5)
    A = rng.random((600, 600)) ** 3
    A = np.maximum(A, A.T)
    src = tmp_path / "mi.exp"
    write_exp(src, A, names)
    prefix = tmp_path / "gnw"
    out = run("mcpnet", "-i", src, "-o", prefix, "--mi-input", "--tf-input", gold / "infergs_yeast_tfs.txt",
              "-m", "1", "2", "3", "--format", "exp")
    assert f"TFs specified {len(tf_list)}, found {len(tf_idx)}" in out
    Ain = read_exp(src)[1]
    np.fill_diagonal(Ain, 0.0)
    mask = np.full(600, -np.finfo(np.float64).max)
    mask[tf_idx] = np.finfo(np.float64).max
    T = np.ascontiguousarray(Ain[tf_idx])
    mm = oracle.maxmin
    G1 = mm(Ain, Ain, mask)
    want = {2: mm(T, Ain, mask), 3: mm(T, G1, mask), 4: mm(mm(T, Ain, mask), G1, mask)}
    for k in (2, 3, 4):
        got_names, got = read_exp(f"{prefix}_tf_inmi.mcp{k}.exp")
        assert got_names == [names[t] for t in tf_idx]
        assert np.array_equal(got, oracle.ratio(T, want[k])), k


def test_mcpnet_tf_mode(tmp_path):
    """mcpnet --tf-input: MCP2/3/4 from the TF rows (compute_mcp{2,3,4}_tfs with transitions 1, 2, 3)
    and the TF ensemble (compute_maxmins_tfs, combo on the TF rows, compute_stouffer_tf), in
    double as the reference runs by default."""
    X = grn.grn_expression(40, 256, 21)[0]
    A = oracle.bspline_mi_f64(X)
    names = [f"g{i}" for i in range(40)]
    src = tmp_path / "mi.exp"
    write_exp(src, A, names)
    tf_idx = [2, 7, 11, 19, 33]
    tf_file = tmp_path / "tfs.txt"
    tf_file.write_text("\n".join(names[t] for t in tf_idx) + "\nnot_a_gene\n")
    cf = np.array([[0.25, 0.25, 0.25, 0.25], [0.5, 0, 0.5, 0]])
    cfile = tmp_path / "coeffs.exp"
    write_exp(cfile, cf, ["r0", "r1"])
    prefix = tmp_path / "net"
    out = run("mcpnet", "-i", src, "-o", prefix, "--mi-input", "--tf-input", tf_file, "-m", "1", "2", "3", "4",
              "-f", cfile, "--format", "exp")
    assert "TFs specified 6, found 5" in out
    Ain = read_exp(src)[1]
    np.fill_diagonal(Ain, 0.0)
    mask = np.full(40, -np.finfo(np.float64).max)
    mask[tf_idx] = np.finfo(np.float64).max
    T = np.ascontiguousarray(Ain[tf_idx])
    mm = oracle.maxmin
    G1 = mm(Ain, Ain, mask)                                  # gene-tf-gene
    want = {
        2: mm(T, Ain, mask),                                 # tf-tf-gene
        3: mm(T, G1, mask),                                  # tf-tf-tf-gene
        4: mm(mm(T, Ain, mask), G1, mask),                   # (tf-tf-gene)-tf-gene
    }
    base = f"{prefix}_tf_inmi"
    for k in (2, 3, 4):
        got_names, got = read_exp(f"{base}.mcp{k}.exp")
        assert got_names == [names[t] for t in tf_idx]
        assert np.array_equal(got, oracle.ratio(T, want[k])), k
    M1 = mm(T, Ain, mask)
    M2 = mm(M1, Ain, mask)
    M3 = mm(M1, G1, mask)
    for row in cf:
        tag = "_".join(f"{v:.3f}" for v in row)
        got = read_exp(f"{base}.mcp_{tag}.exp")[1]
        w = oracle.combo(T, M1, M2, M3, row)
        assert np.array_equal(got, w)
        gst = read_exp(f"{base}.mcpst_{tag}.exp")[1]
        assert np.allclose(gst, oracle.stouffer_tf(w), rtol=1e-12, atol=1e-12)
