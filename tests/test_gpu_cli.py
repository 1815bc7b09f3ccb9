"""cuda-comp end to end: the reference's command line and .dist output.

`cuda-comp -T table -D G -o out.dist [-r F] [-d F]` with ParallelMETA pointing
at the golden database must write the same bytes as the reference `comp -M 0`
(cms_cuda.cu:1305-1342 = comp_sam.cpp:184-219 output format).
"""
import os
import subprocess

import pytest

from conftest import GOLDEN_CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("tag,extra", [("dist_rT", []), ("dist_rF", ["-r", "F"]), ("sim_rT", ["-d", "F"])])
def test_cuda_comp_matches_reference_bytes(cms, gpu_ok, golden, tmp_path, case, tag, extra):
    print(f"cuda-comp {case} {tag}", flush=True)
    g = golden(case)
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    out = str(tmp_path / "o.dist")
    env = dict(os.environ, ParallelMETA=g.root)
    r = subprocess.run([exe, "-T", g.table, "-D", "G", "-o", out, "-M", "0"] + extra, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert open(out, "rb").read() == g.read(tag + ".dist")


def test_cuda_comp_sharedmin_kernel_close(cms, gpu_ok, golden, tmp_path):
    """--kernel sharedmin: same layout, values within the 6-digit print of the walk."""
    g = golden("config0")
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    out = str(tmp_path / "o.dist")
    env = dict(os.environ, ParallelMETA=g.root)
    r = subprocess.run([exe, "-T", g.table, "-D", "G", "-o", out, "--kernel", "sharedmin"], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = [line.split("\t") for line in open(out).read().splitlines()]
    ref = [line.split("\t") for line in g.read("dist_rT.dist").decode().splitlines()]
    assert got[0] == ref[0] and len(got) == len(ref)
    worst = max(abs(float(a) - float(b)) for rg, rr in zip(got[1:], ref[1:]) for a, b in zip(rg[1:], rr[1:]))
    assert worst <= 2e-5


def test_cuda_comp_errors(cms, tmp_path):
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    r = subprocess.run([exe, "-T", "/nonexistent.txt", "-o", str(tmp_path / "x")],
                       env=dict(os.environ, ParallelMETA=str(tmp_path)), capture_output=True, text=True)
    assert r.returncode == 1 and "Error" in r.stderr
    r = subprocess.run([exe, "-Z", "1"], capture_output=True, text=True)
    assert "Unrec argument" in r.stderr


@pytest.mark.parametrize("case", ["config0", "core"])
@pytest.mark.parametrize("kernel,extra", [("walk", []), ("sharedmin", []), ("sharedmin", ["--stream-bands", "1000"])])
def test_cuda_comp_binary_output(cms, gpu_ok, golden, tmp_path, case, kernel, extra):
    """--format binary: packed float32 triangle of the values .dist prints.
    Walk: exactly (float)(1 - (double)sim) of the reference's Calc_sim; the
    shared-minimum kernel whole (binned) or streamed to the file in bands
    (<= 1e-5)."""
    import numpy as np
    print(f"cuda-comp binary {case} {kernel} {extra}", flush=True)
    g = golden(case)
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    out = str(tmp_path / "o.bin")
    env = dict(os.environ, ParallelMETA=g.root)
    r = subprocess.run([exe, "-T", g.table, "-D", "G", "-o", out, "--format", "binary", "--kernel", kernel] + extra,
                       env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    names = open(out + ".names").read().split("\n")[:-1]
    assert names == [ln.split("\t")[0] for ln in g.read("dist_rT.dist").decode().splitlines()[1:]]
    got = np.fromfile(out, dtype="<f4")
    i, j, s = g.hex_sims(True)
    n = len(names)
    assert got.size == n * (n - 1) // 2
    exp = (1.0 - s.astype(np.float64)).astype(np.float32)
    p = cms.packed_index(i, j, n)
    if kernel == "walk":
        np.testing.assert_array_equal(got[p].view(np.uint32), exp.view(np.uint32))
    else:
        assert np.abs(got[p].astype(np.float64) - exp).max() <= 1e-5


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("tag,extra", [("dist_rT", []), ("sim_rT", ["-d", "F"])])
def test_cuda_comp_sparse_table_input(cms, gpu_ok, golden, tmp_path, case, tag, extra):
    """The reference's table saved as a sparse table file: -T reads it and
    cuda-comp writes the reference's .dist bytes."""
    print(f"cuda-comp sparse input {case} {tag}", flush=True)
    g = golden(case)
    sparse = str(tmp_path / "table.cmst")
    cms.Table(g.table).save_csr(sparse)
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    out = str(tmp_path / "o.dist")
    env = dict(os.environ, ParallelMETA=g.root)
    r = subprocess.run([exe, "-T", sparse, "-D", "G", "-o", out, "-M", "0"] + extra, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert open(out, "rb").read() == g.read(tag + ".dist")
