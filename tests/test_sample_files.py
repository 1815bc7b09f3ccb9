"""Single-file inputs of `cuda-comp -i` / `-l` (Single_Comp / Multi_Comp,
cms_cuda.cu:1344-1403): classification files loaded by
_Comp_Tree::Load_abd(const char *, float *, bool) (comp.h:194-236) over
_OTU_Parser::Load_file_to_hash (otu_parser.h:104-157), and sample lists read
by Load_List (utility.h:344-391).

CPU: the loader's abundances equal the compiled reference's bit for bit on
every golden file (count and per-read formats, the loader's quirks, hash-map
growth); list ids and files match the reference's .dist headers; the oracle's
pair values over the loaded samples, written by cms_write_dist, reproduce the
reference's `comp -l` files byte for byte and its `comp -i` output.
GPU: `cuda-comp -l` / `-i` give the same bytes (walk) or values within 1e-5.
"""
import os
import subprocess

import numpy as np
import pytest

LISTS = {  # list -> (-p prefix, [(tag, metric, cp_correct, is_sim)])
    "ids": ("", [("M0", 0, True, False), ("rF", 0, False, False), ("M1", 1, True, False), ("M4", 4, True, False)]),
    "paths": ("", [("sim", 0, True, True)]),
    "prefix": ("samples/", [("M0", 0, True, False), ("M5", 5, True, False)]),
}
PAIRS = [  # tag, file1, file2, metric, cp_correct, is_sim
    ("M0", "samples/S0/classification.txt", "samples/S1/classification.txt", 0, True, False),
    ("sim_rF", "samples/S2/classification.txt", "samples/R0/classification.txt", 0, False, True),
    ("M1", "samples/S3/classification.txt", "samples/R1/classification.txt", 1, True, False),
    ("M4", "samples/S4/classification.txt", "samples/S5/classification.txt", 4, True, False),
    ("M5", "samples/R2/classification.txt", "samples/S6/classification.txt", 5, True, False),
]


@pytest.fixture(scope="module")
def files(golden):
    return golden("files")


def _ref_abd(g, cp, leaf_n):
    """abd_rT / abd_rF: per file, the mapped count and the dense abundances."""
    out = []
    for ln in g.read("abd_rT.txt" if cp else "abd_rF.txt").decode().splitlines():
        if ln.startswith("#"):
            _, k, m = ln.split()
            out.append((int(m), np.zeros(leaf_n, np.float32)))
        else:
            leaf, h = ln.split()
            out[-1][1][int(leaf)] = float.fromhex(h)
    return out


def _all_files(g):
    return [ln for ln in open(os.path.join(g.root, "files_all.txt")).read().splitlines() if ln]


@pytest.mark.parametrize("cp", [True, False])
def test_loader_bit_exact(cms, files, cp):
    db = cms.Database(files.root, "G")
    ref = _ref_abd(files, cp, db.leaf_n)
    paths = [os.path.join(files.root, p) for p in _all_files(files)]
    assert len(ref) == len(paths)
    for (m, a), p in zip(ref, paths):
        got, gm = cms.sample_file_abundance(db, p, cp)
        assert gm == m, p
        np.testing.assert_array_equal(got.view(np.uint32), a.view(np.uint32), err_msg=p)
    # the CSR form: the same values, zeros dropped, mapped counts alike
    s, mapped = cms.sample_files_csr(db, paths, cp)
    np.testing.assert_array_equal(mapped, [m for m, _ in ref])
    dense = s.dense(db.leaf_n)
    for k, (_, a) in enumerate(ref):
        np.testing.assert_array_equal(dense[k].view(np.uint32), a.view(np.uint32))


def _parse_counts(path):
    """The count format read naively, in file order (no hash map)."""
    lines = open(path, newline="").read().replace("\r", "").split("\n")
    counts = {}
    for ln in lines[1:]:
        f = ln.split()
        if not f:
            continue
        c = 0
        if len(f) > 1:
            digits = ""
            for ch in f[1]:
                if ch.isdigit() or (not digits and ch in "+-"):
                    digits += ch
                else:
                    break
            c = int(digits) if digits.strip("+-") else 0
        counts[f[0]] = c
    return counts


def test_total_order_matters(cms, files):
    """A planted mistake: the FP32 total summed in file order instead of the
    reference's hash-map order changes some abundance bits, so the bit-exact
    comparison above pins the order."""
    db = cms.Database(files.root, "G")
    ref = _ref_abd(files, True, db.leaf_n)
    leaf_of = {db.leaf_name(i): i for i in range(db.leaf_n)}
    cp = {}
    with open(os.path.join(files.root, "databases", "files", "copy_number.txt")) as f:
        for ln in f.read().splitlines()[1:]:
            k, v = ln.split()[:2]
            cp.setdefault(k, np.float32(v))
    differs = 0
    for k, rel in enumerate(_all_files(files)):
        if "/S" not in rel:
            continue
        counts = _parse_counts(os.path.join(files.root, rel))
        x = {key: np.float32(np.float32(c) / cp.get(key, np.float32(1))) for key, c in counts.items()}
        total = np.float32(0)
        for v in x.values():
            total = np.float32(total + v)
        total = np.float32(np.float64(total) / 100.0)
        a = np.zeros(db.leaf_n, np.float32)
        for key, v in x.items():
            if key in leaf_of:
                a[leaf_of[key]] = np.float32(v / total)
        assert np.abs(a - ref[k][1]).max() < 1e-5      # same values ...
        differs += int(np.any(a.view(np.uint32) != ref[k][1].view(np.uint32)))
    assert differs > 0                                 # ... but not the same bits


def test_list_names_and_paths(cms, files):
    for name, (prefix, runs) in LISTS.items():
        lst = cms.SampleList(os.path.join(files.root, f"list_{name}.txt"), prefix)
        hdr = files.read(f"l_{name}_{runs[0][0]}.dist").decode().split("\n")[0].split("\t")[1:]
        assert lst.names == hdr
        for p in lst.paths:
            assert os.path.exists(os.path.join(files.root, p)), p
    lst = cms.SampleList(os.path.join(files.root, "list_paths.txt"))
    assert lst.names == ["S3", "S1", "S4", "S1", "S5", "named", "S8"]
    assert lst.paths[-1] == "./samples/S8/classification.txt"
    with pytest.raises(cms.CmsError):
        cms.SampleList(os.path.join(files.root, "no_such_list.txt"))


def _oracle_sims(orc, tree, abd, metric):
    n = abd.shape[0]
    i, j = np.triu_indices(n, 1)
    i, j = i.astype(np.int32), j.astype(np.int32)
    if metric == 0:
        return orc.pairs_f32(tree, abd, i, j)
    if metric == 1:
        return orc.pairs_unweight(tree, abd, i, j)[0]
    return orc.pairs_leaf(metric, abd, i, j)[0]


def test_list_dist_bytes_via_oracle(cms, orc, files, tmp_path):
    """comp -l end to end on the CPU: list -> loader -> oracle pair values
    (FP32 as the reference) -> .dist writer equals the reference's bytes."""
    db = cms.Database(files.root, "G")
    tree = db.tree()
    for name, (prefix, runs) in LISTS.items():
        lst = cms.SampleList(os.path.join(files.root, f"list_{name}.txt"), prefix)
        paths = [os.path.join(files.root, p) for p in lst.paths]
        for tag, metric, cp, is_sim in runs:
            s, _ = cms.sample_files_csr(db, paths, cp)
            abd = s.dense(db.leaf_n)
            sims = _oracle_sims(orc, tree, abd, metric)
            out = str(tmp_path / f"{name}_{tag}.dist")
            cms.write_dist(out, lst.names, sims, is_sim)
            assert open(out, "rb").read() == files.read(f"l_{name}_{tag}.dist"), (name, tag)


def test_single_pair_output_via_oracle(cms, orc, files):
    db = cms.Database(files.root, "G")
    tree = db.tree()
    for tag, f1, f2, metric, cp, is_sim in PAIRS:
        a1, m1 = cms.sample_file_abundance(db, os.path.join(files.root, f1), cp)
        a2, m2 = cms.sample_file_abundance(db, os.path.join(files.root, f2), cp)
        sim = float(_oracle_sims(orc, tree, np.stack([a1, a2]), metric)[0])
        v = sim if is_sim else 1.0 - sim
        exp = files.read(f"i_{tag}.txt").decode()
        assert exp == f"{m1} OTUs in file 1\n{m2} OTUs in file 2\n{v:.6g}\n", tag


# ------------------------------------------------------------------ GPU

def _run(cms, files, args):
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    return subprocess.run([exe] + args, cwd=files.root, env=dict(os.environ, ParallelMETA=files.root),
                          capture_output=True, text=True, timeout=120)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(LISTS))
def test_cuda_comp_list(cms, gpu_ok, files, tmp_path, name):
    """cuda-comp -l (default kernels: the reference's arithmetic): every
    metric byte-identical to comp -l."""
    prefix, runs = LISTS[name]
    for tag, metric, cp, is_sim in runs:
        print(f"cuda-comp -l {name} {tag}", flush=True)
        out = str(tmp_path / f"{name}_{tag}.dist")
        args = ["-l", f"list_{name}.txt", "-D", "G", "-o", out, "-M", str(metric), "-r", "T" if cp else "F",
                "-d", "F" if is_sim else "T"] + (["-p", prefix] if prefix else [])
        r = _run(cms, files, args)
        assert r.returncode == 0, r.stderr
        ref = files.read(f"l_{name}_{tag}.dist")
        assert open(out, "rb").read() == ref


@pytest.mark.gpu
def test_cuda_comp_single_pair(cms, gpu_ok, files):
    for tag, f1, f2, metric, cp, is_sim in PAIRS:
        print(f"cuda-comp -i {tag}", flush=True)
        r = _run(cms, files, ["-i", f1, f2, "-D", "G", "-M", str(metric), "-r", "T" if cp else "F",
                              "-d", "F" if is_sim else "T"])
        assert r.returncode == 0, r.stderr
        exp = files.read(f"i_{tag}.txt").decode().splitlines()
        got = r.stdout.splitlines()
        assert got == exp


def test_loader_errors(cms, files, tmp_path):
    db = cms.Database(files.root, "G")
    with pytest.raises(cms.CmsError):
        cms.sample_file_abundance(db, str(tmp_path / "missing.txt"))
    with pytest.raises(cms.CmsError):
        cms.sample_files_csr(db, [os.path.join(files.root, _all_files(files)[0]), str(tmp_path / "missing.txt")])
    # an empty file and a title-only file are all-zero samples with no OTUs
    for text in ("", "#Database_OTU\tCount\n"):
        p = tmp_path / "e.txt"
        p.write_text(text)
        a, m = cms.sample_file_abundance(db, str(p))
        assert m == 0 and not a.any()
    # counts that sum to zero: mapped OTUs, all-zero abundances
    p = tmp_path / "z.txt"
    p.write_text("#Database_OTU\tCount\nT1\t0\nT2\t0\n")
    a, m = cms.sample_file_abundance(db, str(p))
    assert m == 2 and not a.any()


def test_cuda_comp_input_errors(cms, files, tmp_path):
    """Errors that end cuda-comp before it touches the GPU: exit status 1 and
    a message (the reference prints and goes on with empty data)."""
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    env = dict(os.environ, ParallelMETA=files.root)

    def run(args):
        return subprocess.run([exe] + args, cwd=files.root, env=env, capture_output=True, text=True, timeout=60)

    r = run(["-i", "samples/S0/classification.txt"])          # -i needs two files
    assert r.returncode == 1 and "missing value" in r.stderr
    r = run(["-l", "no_such.list", "-o", str(tmp_path / "o.dist")])
    assert r.returncode == 1 and "Cannot open file" in r.stderr
    r = run(["-i", "samples/S0/classification.txt", "samples/none/classification.txt"])
    assert r.returncode == 1 and "Cannot open input file" in r.stderr
    r = run(["-l", "list_ids.txt"])                            # no -o
    assert r.returncode == 1 and "Cannot open output file" in r.stderr
    r = run(["-D", "G"])                                       # no input at all
    assert r.returncode == 1 and "an input is required" in r.stderr
