"""Host-side logic on the CPU: C-ABI exports, schedule validation, database
configuration, table parsing, the generator and the .dist writer."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def test_every_declared_symbol_is_exported(cms):
    hdr = open(os.path.join(ROOT, "include", "cms.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"\b(cms_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations parsed"
    missing = [s for s in sorted(declared) if not hasattr(cms.lib, s)]
    assert not missing, missing
    assert declared == set(cms.EXPORTED)


def _toy(cms):
    return cms.Tree(4, np.array([0, 2, -1], np.int32), np.array([0.1, 0.4, 0.3], np.float32),
                    np.array([1, 3, -2], np.int32), np.array([0.2, 0.5, 0.6], np.float32),
                    np.array([-1, -2, -1], np.int32))


def test_validate_tree(cms):
    ns, is_tree, nonneg = cms.validate_tree(_toy(cms))
    assert (ns, is_tree, nonneg) == (2, True, True)
    t = cms.synth_tree(1001, 99322, 0.02)
    ns, is_tree, nonneg = cms.validate_tree(t)
    assert is_tree and nonneg and ns <= t.max_slots <= 16


@pytest.mark.parametrize("o1,o2,od,what", [
    ([0, -5], [1, 2], [-1, -2], "read before write"),
    ([0, 9], [1, 2], [-1, -2], "leaf out of range"),
    ([0, -151], [1, 2], [-1, -2], "register below -150"),
    ([0, 2], [1, 3], [-1, 0], "destination not negative"),
])
def test_validate_tree_rejects(cms, o1, o2, od, what):
    t = cms.Tree(4, np.array(o1, np.int32), np.zeros(2, np.float32), np.array(o2, np.int32),
                 np.zeros(2, np.float32), np.array(od, np.int32))
    with pytest.raises(cms.CmsError) as e:
        cms.validate_tree(t)
    assert e.value.code == cms.E_SCHEDULE, what


def test_validate_tree_dag_and_negative_branch(cms):
    # node 2 reads register -1 twice (a DAG): valid walk, not a tree
    t = cms.Tree(4, np.array([0, -1, -1], np.int32), np.array([0.1, 0.1, 0.1], np.float32),
                 np.array([1, 2, -1], np.int32), np.array([0.1, 0.1, 0.1], np.float32),
                 np.array([-1, -2, -3], np.int32))
    assert cms.validate_tree(t)[1] is False
    t2 = _toy(cms)
    t2.dist_1[0] = -0.5
    assert cms.validate_tree(t2)[2] is False


def _write_config(root, text):
    os.makedirs(os.path.join(root, "databases", "x", "tree"), exist_ok=True)
    with open(os.path.join(root, "databases", "db.config"), "w") as f:
        f.write(text)
    with open(os.path.join(root, "databases", "x", "tree", "id.txt"), "w") as f:
        f.write("A\nB\n\nC\n")
    with open(os.path.join(root, "databases", "x", "tree", "order.txt"), "w") as f:
        f.write("0 0.1 1 0.2 -1\n\n-1 0.3 2 0.4 -1\n")


def test_db_config_rules(cms, tmp_path):
    root = str(tmp_path)
    _write_config(root, "# comment\n\nG\n0\nx\ndesc\nN\nY\nN\nN\n#dup G ignored\nG\n0\nnope\nd\nN\nN\nN\nN\n")
    db = cms.Database(root, "Q")  # unknown letter falls back to G (db.h:44)
    assert db.leaf_n == 3 and db.order_n == 2 and db.has_tree and not db.has_cp
    assert [db.leaf_name(i) for i in range(3)] == ["A", "B", "C"]
    assert db.leaf_name(7) == "NULL"
    tr = db.tree()
    assert list(tr.order_1) == [0, -1] and list(tr.order_d) == [-1, -1]


def test_db_config_without_g_is_an_error(cms, tmp_path):
    root = str(tmp_path)
    _write_config(root, "S\n0\nx\ndesc\nN\nY\nN\nN\n")
    with pytest.raises(cms.CmsError) as e:
        cms.Database(root, "S")
    assert e.value.code == cms.E_IO and "'G'" in str(e.value)


def test_db_missing_config(cms, tmp_path):
    with pytest.raises(cms.CmsError) as e:
        cms.Database(str(tmp_path), "G")
    assert e.value.code == cms.E_IO


def test_table_parse_rules(cms, tmp_path):
    p = tmp_path / "t.txt"
    p.write_text("title\tA\tB\tC\r\n"
                 "s1\t1\t2e1\t0\r\n"
                 "s2\t1\t2\n"            # too few: skipped
                 "s3\t1\t2\t3\t4\n"      # too many: skipped
                 "s4 0.5 1.5 +2\n"
                 "s5\t1\tx\t2\n"         # non-number stops the row: skipped
                 "s6\t1\t2\t3")          # no final newline
    t = cms.Table(str(p))
    assert t.feature_names() == ["A", "B", "C"]
    assert t.sample_names() == ["s1", "s4", "s6"]
    assert t.skipped == 3


def test_abundance_copy_number_and_prefix(cms, tmp_path, orc):
    root = str(tmp_path)
    dbdir = os.path.join(root, "databases", "toy")
    os.makedirs(os.path.join(dbdir, "tree"))
    with open(os.path.join(root, "databases", "db.config"), "w") as f:
        f.write("G\n0\ntoy\nToy\nN\nY\nY\nN\n")
    with open(os.path.join(dbdir, "tree", "id.txt"), "w") as f:
        f.write("A\nB\nC\nD\n")
    with open(os.path.join(dbdir, "tree", "order.txt"), "w") as f:
        f.write("0 0.1 1 0.2 -1\n2 0.4 3 0.5 -2\n-1 0.3 -2 0.6 -1\n")
    with open(os.path.join(dbdir, "copy_number.txt"), "w") as f:
        f.write("ID\tcp\nA\t2\nA\t5\nC\t4\n")   # duplicate A keeps the first
    db = cms.Database(root, "G")
    t = cms.Table(counts=np.array([[2, 1, 4, 0, 3]], np.float32), samples=["s"],
                  features=["OTU_A", "B", "otu_C", "D", "nothere"])
    abd, mapped = t.abundance(db, True)
    # x = [1, 1, 1, 0] (A: 2/2, C: 4/4); total = 3 -> 100/3 each
    np.testing.assert_allclose(abd[0], [100 / 3, 100 / 3, 100 / 3, 0], rtol=1e-6)
    assert mapped[0] == 3
    abd2, _ = t.abundance(db, False)
    np.testing.assert_allclose(abd2[0], [200 / 7, 100 / 7, 400 / 7, 0], rtol=1e-6)


def test_generator_deterministic_and_well_formed(cms):
    a = cms.synth_tree(99, 3000, 0.05)
    b = cms.synth_tree(99, 3000, 0.05)
    for x, y in zip(a.arrays(), b.arrays()):
        np.testing.assert_array_equal(x, y)
    leaves = np.concatenate([a.order_1[a.order_1 >= 0], a.order_2[a.order_2 >= 0]])
    assert sorted(leaves.tolist()) == list(range(3000))  # every leaf exactly once
    assert np.all((a.dist_1 >= 0) & (a.dist_1 <= 0.99))
    s1 = cms.synth_counts(5, 50, 3000, 20, 80, 1.1, 40, 0.4)
    s2 = cms.synth_counts(5, 50, 3000, 20, 80, 1.1, 40, 0.4)
    np.testing.assert_array_equal(s1.col, s2.col)
    np.testing.assert_array_equal(s1.val, s2.val)
    nnz = np.diff(s1.rowptr)
    assert nnz.min() >= 20 and nnz.max() <= 80
    assert np.all(s1.val >= 1) and np.all(s1.val == np.round(s1.val))
    for s in range(s1.n):
        c = s1.col[s1.rowptr[s]:s1.rowptr[s + 1]]
        assert np.all(np.diff(c) > 0)
    # Zipf popularity: the most common leaf is in far more samples than the median
    counts = np.bincount(s1.col, minlength=3000)
    assert counts.max() > 10 * max(np.median(counts), 1)


def test_negative_binomial_moments(cms):
    s = cms.synth_counts(8, 400, 2000, 200, 200, 0.0, 40, 0.4)
    # zero-truncated NB(mean 40, r 0.4): mean = 40 / (1 - P0), P0 = (r/(r+mu))^r
    p0 = (0.4 / 40.4) ** 0.4
    assert abs(s.val.mean() - 40 / (1 - p0)) / (40 / (1 - p0)) < 0.05


def test_counts_to_abundance_matches_oracle(cms, orc):
    tree, cp, counts, abd = cms.workload(0, samples=10)
    for s in range(10):
        row = np.zeros(tree.leaf_n, np.float32)
        sl = slice(counts.rowptr[s], counts.rowptr[s + 1])
        row[counts.col[sl]] = counts.val[sl]
        ref, _ = orc.load_abd(row, np.arange(tree.leaf_n, dtype=np.int32), cp, tree.leaf_n)
        got = np.zeros(tree.leaf_n, np.float32)
        got[abd.col[sl]] = abd.val[sl]
        np.testing.assert_array_equal(ref.view(np.uint32), got.view(np.uint32))


def test_db_round_trip_through_files(cms, tmp_path):
    tree, cp, counts, abd = cms.workload(0, samples=5)
    root = str(tmp_path)
    cms.write_db(root, "w0", tree, cp)
    cms.write_table(str(tmp_path / "t.txt"), counts, tree.leaf_n)
    db = cms.Database(root, "G")
    t2 = db.tree()
    for x, y in zip(tree.arrays(), t2.arrays()):
        np.testing.assert_array_equal(x, y)
    tab = cms.Table(str(tmp_path / "t.txt"))
    dense, _ = tab.abundance(db, True)
    np.testing.assert_array_equal(dense.view(np.uint32), abd.dense(tree.leaf_n).view(np.uint32))


def test_write_dist_format(cms, tmp_path):
    p = str(tmp_path / "o.dist")
    sim = np.array([0.5, 1e-6, 0.123456789], np.float32)
    cms.write_dist(p, ["a", "b", "c"], sim, False)
    lines = open(p).read().split("\n")
    assert lines[0] == "\ta\tb\tc"
    assert lines[1] == "a\t0\t0.5\t0.999999"
    assert lines[2].split("\t")[3] == "%g" % (1.0 - float(np.float32(0.123456789)))
    cms.write_dist(p, ["a", "b", "c"], sim, True)
    lines = open(p).read().split("\n")
    assert lines[1] == "a\t1\t0.5\t1e-06"


def test_package_refuses_without_library(tmp_path):
    """The package raises on import when libcms.so is missing (no fallback)."""
    import shutil
    import subprocess
    import sys
    pkg = os.path.join(ROOT, "coordinated-meta-storms_amd")
    dst = tmp_path / "coordinated-meta-storms_amd"
    dst.mkdir()
    shutil.copy(os.path.join(pkg, "__init__.py"), dst / "__init__.py")
    r = subprocess.run([sys.executable, "-c", "import importlib; importlib.import_module('coordinated-meta-storms_amd')"],
                       cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode != 0 and "libcms.so not built" in r.stderr


@pytest.mark.parametrize("case", ["quirks", "config0", "core"])
def test_abundance_csr_equals_dense(cms, golden, case):
    # the CSR loader gives exactly the dense loader's values (duplicated OTU
    # columns, prefixes and unmapped columns of the quirks table included)
    g = golden(case)
    db = cms.Database(g.root, "G")
    t = cms.Table(g.table)
    for cp in (True, False):
        dense, mapped = t.abundance(db, cp)
        csr = t.abundance_csr(db, cp)
        np.testing.assert_array_equal(csr.dense(db.leaf_n), dense)
        np.testing.assert_array_equal(np.diff(csr.rowptr), mapped)
        for s in range(csr.n):
            c = csr.col[csr.rowptr[s]:csr.rowptr[s + 1]]
            assert np.all(np.diff(c) > 0)


@pytest.mark.parametrize("case", ["toy", "quirks", "config0", "core"])
def test_sparse_table_file_same_abundances(cms, golden, tmp_path, case):
    # the reference's text tables saved as sparse table files load to the
    # same names and bit-identical abundances (dense and CSR, copy numbers on
    # and off): duplicated, prefixed and unmapped columns and skipped rows of
    # the quirks table included (skipped rows are not saved)
    g = golden(case)
    db = cms.Database(g.root, "G")
    t = cms.Table(g.table)
    p = str(tmp_path / "t.cmst")
    t.save_csr(p)
    s = cms.Table(p)
    assert s.is_sparse and not t.is_sparse
    assert s.sample_names() == t.sample_names() and s.feature_names() == t.feature_names()
    assert s.skipped == 0
    for cp in (True, False):
        d0, m0 = t.abundance(db, cp)
        d1, m1 = s.abundance(db, cp)
        np.testing.assert_array_equal(d1.view(np.uint32), d0.view(np.uint32))
        np.testing.assert_array_equal(m1, m0)
        c0, c1 = t.abundance_csr(db, cp), s.abundance_csr(db, cp)
        np.testing.assert_array_equal(c1.rowptr, c0.rowptr)
        np.testing.assert_array_equal(c1.col, c0.col)
        np.testing.assert_array_equal(c1.val.view(np.uint32), c0.val.view(np.uint32))
    # saving the sparse table again gives the same bytes
    p2 = str(tmp_path / "t2.cmst")
    s.save_csr(p2)
    assert open(p, "rb").read() == open(p2, "rb").read()


def test_sparse_table_writer_matches_text_writer(cms, tmp_path):
    counts = cms.synth_counts(77, 40, 300, 5, 60, 1.1, 30.0, 0.5)
    tp, cp_ = str(tmp_path / "t.txt"), str(tmp_path / "t.cmst")
    cms.write_table(tp, counts, 300)
    cms.write_table_csr(cp_, counts, 300)
    a, b = cms.Table(tp), cms.Table(cp_)
    assert a.sample_names() == b.sample_names() and a.feature_names() == b.feature_names()
    tree = cms.synth_tree(78, 300, 0.05)
    root = str(tmp_path / "db")
    cms.write_db(root, "x", tree, cms.synth_copy_numbers(79, 300))
    db = cms.Database(root, "G")
    np.testing.assert_array_equal(a.abundance(db)[0].view(np.uint32), b.abundance(db)[0].view(np.uint32))
    assert os.path.getsize(cp_) < os.path.getsize(tp)


def test_sparse_table_rejects_bad_files(cms, tmp_path):
    counts = cms.synth_counts(80, 6, 50, 2, 10, 1.1, 30.0, 0.5)
    good = str(tmp_path / "g.cmst")
    cms.write_table_csr(good, counts, 50)
    raw = open(good, "rb").read()
    nnz = int(counts.rowptr[-1])
    cases = {
        "truncated": raw[:-3],
        "trailing": raw + b"\0",
        # the first column index (after the row offsets) set out of range
        "column": raw[: len(raw) - 8 * nnz] + (1000).to_bytes(4, "little") + raw[len(raw) - 8 * nnz + 4:],
        # the first row's columns made descending when it has two or more
        "order": (raw[: len(raw) - 8 * nnz] + raw[len(raw) - 8 * nnz + 4: len(raw) - 8 * nnz + 8]
                  + raw[len(raw) - 8 * nnz: len(raw) - 8 * nnz + 4] + raw[len(raw) - 8 * nnz + 8:]),
    }
    assert counts.rowptr[1] >= 2
    for name, data in cases.items():
        p = str(tmp_path / f"{name}.cmst")
        open(p, "wb").write(data)
        with pytest.raises(cms.CmsError) as e:
            cms.Table(p)
        assert e.value.code == cms.E_IO, name
    # built from memory: bad rows are refused
    with pytest.raises(cms.CmsError):
        cms.Table(csr=(np.array([0, 2]), np.array([3, 1]), np.array([1.0, 2.0])), samples=["a"],
                  features=[f"f{j}" for j in range(5)])


def test_cms_gen_writes_sparse_table(cms, tmp_path):
    import subprocess
    exe = os.path.join(cms.BIN_DIR, "cms-gen")
    t, c = str(tmp_path / "t.txt"), str(tmp_path / "t.cmst")
    r = subprocess.run([exe, "--root", str(tmp_path), "--table", t, "--csr", c, "--leaves", "200", "--samples", "9"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    a, b = cms.Table(t), cms.Table(c)
    assert b.is_sparse and a.n == b.n == 9 and a.f == b.f == 200
    db = cms.Database(str(tmp_path), "G")
    np.testing.assert_array_equal(a.abundance(db)[0].view(np.uint32), b.abundance(db)[0].view(np.uint32))


@pytest.mark.parametrize("field,value", [("embed_input", 2), ("embed_input", -1), ("embed_slice", -64),
                                         ("variant", 17), ("chunk", 12), ("tile_a", 64), ("wide_rows", 2),
                                         ("wide_rows", -1)])
def test_context_options_refused_before_any_device_call(cms, field, value):
    # cms_create checks the options first: an out-of-range field is
    # CMS_E_INVAL on any machine, GPU or not
    with pytest.raises(cms.CmsError) as e:
        cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, **{field: value})
    assert e.value.code == cms.E_INVAL
