"""Database builder: Newick -> tree/id.txt + tree/order.txt (cms_newick_*).

Pins:
- core.tre (the reference's only real tree, databases/oral_core/tree/core.tre,
  copied to tests/golden/core.tre): the id.txt written here is byte-identical
  to the reference's own id.txt, whose SHA-256 and size its git-lfs pointer
  records (databases/oral_core/tree/id.txt); 1,045 leaves, 1,044 internal
  nodes, and the Sethi-Ullman schedule needs 6 registers.
- The schedule computes the walk of comp.h:332-375: the FP64 walk along the
  built order.txt (oracle) equals a recursive evaluation over the parsed tree
  written here independently of any schedule or register.
- The reference's own comp reads the built database (golden case "core",
  tests/test_oracle_golden.py and the GPU CLI tests).
"""
import hashlib
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

CORE = os.path.join(GOLDEN, "core.tre")
# git-lfs pointer of the reference's databases/oral_core/tree/id.txt
CORE_ID_SHA256 = "ed2a925596fd4e9fa982e47627f4001dfb6dfbdfb3776ee4d745d76832f9bb02"
CORE_ID_SIZE = 9226
MIN_DIST = 1e-5


# ------------------------------------------------ independent evaluation

def parse(text):
    """Minimal Newick reader for the tests: (children, name, length) tuples."""
    text = re.sub(r"\[[^\]]*\]", "", text).strip()
    pos = 0

    def node():
        nonlocal pos
        if text[pos] == "(":
            pos += 1
            ch = [node()]
            while text[pos] == ",":
                pos += 1
                ch.append(node())
            assert text[pos] == ")"
            pos += 1
            m = re.match(r"[^(),:;]*", text[pos:])
            pos += len(m.group(0))
            name = None
        else:
            m = re.match(r"[^(),:;]+", text[pos:])
            name = m.group(0).strip()
            pos += len(m.group(0))
            ch = None
        length = 0.0
        if text[pos] == ":":
            pos += 1
            m = re.match(r"[^(),;]+", text[pos:])
            length = float(m.group(0))
            pos += len(m.group(0))
        return (ch, name, length)

    return node()


def leaves(t, out=None):
    out = [] if out is None else out
    if t[0] is None:
        out.append(t[1])
    else:
        for c in t[0]:
            leaves(c, out)
    return out


def walk(t, a, b, idx):
    """comp.h:332-375 in FP64 by recursion: (leftover a, leftover b, total).
    A node with k children is read as a left-leaning chain of binary nodes
    joined by zero-length edges; one child is merged into its child's edge."""
    ch, name, _ = t
    if ch is None:
        i = idx[name]
        return a[i], b[i], 0.0
    if len(ch) == 1:
        return walk(ch[0], a, b, idx)
    parts = [(walk(c, a, b, idx), c[2] if c[0] is None or len(c[0]) != 1 else c[2] + c[0][0][2]) for c in ch]

    def join(x, y):
        (xa, xb, xt), dx = x
        (ya, yb, yt), dy = y
        wx, wy = 1.0 - np.float32(dx), 1.0 - np.float32(dy)
        wx = MIN_DIST if wx < 0 else wx
        wy = MIN_DIST if wy < 0 else wy
        mx, my = min(xa, xb), min(ya, yb)
        return ((xa - mx) * wx + (ya - my) * wy, (xb - mx) * wx + (yb - my) * wy, xt + yt + mx + my)

    acc = join(parts[0], parts[1])
    for p in parts[2:]:
        acc = join((acc, 0.0), p)
    return acc


def sim64(t, a, b, idx):
    return min(1.0, max(0.0, walk(t, a, b, idx)[2] / 100.0))


def random_abd(rng, n, leaf_n, k):
    abd = np.zeros((n, leaf_n), np.float32)
    for s in range(n):
        cols = rng.choice(leaf_n, size=min(k, leaf_n), replace=False)
        v = rng.gamma(0.5, 1.0, len(cols)).astype(np.float32) + np.float32(1e-3)
        abd[s, cols] = v / v.sum() * 100
    return abd


# ------------------------------------------------------------------ tests

def test_core_id_txt_is_the_references(cms, tmp_path):
    nw = cms.Newick(CORE)
    nw.write_db(str(tmp_path), "oral_core")
    data = open(tmp_path / "databases" / "oral_core" / "tree" / "id.txt", "rb").read()
    assert len(data) == CORE_ID_SIZE
    assert hashlib.sha256(data).hexdigest() == CORE_ID_SHA256


def test_core_schedule_shape(cms):
    nw = cms.Newick(CORE)
    assert nw.leaf_n == 1045 and nw.order_n == 1044
    info = nw.info()
    assert info == {"nslots": 6, "multifurcations": 0, "unary_merged": 0}
    t = nw.tree()
    cms.validate_tree(t)
    # every leaf read exactly once, every register value read exactly once,
    # the root last
    reads = np.concatenate([t.order_1, t.order_2])
    assert sorted(reads[reads >= 0].tolist()) == list(range(1045))
    assert (t.order_d < 0).all() and (t.order_d >= -6).all()
    assert t.order_1.dtype == np.int32


def test_core_schedule_computes_the_walk(cms, orc):
    text = open(CORE).read()
    ref_tree = parse(text)
    nw = cms.Newick(CORE)
    names = nw.names
    assert names == leaves(ref_tree)
    idx = {n: i for i, n in enumerate(names)}
    rng = np.random.default_rng(5)
    abd = random_abd(rng, 12, len(names), 60)
    ii, jj = np.triu_indices(12, 1)
    got = orc.pairs_f64(nw.tree(), abd, ii.astype(np.int32), jj.astype(np.int32))
    exp = np.array([sim64(ref_tree, abd[i].astype(np.float64), abd[j].astype(np.float64), idx)
                    for i, j in zip(ii, jj)])
    assert np.abs(got - exp).max() < 1e-12
    assert exp.min() > 0.0  # the pairs share mass: the check is not vacuous


def test_multifurcating_and_unary_trees(cms, orc):
    text = "((A:0.1,B:0.2,C:0.05,D:0.3):0.2,((E:0.4):0.1,F:0.15):0.05,G:0.7);"
    nw = cms.Newick(text=text)
    info = nw.info()
    assert info["multifurcations"] == 2 and info["unary_merged"] == 1
    assert nw.names == list("ABCDEFG")
    t = nw.tree()
    assert nw.order_n == 6  # binary tree on 7 leaves
    cms.validate_tree(t)
    # the merged unary edge carries 0.4 + 0.1
    k = int(np.nonzero((t.order_1 == 4) | (t.order_2 == 4))[0][0])
    d = t.dist_1[k] if t.order_1[k] == 4 else t.dist_2[k]
    assert d == np.float32(0.5)
    ref_tree = parse(text)
    idx = {n: i for i, n in enumerate(nw.names)}
    rng = np.random.default_rng(9)
    abd = random_abd(rng, 8, 7, 4)
    ii, jj = np.triu_indices(8, 1)
    got = orc.pairs_f64(t, abd, ii.astype(np.int32), jj.astype(np.int32))
    exp = np.array([sim64(ref_tree, abd[i].astype(np.float64), abd[j].astype(np.float64), idx)
                    for i, j in zip(ii, jj)])
    assert np.abs(got - exp).max() < 1e-12


def test_planted_mistake_is_caught(cms, orc):
    # swapping two branch lengths of the built schedule moves the walk away
    # from the recursive evaluation: the comparison above has teeth
    text = open(CORE).read()
    ref_tree = parse(text)
    nw = cms.Newick(CORE)
    t = nw.tree()
    t.dist_1[10], t.dist_2[10] = t.dist_2[10], t.dist_1[10]
    idx = {n: i for i, n in enumerate(nw.names)}
    rng = np.random.default_rng(5)
    abd = random_abd(rng, 12, nw.leaf_n, 400)
    ii, jj = np.triu_indices(12, 1)
    got = orc.pairs_f64(t, abd, ii.astype(np.int32), jj.astype(np.int32))
    exp = np.array([sim64(ref_tree, abd[i].astype(np.float64), abd[j].astype(np.float64), idx)
                    for i, j in zip(ii, jj)])
    assert np.abs(got - exp).max() > 1e-6


def test_syntax(cms):
    # quoted labels, comments, internal labels, exponents, missing lengths
    nw = cms.Newick(text="[tree 1] (('a b':1e-2,'it''s':2.5E-1)0.95:0.1 , c)root;\n")
    assert nw.names == ["a b", "it's", "c"]
    t = nw.tree()
    assert nw.order_n == 2
    assert np.float32(0.01) in t.dist_1.tolist() + t.dist_2.tolist()
    assert np.float32(0.25) in t.dist_1.tolist() + t.dist_2.tolist()
    assert cms.Newick(text="A;").order_n == 0
    for bad in ["", "(A,B)", "(A,B));", "((A,B);", "(A,);", "(A:x,B);", "(A,B)C D;", "[open (A,B);"]:
        with pytest.raises(cms.CmsError) as e:
            cms.Newick(text=bad)
        assert e.value.code == cms.E_INVAL, bad


def test_deep_caterpillar(cms):
    # 20,000 levels: explicit stacks, and one register suffices
    n = 20000
    text = "(" * (n - 1) + "L0:0.1" + "".join(f",L{i}:0.1):0.01" for i in range(1, n - 1)) + f",L{n - 1}:0.1);"
    nw = cms.Newick(text=text)
    assert nw.leaf_n == n and nw.order_n == n - 1
    assert nw.info()["nslots"] == 1
    cms.validate_tree(nw.tree())


def test_cli_writes_the_same_database(cms, tmp_path):
    exe = os.path.join(ROOT, "coordinated-meta-storms_amd", "bin", "cms-make-db")
    cp = tmp_path / "cp.txt"
    cp.write_text("ID\tcp\nHM358615\t3\n")
    r = subprocess.run([exe, "-i", CORE, "-o", str(tmp_path / "cli"), "-n", "oral_core", "-c", str(cp)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "1045 leaves, 1044 internal nodes, 6 registers" in r.stdout
    cms.Newick(CORE).write_db(str(tmp_path / "api"), "oral_core", str(cp))
    for rel in ["databases/db.config", "databases/oral_core/tree/id.txt", "databases/oral_core/tree/order.txt",
                "databases/oral_core/copy_number.txt"]:
        assert open(tmp_path / "cli" / rel, "rb").read() == open(tmp_path / "api" / rel, "rb").read()
    db = cms.Database(str(tmp_path / "cli"), "G")
    assert db.leaf_n == 1045 and db.order_n == 1044
    r = subprocess.run([exe, "-i", str(tmp_path / "missing.tre"), "-o", str(tmp_path)], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "Cannot open" in r.stderr
