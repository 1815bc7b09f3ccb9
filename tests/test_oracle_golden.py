"""The oracle and the host loader pinned against the compiled reference.

Golden fixtures (tests/golden/*.tar.gz, made by make_golden.py from the
reference `comp` and its own _Comp_Tree::Calc_sim) fix, bit for bit:
the FP32 operation order of Calc_sim (comp.h:332-375), the Load_abd
arithmetic and parse quirks (comp.h:243-279, table_format.h:110-165), and the
.dist text (comp_sam.cpp:184-219). Closed forms fix the rest.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN_CASES


def _loaded(cms, g, cp=True):
    db = cms.Database(g.root, "G")
    tree = db.tree()
    tab = cms.Table(g.table)
    abd, mapped = tab.abundance(db, cp)
    return db, tree, tab, abd


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("cp", [True, False])
def test_oracle_f32_bit_exact_vs_reference(cms, orc, golden, case, cp):
    g = golden(case)
    db, tree, tab, abd = _loaded(cms, g, cp)
    i, j, ref = g.hex_sims(cp)
    got = orc.pairs_f32(tree, abd, i, j)
    assert got.dtype == np.float32
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_oracle_f64_and_closed_form_agree(cms, orc, golden, case):
    g = golden(case)
    db, tree, tab, abd = _loaded(cms, g, True)
    i, j, ref = g.hex_sims(True)
    walk64 = orc.pairs_f64(tree, abd, i, j)
    closed = orc.pairs_f64(tree, abd, i, j, closed_form=True)
    np.testing.assert_allclose(walk64, closed, rtol=0, atol=1e-12)
    # FP32 reference vs FP64 truth: the reference's own rounding noise
    assert np.abs(ref.astype(np.float64) - walk64).max() < 1e-5


def test_known_answers_toy(cms, orc, golden):
    """((A:0.1,B:0.2):0.3,(C:0.4,D:0.5):0.6): distances by hand."""
    g = golden("toy")
    for cp, expect in [(False, [0.5, 0.5, 0.1]), (True, [2 / 3, 0.5, 0.2])]:
        db, tree, tab, abd = _loaded(cms, g, cp)
        i = np.array([0, 0, 1], np.int32)
        j = np.array([1, 2, 2], np.int32)
        for fn in (lambda: orc.pairs_f64(tree, abd, i, j), lambda: orc.pairs_f64(tree, abd, i, j, True)):
            np.testing.assert_allclose(1 - fn(), expect, atol=1e-7)


def _py_walk(tree, a, b, mistake=None):
    """Pure-Python FP64 walk with an optional planted mistake."""
    R1, R2 = [0.0] * 150, [0.0] * 150
    total = 0.0
    for i in range(tree.order_n):
        c1, c2, dst = int(tree.order_1[i]), int(tree.order_2[i]), int(tree.order_d[i]) + 150
        # weights in FP32 with the MIN_DIST clamp, as the reference forms them
        w1 = float(np.float32(1) - tree.dist_1[i])
        w2 = float(np.float32(1) - tree.dist_2[i])
        w1 = float(np.float32(1e-5)) if w1 < 0 else w1
        w2 = float(np.float32(1e-5)) if w2 < 0 else w2
        if mistake == "no_decay":
            w1 = w2 = 1.0
        if mistake == "swap_w":
            w1, w2 = w2, w1
        a1, b1 = (float(a[c1]), float(b[c1])) if c1 >= 0 else (R1[c1 + 150], R2[c1 + 150])
        a2, b2 = (float(a[c2]), float(b[c2])) if c2 >= 0 else (R1[c2 + 150], R2[c2 + 150])
        m1, m2 = min(a1, b1), min(a2, b2)
        total += m1 + (m2 if mistake != "drop_m2" else 0.0)
        if mistake == "max":
            m1, m2 = max(a1, b1), max(a2, b2)
        R1[dst] = (a1 - m1) * w1 + (a2 - m2) * w2
        R2[dst] = (b1 - m1) * w1 + (b2 - m2) * w2
    return min(max(total / 100.0, 0.0), 1.0)


def test_closed_form_catches_planted_mistakes(cms, orc):
    """The closed form (a different algorithm) agrees with a correct walk to
    1e-12 and disagrees with walks carrying plausible mistakes."""
    tree = cms.synth_tree(7, 300, 0.1, 0.99)
    rng = np.random.default_rng(0)
    a = np.zeros(300, np.float32)
    b = np.zeros(300, np.float32)
    a[rng.choice(300, 60, replace=False)] = rng.random(60).astype(np.float32)
    b[rng.choice(300, 60, replace=False)] = rng.random(60).astype(np.float32)
    a *= 100 / a.sum()
    b *= 100 / b.sum()
    closed = orc.sharedmin_f64(tree, a, b)
    assert abs(_py_walk(tree, a, b) - closed) < 1e-12
    assert abs(orc.calc_sim_f64(tree, a, b) - closed) < 1e-12
    for m in ("no_decay", "swap_w", "drop_m2", "max"):
        assert abs(_py_walk(tree, a, b, m) - closed) > 1e-6, m


def test_bruteforce_caterpillar(cms, orc):
    """Caterpillar tree: node k joins the previous node and leaf k+1. The
    decayed masses x and the shared-minimum sum are written out node by node
    in numpy; only the two root edges keep beta = 1."""
    rng = np.random.default_rng(3)
    L = 9
    bl = rng.random(2 * L).astype(np.float32) * 0.5
    # caterpillar: node0 = (leaf0, leaf1), node k = (node k-1, leaf k+1)
    o1 = np.array([0] + [-1] * (L - 2), np.int32)
    o2 = np.arange(1, L, dtype=np.int32)
    od = np.full(L - 1, -1, np.int32)
    d1 = bl[: L - 1].copy()
    d2 = bl[L - 1: 2 * L - 2].copy()
    tree = cms.Tree(L, o1, d1, o2, d2, od)
    a = rng.random(L).astype(np.float32)
    b = rng.random(L).astype(np.float32)
    a *= 100 / a.sum()
    b *= 100 / b.sum()
    # brute force: x(node) = sum over leaves below of value * prod of w on the path
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    w1 = (np.float32(1) - d1).astype(np.float64)
    w2 = (np.float32(1) - d2).astype(np.float64)
    total = 0.0
    xa = xb = None
    for k in range(L - 1):
        ca, cb = (a64[0], b64[0]) if k == 0 else (xa, xb)
        la, lb = a64[k + 1], b64[k + 1]
        beta1 = (1 - w1[k]) if k < L - 2 else 1.0
        beta2 = (1 - w2[k]) if k < L - 2 else 1.0
        total += beta1 * min(ca, cb) + beta2 * min(la, lb)
        xa = w1[k] * ca + w2[k] * la
        xb = w1[k] * cb + w2[k] * lb
    assert abs(orc.calc_sim_f64(tree, a, b) - total / 100) < 1e-12


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("cp", [True, False])
def test_loader_matches_oracle_load_abd(cms, orc, golden, case, cp):
    """Product loader (cms_table_abundance) == oracle restatement of Load_abd."""
    g = golden(case)
    db, tree, tab, abd = _loaded(cms, g, cp)
    feats = tab.feature_names()
    ids = {db.leaf_name(k): k for k in range(db.leaf_n)}
    cps = {}
    if cp:
        cpfile = [os.path.join(dp, f) for dp, _, fs in os.walk(os.path.join(g.root, "databases")) for f in fs
                  if f == "copy_number.txt"][0]
        with open(cpfile) as f:
            next(f)
            for line in f:
                p = line.split()
                if len(p) >= 2 and p[0] not in cps:
                    cps[p[0]] = np.float32(float(p[1]))

    def strip(s):
        return s[4:] if len(s) > 4 and s[:4] in ("otu_", "OTU_") else s

    col_leaf = np.array([ids.get(strip(f), -1) for f in feats], np.int32)
    col_cp = np.array([cps.get(strip(f), np.float32(1.0)) if cp else 1.0 for f in feats], np.float32)
    counts = []
    with open(g.table) as f:
        next(f)
        for line in f:
            p = line.split()
            if len(p) - 1 == len(feats):
                counts.append(np.array([float(x) for x in p[1:]], np.float32))
    assert len(counts) == tab.n
    for s in range(tab.n):
        ref, _ = orc.load_abd(counts[s], col_leaf, col_cp, db.leaf_n)
        np.testing.assert_array_equal(ref.view(np.uint32), abd[s].view(np.uint32))


def test_quirks_parse(cms, golden):
    g = golden("quirks")
    tab = cms.Table(g.table)
    assert tab.skipped == 1 and tab.n == 12
    db = cms.Database(g.root, "G")
    abd, mapped = tab.abundance(db, True)
    # a duplicated OTU column adds to the total twice: rows do not sum to 100
    assert np.all(np.abs(abd.sum(1) - 100) < 100)


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("tag,cp,is_sim", [("dist_rT", True, False), ("dist_rF", False, False),
                                           ("sim_rT", True, True)])
def test_dist_writer_byte_identical(cms, golden, tmp_path, case, tag, cp, is_sim):
    """Output_Matrix text from the reference's FP32 sims (hex fixture)."""
    g = golden(case)
    tab = cms.Table(g.table)
    i, j, s = g.hex_sims(cp)
    n = tab.n
    packed = np.zeros(n * (n - 1) // 2, np.float32)
    packed[cms.packed_index(i, j, n)] = s
    out = str(tmp_path / "o.dist")
    cms.write_dist(out, tab.sample_names(), packed, is_sim)
    assert open(out, "rb").read() == g.read(tag + ".dist")
