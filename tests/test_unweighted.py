"""Unweighted Meta-Storms (comp -M 1, _Comp_Tree::Calc_sim_unweight,
comp.h:378-446).

CPU: the oracle's FP32 restatement is bit-exact with the compiled reference
(golden hex fixtures) and agrees with the FP64 presence-set ratio
shared / (W_a + W_b - shared).
GPU: the shared-minimum kernels with presence propagation and coefficient w
are within 1e-5 of FP64; cuda-comp -M 1 prints the reference's .dist values
(to the 6-digit print, within one unit of the last digit).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN_CASES

TOL = 1e-5


def _inputs(cms, g, cp=True):
    db = cms.Database(g.root, "G")
    tab = cms.Table(g.table)
    abd, _ = tab.abundance(db, cp)
    return db.tree(), abd


def _hex(g):
    rows = [ln.split() for ln in g.read("hex_M1.txt").decode().splitlines() if ln.strip()]
    i = np.array([int(r[0]) for r in rows], np.int32)
    j = np.array([int(r[1]) for r in rows], np.int32)
    s = np.array([float.fromhex(r[2]) for r in rows], np.float32)
    return i, j, s


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_oracle_unweighted_bit_exact_vs_reference(cms, orc, golden, case):
    g = golden(case)
    tree, abd = _inputs(cms, g)
    i, j, ref = _hex(g)
    s32, s64 = orc.pairs_unweight(tree, abd, i, j)
    np.testing.assert_array_equal(s32.view(np.uint32), ref.view(np.uint32))
    assert np.abs(s64 - ref).max() < 1e-5


def test_unweighted_known_answer(cms, orc, golden):
    """Toy tree, s1 = {A, B}, s2 = {A, C}: shared edges A (0.9) + AB (0.7)
    = 1.6; either: A, B, C, AB, CD = 0.9 + 0.8 + 0.6 + 0.7 + 0.4 = 3.4."""
    g = golden("toy")
    tree, abd = _inputs(cms, g)
    _, s64 = orc.pairs_unweight(tree, abd, np.array([0], np.int32), np.array([1], np.int32))
    assert abs(s64[0] - 1.6 / 3.4) < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_gpu_unweighted_vs_fp64(cms, orc, golden, gpu_ok, case):
    print(f"unweighted golden {case}", flush=True)
    g = golden(case)
    tree, abd = _inputs(cms, g)
    i, j, ref = _hex(g)
    _, s64 = orc.pairs_unweight(tree, abd, i, j)
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, metric=cms.METRIC_UNWEIGHTED) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples(abd)
        got = ctx.all_pairs()
        n = abd.shape[0]
        p = cms.packed_index(i, j, n)
        assert np.abs(got[p] - s64).max() <= TOL
        np.testing.assert_array_equal(ctx.sim_pairs(i, j), got[p])


@pytest.mark.gpu
def test_gpu_unweighted_large_tree(cms, orc, gpu_ok):
    print("unweighted large tree", flush=True)
    tree, cp, counts, abd = cms.workload(1, samples=80)
    ii, jj = np.triu_indices(abd.n, 1)
    s32, s64 = orc.pairs_unweight(tree, abd.dense(tree.leaf_n), ii.astype(np.int32), jj.astype(np.int32))
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, metric=cms.METRIC_UNWEIGHTED) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        got = ctx.all_pairs()
    err = np.abs(got - s64)
    print(f"unweighted max err {err.max():.3e}", flush=True)
    assert err.max() <= TOL
    # the walk: Calc_sim_unweight's FP32 arithmetic, bit for bit
    with cms.Context(0, kernel=cms.KERNEL_WALK, metric=cms.METRIC_UNWEIGHTED) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        walk = ctx.all_pairs()
    np.testing.assert_array_equal(walk.view(np.uint32), s32.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_gpu_unweighted_walk_bit_exact(cms, orc, golden, gpu_ok, case):
    """The walk kernel computes comp -M 1 bit for bit: all pairs, the pair
    list and a rectangle against the reference's hex fixtures."""
    print(f"unweighted walk golden {case}", flush=True)
    g = golden(case)
    tree, abd = _inputs(cms, g)
    i, j, ref = _hex(g)
    n = abd.shape[0]
    with cms.Context(0, kernel=cms.KERNEL_WALK, metric=cms.METRIC_UNWEIGHTED) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples(abd)
        got = ctx.all_pairs()
        p = cms.packed_index(i, j, n)
        np.testing.assert_array_equal(got[p].view(np.uint32), ref.view(np.uint32))
        np.testing.assert_array_equal(ctx.sim_pairs(i, j).view(np.uint32), ref.view(np.uint32))
        t = ctx.tile(0, n, 0, n)
        full = np.ones((n, n), np.float32)
        full[i, j] = ref
        full[j, i] = ref
        np.testing.assert_array_equal(t.view(np.uint32), full.view(np.uint32))


@pytest.mark.gpu
def test_gpu_unweighted_walk_empty_samples(cms, orc, gpu_ok):
    """Two samples with no leaf present divide 0 by 0: the host's sign-set
    NaN, as the reference prints it; one empty sample gives 0."""
    tree = cms.synth_tree(21, 300, 0.05)
    abd = np.zeros((4, tree.leaf_n), np.float32)
    abd[1, :5] = 20.0
    abd[3, 100:110] = 10.0
    with cms.Context(0, kernel=cms.KERNEL_WALK, metric=cms.METRIC_UNWEIGHTED) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples(abd)
        got = ctx.all_pairs()
    ii, jj = np.triu_indices(4, 1)
    s32, _ = orc.pairs_unweight(tree, abd, ii.astype(np.int32), jj.astype(np.int32))
    np.testing.assert_array_equal(got.view(np.uint32), s32.view(np.uint32))
    assert got.view(np.uint32)[cms.packed_index(0, 2, 4)] == 0xFFC00000


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_cuda_comp_unweighted(cms, golden, gpu_ok, tmp_path, case):
    """cuda-comp -M 1: the walk writes comp -M 1's bytes; the shared-minimum
    kernel agrees within the 6-digit print."""
    g = golden(case)
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    for kernel in ("walk", "sharedmin"):
        out = str(tmp_path / f"o_{kernel}.dist")
        r = subprocess.run([exe, "-T", g.table, "-D", "G", "-o", out, "-M", "1", "--kernel", kernel],
                           env=dict(os.environ, ParallelMETA=g.root), capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        if kernel == "walk":
            assert open(out, "rb").read() == g.read("dist_M1.dist")
            continue
        got = [ln.split("\t") for ln in open(out).read().splitlines()]
        ref = [ln.split("\t") for ln in g.read("dist_M1.dist").decode().splitlines()]
        assert got[0] == ref[0] and len(got) == len(ref)
        for rg, rr in zip(got[1:], ref[1:]):
            assert rg[0] == rr[0]
            for a, b in zip(rg[1:], rr[1:]):
                assert abs(float(a) - float(b)) <= 2e-5 + 1e-5 * abs(float(b)), (a, b)
