"""GPU parity: libcms kernels (through the C ABI) against the oracle and the
reference's golden outputs.

Bars:
  walk F32        bit-exact with the reference Calc_sim (comp.h:332-375)
  shared-minimum  |sim - FP64 oracle| <= 1e-5 max (fixed point, ~1e-8 typical)
  BF16 / F16      error vs the FP64 oracle reported and bounded (storage
                  rounding of the abundances: 2^-9 / 2^-12 relative)
"""
import numpy as np
import pytest

from conftest import GOLDEN_CASES

pytestmark = pytest.mark.gpu

SM_TOL = 1e-5


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx(cms, gpu_ok):
    c = cms.Context(0)
    yield c
    c.close()


def _golden_inputs(cms, g, cp):
    db = cms.Database(g.root, "G")
    tree = db.tree()
    tab = cms.Table(g.table)
    abd, _ = tab.abundance(db, cp)
    return tree, abd


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("cp", [True, False])
def test_walk_bit_exact_vs_reference(cms, ctx, golden, case, cp):
    _flush(f"walk golden {case} cp={cp}")
    g = golden(case)
    tree, abd = _golden_inputs(cms, g, cp)
    i, j, ref = g.hex_sims(cp)
    ctx.set_opts(cms.KERNEL_WALK, cms.STORE_F32)
    ctx.set_tree(tree)
    ctx.set_samples(abd)
    got = ctx.all_pairs()
    n = abd.shape[0]
    exp = np.zeros(n * (n - 1) // 2, np.float32)
    exp[cms.packed_index(i, j, n)] = ref
    np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))
    # explicit pair list (the reference kernel boundary) and a tile
    np.testing.assert_array_equal(ctx.sim_pairs(i, j).view(np.uint32), ref.view(np.uint32))
    t = ctx.tile(0, n, 0, n)
    full = np.ones((n, n), np.float32)
    full[i, j] = ref
    full[j, i] = ref
    np.testing.assert_array_equal(t.view(np.uint32), full.view(np.uint32))


def test_walk_large_tree_bit_exact(cms, orc, ctx):
    """99,322-leaf tree (the GreenGenes-sized workload), 72 samples."""
    _flush("walk large tree")
    tree, cp, counts, abd = cms.workload(1, samples=72)
    ctx.set_opts(cms.KERNEL_WALK, cms.STORE_F32)
    ctx.set_tree(tree)
    ctx.set_samples_csr(abd)
    got = ctx.all_pairs()
    dense = abd.dense(tree.leaf_n)
    exp = orc.all_pairs_f32(tree, dense)
    np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))


@pytest.mark.parametrize("n", [0, 1, 2, 7, 9, 33, 65])
def test_walk_sizes_and_tiles(cms, orc, ctx, n):
    """Sample counts that are not multiples of the 8-sample walk block."""
    _flush(f"walk sizes n={n}")
    tree, cp, counts, abd = cms.workload(0, samples=max(n, 1))
    abd = abd.subset(range(n)) if n else abd.subset([])
    ctx.set_opts(cms.KERNEL_WALK, cms.STORE_F32)
    ctx.set_tree(tree)
    ctx.set_samples_csr(abd)
    got = ctx.all_pairs()
    assert got.shape == (n * (n - 1) // 2,)
    if n >= 2:
        exp = orc.all_pairs_f32(tree, abd.dense(tree.leaf_n))
        np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))
        # unaligned rectangle
        r0, nr, c0, nc = 1, n - 1, 0, n - 1
        t = ctx.tile(r0, nr, c0, nc)
        rows, cols = np.meshgrid(np.arange(r0, r0 + nr), np.arange(c0, c0 + nc), indexing="ij")
        ref = np.ones((nr, nc), np.float32)
        m = rows != cols
        lo, hi = np.minimum(rows, cols), np.maximum(rows, cols)
        ref[m] = exp[cms.packed_index(lo[m], hi[m], n)]
        np.testing.assert_array_equal(t.view(np.uint32), ref.view(np.uint32))


def _many_live_tree(cms, cherries):
    """order.txt that keeps `cherries` registers live: every leaf pair is
    joined into its own register first, then the registers are folded into
    one (the walk's slots then exceed what it keeps in VGPRs)."""
    L = 2 * cherries
    o1, d1, o2, d2, od = [], [], [], [], []
    for i in range(cherries):
        o1.append(2 * i); d1.append(0.1 + 0.01 * i); o2.append(2 * i + 1); d2.append(0.2); od.append(-(i + 1))
    for i in range(1, cherries):
        o1.append(-1); d1.append(0.05); o2.append(-(i + 1)); d2.append(0.3 + 0.005 * i); od.append(-1)
    arr = lambda v, t: np.array(v, t)
    return cms.Tree(L, arr(o1, np.int32), arr(d1, np.float32), arr(o2, np.int32), arr(d2, np.float32),
                    arr(od, np.int32))


@pytest.mark.parametrize("cherries", [12, 40])
def test_walk_register_slots_vgpr_and_lds(cms, orc, ctx, cherries):
    """Schedules with up to 16 live registers run with the slots in VGPRs,
    larger ones with the slots in LDS; both bit-exact with the reference
    walk."""
    _flush(f"walk live registers {cherries}")
    tree = _many_live_tree(cms, cherries)
    rng = np.random.default_rng(cherries)
    abd = (rng.random((37, tree.leaf_n)) * (rng.random((37, tree.leaf_n)) < 0.5)).astype(np.float32)
    abd = (abd / np.maximum(abd.sum(1, keepdims=True), 1e-30) * 100.0).astype(np.float32)
    ctx.set_opts(cms.KERNEL_WALK, cms.STORE_F32)
    ctx.set_tree(tree)
    ctx.set_samples(abd)
    assert ctx.stats().nslots == cherries
    got = ctx.all_pairs()
    exp = orc.all_pairs_f32(tree, abd)
    np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))
    t = ctx.tile(3, 20, 0, 37)
    rows, cols = np.meshgrid(np.arange(3, 23), np.arange(37), indexing="ij")
    ref = np.ones((20, 37), np.float32)
    m = rows != cols
    ref[m] = exp[cms.packed_index(np.minimum(rows, cols)[m], np.maximum(rows, cols)[m], 37)]
    np.testing.assert_array_equal(t.view(np.uint32), ref.view(np.uint32))


def test_walk_edge_samples(cms, orc, ctx):
    """All-zero samples (sim 0), identical samples (sim ~1), disjoint samples."""
    _flush("walk edge samples")
    tree = cms.synth_tree(11, 500, 0.05)
    L = tree.leaf_n
    abd = np.zeros((6, L), np.float32)
    abd[1, :10] = 10.0
    abd[2, :10] = 10.0
    abd[3, 100:104] = 25.0
    abd[4, 400:] = 100.0 / (L - 400)
    abd[5, 0] = 100.0
    ctx.set_opts(cms.KERNEL_WALK, cms.STORE_F32)
    ctx.set_tree(tree)
    ctx.set_samples(abd)
    got = ctx.all_pairs()
    exp = orc.all_pairs_f32(tree, abd)
    np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))
    assert got[cms.packed_index(0, 1, 6)] == 0.0
    assert abs(got[cms.packed_index(1, 2, 6)] - 1.0) < 1e-5


def test_sharedmin_vs_fp64(cms, orc, ctx):
    _flush("sharedmin vs fp64 large tree")
    tree, cp, counts, abd = cms.workload(1, samples=96)
    ctx.set_opts(cms.KERNEL_SHAREDMIN)
    ctx.set_tree(tree)
    ctx.set_samples_csr(abd)
    got = ctx.all_pairs()
    n = abd.n
    ii, jj = np.triu_indices(n, 1)
    dense = abd.dense(tree.leaf_n)
    exp64 = orc.pairs_f64(tree, dense, ii.astype(np.int32), jj.astype(np.int32))
    err = np.abs(got.astype(np.float64) - exp64)
    _flush(f"sharedmin max err {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e}")
    assert err.max() <= SM_TOL
    assert np.sqrt((err ** 2).mean()) <= 1e-6
    # pair list and tile give the same integers, hence the same floats
    sel = np.arange(0, len(ii), 37)
    np.testing.assert_array_equal(ctx.sim_pairs(ii[sel], jj[sel]), got[sel])
    t = ctx.tile(3, 50, 40, 56)
    for a in range(0, 50, 7):
        for b in range(0, 56, 5):
            r, c = 3 + a, 40 + b
            if r != c:
                assert t[a, b] == got[cms.packed_index(min(r, c), max(r, c), n)]


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_sharedmin_golden(cms, orc, ctx, golden, case):
    _flush(f"sharedmin golden {case}")
    g = golden(case)
    tree, abd = _golden_inputs(cms, g, True)
    i, j, ref = g.hex_sims(True)
    ctx.set_opts(cms.KERNEL_SHAREDMIN)
    ctx.set_tree(tree)
    ctx.set_samples(abd)
    got = ctx.all_pairs()
    exp64 = orc.pairs_f64(tree, abd, i, j)
    n = abd.shape[0]
    assert np.abs(got[cms.packed_index(i, j, n)] - exp64).max() <= SM_TOL


@pytest.mark.parametrize("storage,tol", [(1, 2e-3), (2, 3e-4)])
def test_walk_16bit_storage(cms, orc, ctx, storage, tol):
    _flush(f"walk storage {storage}")
    tree, cp, counts, abd = cms.workload(0)
    ctx.set_opts(cms.KERNEL_WALK, storage)
    ctx.set_tree(tree)
    ctx.set_samples_csr(abd)
    got = ctx.all_pairs()
    n = abd.n
    ii, jj = np.triu_indices(n, 1)
    exp64 = orc.pairs_f64(tree, abd.dense(tree.leaf_n), ii.astype(np.int32), jj.astype(np.int32))
    err = np.abs(got - exp64)
    _flush(f"storage {storage}: max err {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e}")
    assert err.max() < tol


def test_errors_through_abi(cms, ctx):
    _flush("abi errors")
    bad = cms.Tree(4, np.array([0, -3], np.int32), np.zeros(2, np.float32), np.array([1, 2], np.int32),
                   np.zeros(2, np.float32), np.array([-1, -2], np.int32))
    with pytest.raises(cms.CmsError) as e:
        ctx.set_tree(bad)
    assert e.value.code == cms.E_SCHEDULE
    tree = cms.synth_tree(5, 50, 0.05)
    ctx.set_opts(cms.KERNEL_WALK)
    ctx.set_tree(tree)
    ctx.set_samples(np.full((3, 50), 2.0, np.float32))
    with pytest.raises(cms.CmsError) as e:
        ctx.sim_pairs([0, 3], [1, 1])
    assert e.value.code == cms.E_INVAL
    with pytest.raises(cms.CmsError):
        ctx.tile(2, 5, 0, 1)
