"""The precision, batched and serving forms on their own workloads: the
99,322-leaf tree (Exp(0.02) branch lengths, seeds 1000+k) with each
workload's sample generator, checked directly against the FP64 oracle (the
walk of comp.h:332-375 in double, orc_pairs_f64).

- Workload 2 (precision): samples with 200-2,000 nonzero OTUs, Zipf 1.1,
  NB(40, 0.4). The walk with F32, BF16 and F16 abundance storage (FP32
  arithmetic, Kahan total, FP64 final) and the fixed-point shared minimum,
  every pair of 96 samples.
- Workload 3 (batched): log-uniform 50-20,000 nonzero OTUs, Zipf 1.0,
  NB(30, 0.3). 2,048 samples binned by sparsity, autotuned persistent
  kernel, hipGraph-replayed bands streamed to host memory; 2,000 random
  pairs vs the oracle and the whole streamed triangle vs the unbinned
  all-pairs run.
- Workload 4 (serving): references from the same generator resident, query
  sets of 1-130 queries from another seed; random query-reference pairs vs
  the oracle, and each set's output bit-identical alone and in the mix.

Tolerances: 1e-5 max and 1e-6 RMS absolute for the F32 walk and the
shared-minimum forms; the 16-bit storage forms carry their storage rounding
(measured at this workload: BF16 max 9.3e-4 RMS 1.2e-4, F16 max 1.3e-4 RMS
1.6e-5), bounded here at BF16 2e-3 / 3e-4 and F16 3e-4 / 4e-5.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL, RMS = 1e-5, 1e-6


def _flush(msg):
    print(msg, flush=True)


def _oracle_pairs(cms, orc, tree, abd, ii, jj):
    """FP64 oracle on the pairs (ii[k], jj[k]) of `abd`, from a dense matrix
    of only the samples those pairs use."""
    ii, jj = np.asarray(ii), np.asarray(jj)
    used = np.unique(np.concatenate([ii, jj]))
    loc = np.full(abd.n, -1, np.int64)
    loc[used] = np.arange(len(used))
    dense = abd.subset(used).dense(tree.leaf_n)
    return orc.pairs_f64(tree, dense, loc[ii].astype(np.int32), loc[jj].astype(np.int32))


def _errs(got, exp):
    err = np.abs(np.asarray(got, np.float64) - exp)
    return float(err.max()), float(np.sqrt((err ** 2).mean()))


# ------------------------------------------------------------ workload 2

@pytest.fixture(scope="module")
def wl2(cms, gpu_ok):
    tree, cp, counts, abd = cms.workload(2, samples=96)
    return tree, abd


@pytest.mark.parametrize("form,storage,tmax,trms", [
    ("walk", 0, TOL, RMS), ("walk", 1, 2e-3, 3e-4), ("walk", 2, 3e-4, 4e-5), ("sharedmin", 0, TOL, RMS)])
def test_workload2_precision_vs_fp64(cms, orc, wl2, form, storage, tmax, trms):
    _flush(f"workload 2 {form} storage={storage}")
    tree, abd = wl2
    n = abd.n
    ii, jj = np.triu_indices(n, 1)
    exp = _oracle_pairs(cms, orc, tree, abd, ii, jj)
    kernel = cms.KERNEL_WALK if form == "walk" else cms.KERNEL_SHAREDMIN
    with cms.Context(0, kernel=kernel, storage=storage) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        got = ctx.all_pairs()
    mx, rms = _errs(got, exp)
    _flush(f"workload 2 {form} storage={storage}: max {mx:.3e} rms {rms:.3e} over {len(ii)} pairs")
    assert np.isfinite(got).all()
    assert mx <= tmax and rms <= trms


# ------------------------------------------------------------ workload 3

@pytest.fixture(scope="module")
def wl3(cms, gpu_ok):
    tree, cp, counts, abd = cms.workload(3, samples=2048)
    return tree, abd


def test_workload3_batched_stream_vs_fp64(cms, orc, wl3):
    _flush("workload 3 batched stream (bin, autotune, hipGraph) vs fp64")
    tree, abd = wl3
    n = abd.n
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        plain = ctx.all_pairs()
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1, autotune=1, use_graph=1) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        stored = np.full(n * (n - 1) // 2, np.nan, np.float32)
        bands = []

        def fn(first, count, sims):
            stored[first:first + count] = sims
            bands.append(count)

        ctx.all_pairs_stream(fn, band_pairs=400_000)
        order = ctx.sample_order()
        st = ctx.stats()
    assert st.bands == len(bands) >= 4
    assert not np.isnan(stored).any()
    # stored packed (p, q) -> caller's pair (order[p], order[q])
    pi, qi = np.triu_indices(n, 1)
    a, b = order[pi], order[qi]
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    caller = np.empty_like(stored)
    caller[cms.packed_index(lo, hi, n)] = stored
    # binned, tuned and replayed bands give the integers of the plain run
    np.testing.assert_array_equal(caller.view(np.uint32), plain.view(np.uint32))
    # and both are the oracle's similarities
    rng = np.random.default_rng(303)
    ri = rng.integers(0, n, 2000)
    rj = rng.integers(0, n, 2000)
    keep = ri != rj
    ri, rj = np.minimum(ri, rj)[keep], np.maximum(ri, rj)[keep]
    exp = _oracle_pairs(cms, orc, tree, abd, ri, rj)
    mx, rms = _errs(caller[cms.packed_index(ri, rj, n)], exp)
    _flush(f"workload 3: {len(ri)} random pairs max {mx:.3e} rms {rms:.3e}, {len(bands)} bands")
    assert mx <= TOL and rms <= RMS


# ------------------------------------------------------------ workload 4

@pytest.fixture(scope="module")
def wl4(cms, gpu_ok):
    tree, cp, counts, refs = cms.workload(4, samples=2048)
    w = cms.WORKLOADS[4]
    qc = cms.synth_counts(4004, 300, w["leaves"], w["nnz"][0], w["nnz"][1], w["zipf"], w["nb"][0], w["nb"][1],
                          w["log_uniform"])
    return tree, refs, cms.counts_to_abundance(qc, cp)


def test_workload4_serving_vs_fp64_and_mix(cms, orc, wl4):
    _flush("workload 4 serving vs fp64, alone vs mix")
    tree, refs, qs = wl4
    sizes = [1, 7, 40, 130, 3, 64, 19]
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    sets = [np.arange(s, s + k) for s, k in zip(starts, sizes)]
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(refs)
        with cms.Server(ctx, slots=256, bands=6) as srv:
            alone = [srv.submit(qs.subset(idx)).wait(timeout=120).copy() for idx in sets]
        # all sets submitted back to back: they join mid-cycle, share steps,
        # and wait for slots (256) and memory as others leave
        with cms.Server(ctx, slots=256, bands=6, memory_budget=256 << 20) as srv:
            reqs = []
            for idx in sets:
                reqs.append(srv.submit(qs.subset(idx)))
                time.sleep(0.001)
            mixed = [r.wait(timeout=120) for r in reqs]
            st = srv.stats()
    assert st.completed == len(sets) and st.failed == 0
    for a, m in zip(alone, mixed):
        np.testing.assert_array_equal(m.view(np.uint32), a.view(np.uint32))
    # random (query, reference) pairs of every set vs the oracle
    both = cms.Samples(*_stack(refs, qs))
    rng = np.random.default_rng(404)
    gi, gj, got = [], [], []
    for idx, out in zip(sets, alone):
        k = min(400, out.size)
        q = rng.integers(0, len(idx), k)
        r = rng.integers(0, refs.n, k)
        gi.append(refs.n + idx[q])
        gj.append(r)
        got.append(out[q, r])
    gi, gj, got = np.concatenate(gi), np.concatenate(gj), np.concatenate(got)
    exp = _oracle_pairs(cms, orc, tree, both, gi, gj)
    mx, rms = _errs(got, exp)
    _flush(f"workload 4: {len(gi)} query-reference pairs max {mx:.3e} rms {rms:.3e}")
    assert mx <= TOL and rms <= RMS


def _stack(a, b):
    rp = np.concatenate([a.rowptr, a.rowptr[-1] + b.rowptr[1:]]).astype(np.int64)
    return rp, np.concatenate([a.col, b.col]).astype(np.int32), np.concatenate([a.val, b.val]).astype(np.float32)


# ------------------------------------------------------------ workload 1

def test_workload1_full_triangle(cms, orc, gpu_ok):
    """The headline configuration at its full size: 10,000 samples, all
    49,995,000 pairs with the default shared-minimum path (binned, split
    tiles, leaf-list embedding input). 2,000 random pairs against the FP64
    oracle; over the whole triangle every value is finite and in [0, 1], and
    the bits equal a second run through different code (plain tiles,
    variant 1, and the embedding read from the F32 walk layout)."""
    tree, cp, counts, abd = cms.workload(1)
    del counts
    n = abd.n
    _flush(f"workload 1: {n} samples, nnz {abd.rowptr[-1]}")
    t0 = time.perf_counter()
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        got = ctx.all_pairs()
    _flush(f"default path {time.perf_counter() - t0:.1f} s")
    assert got.shape == (n * (n - 1) // 2,)
    assert np.isfinite(got).all()
    assert got.min() >= 0.0 and got.max() <= 1.0
    rng = np.random.default_rng(11)
    ii = rng.integers(0, n, 2000)
    jj = rng.integers(0, n, 2000)
    keep = ii != jj
    ii, jj = np.minimum(ii, jj)[keep], np.maximum(ii, jj)[keep]
    exp = _oracle_pairs(cms, orc, tree, abd, ii, jj)
    mx, rms = _errs(got[cms.packed_index(ii, jj, n)], exp)
    _flush(f"vs FP64 oracle on {len(ii)} pairs: max {mx:.3g} rms {rms:.3g}")
    assert mx <= TOL and rms <= RMS
    t0 = time.perf_counter()
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1, variant=1, embed_input=1) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        other = ctx.all_pairs()
    _flush(f"plain tiles, walk-layout input {time.perf_counter() - t0:.1f} s")
    np.testing.assert_array_equal(got.view(np.uint32), other.view(np.uint32))


def test_split_tiles_row_forms(cms, orc, gpu_ok):
    """The split tiles read a staged row through one buffer resource per
    bitmap window with the row's 32-bit byte offset (below 699,051 padded
    samples, where every list row times the stride times 4 fits in 32
    bits; a short last chunk's empty slots load the window's first row and
    are cleared), or from a 64-bit row address (larger sets, or
    cms_opts.wide_rows = 1, which takes it at any size). On workload 1's
    generator at 3,000 samples both forms must give the plain tiles' bits
    over the whole triangle, and the FP64 oracle on 2,000 random pairs."""
    tree, cp, counts, abd = cms.workload(1, samples=3000)
    del counts
    n = abd.n
    _flush(f"row forms: {n} samples, nnz {abd.rowptr[-1]}")
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1, variant=1) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        plain = ctx.all_pairs()
    rng = np.random.default_rng(12)
    ii = rng.integers(0, n, 2000)
    jj = rng.integers(0, n, 2000)
    keep = ii != jj
    ii, jj = np.minimum(ii, jj)[keep], np.maximum(ii, jj)[keep]
    exp = _oracle_pairs(cms, orc, tree, abd, ii, jj)
    for wide in (0, 1):
        with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1, wide_rows=wide) as ctx:
            ctx.set_tree(tree)
            ctx.set_samples_csr(abd)
            got = ctx.all_pairs()
            assert ctx.stats().variant >= 7  # a split-tile variant ran
        np.testing.assert_array_equal(got.view(np.uint32), plain.view(np.uint32))
        mx, rms = _errs(got[cms.packed_index(ii, jj, n)], exp)
        _flush(f"wide_rows={wide}: vs FP64 oracle on {len(ii)} pairs: max {mx:.3g} rms {rms:.3g}")
        assert mx <= TOL and rms <= RMS
