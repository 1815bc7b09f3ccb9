"""Batched form of the shared-minimum path: samples binned by sparsity,
persistent-kernel variants and the autotuner, banded streaming to host memory
(with and without hipGraph replay). All must give the same integers, hence
bit-identical floats, as the plain all-pairs run; the rectangle and pair-list
entry points must agree with it under binning. The plain run itself is
checked against the FP64 oracle (comp.h:332-375 and comp.h:378-446 in
double) on 2,000 random pairs, so every comparison below is anchored to the
oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data(cms, gpu_ok):
    # uneven sparsity (log-uniform nonzero counts) as in the batched workload
    tree = cms.synth_tree(1003, 20000, 0.02)
    cp = cms.synth_copy_numbers(3003, 20000, 15)
    counts = cms.synth_counts(2003, 203, 20000, 20, 4000, 1.0, 30, 0.3, log_uniform=True)
    return tree, cms.counts_to_abundance(counts, cp)


def _plain(cms, tree, abd, metric=0):
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, metric=metric) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        return ctx.all_pairs()


@pytest.mark.parametrize("metric", [0, 1])
def test_plain_vs_fp64(cms, orc, data, metric):
    """The reference run of this module vs the FP64 oracle: <= 1e-5 max,
    <= 1e-6 RMS absolute."""
    print(f"plain vs fp64 metric={metric}", flush=True)
    tree, abd = data
    got = _plain(cms, tree, abd, metric)
    n = abd.n
    rng = np.random.default_rng(41)
    ii = rng.integers(0, n, 2000)
    jj = rng.integers(0, n, 2000)
    keep = ii != jj
    ii, jj = ii[keep], jj[keep]
    lo, hi = np.minimum(ii, jj), np.maximum(ii, jj)
    dense = abd.dense(tree.leaf_n)
    if metric == 0:
        exp = orc.pairs_f64(tree, dense, lo.astype(np.int32), hi.astype(np.int32))
    else:
        _, exp = orc.pairs_unweight(tree, dense, lo.astype(np.int32), hi.astype(np.int32))
    err = np.abs(got[cms.packed_index(lo, hi, n)].astype(np.float64) - exp)
    assert err.max() <= 1e-5 and np.sqrt((err ** 2).mean()) <= 1e-6, (err.max(), np.sqrt((err ** 2).mean()))


@pytest.mark.parametrize("metric", [0, 1])
def test_binned_and_variants_identical(cms, data, metric):
    print(f"binned variants metric={metric}", flush=True)
    tree, abd = data
    ref = _plain(cms, tree, abd, metric)
    visited = None
    # 7..16 take the nearly empty entries out of the staged path; 13..16 share
    # each tile among 2 or 4 waves
    for variant in range(1, 17):
        with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, metric=metric, bin=1, variant=variant) as ctx:
            ctx.set_tree(tree)
            ctx.set_samples_csr(abd)
            got = ctx.all_pairs()
            st = ctx.stats()
            assert st.variant == variant
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
        # every variant visits the same entries; the sparse ones stage fewer
        visited = st.entries_visited if visited is None else visited
        assert st.entries_visited == visited
        assert st.entries_staged + st.entries_sparse + st.entries_half == st.entries_visited
        if variant >= 7:
            assert st.entries_sparse > 0 and st.sparse_pairs >= st.entries_sparse
            assert st.entries_half > 0 and st.half_pairs >= st.entries_half
        else:
            assert st.entries_sparse == 0 and st.entries_half == 0


def test_autotune_picks_a_variant(cms, data):
    print("autotune", flush=True)
    tree, abd = data
    ref = _plain(cms, tree, abd)
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1, autotune=1) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        got = ctx.all_pairs()
        st = ctx.stats()
        # 28 tiles: the tuner times the variants that share tiles among waves
        assert 13 <= st.variant <= 16
        assert all(t > 0 for t in st.tune_ms[12:]) and all(t == 0 for t in st.tune_ms[:12])
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("use_graph,bin_,autotune", [(0, 0, 0), (0, 1, 0), (1, 0, 0), (1, 1, 0), (0, 1, 1), (1, 1, 1)])
def test_stream_reassembles_triangle(cms, data, use_graph, bin_, autotune):
    # autotune: the tuner times variants inside the first band's buffer, which
    # is smaller than the whole triangle here
    print(f"stream graph={use_graph} bin={bin_} autotune={autotune}", flush=True)
    tree, abd = data
    ref = _plain(cms, tree, abd)
    n = abd.n
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=bin_, use_graph=use_graph, autotune=autotune) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        stored = np.full(n * (n - 1) // 2, np.nan, np.float32)
        seen = []

        def fn(first, count, sims):
            stored[first:first + count] = sims
            seen.append((first, count))

        ctx.all_pairs_stream(fn, band_pairs=3000)  # many small bands
        order = ctx.sample_order()
        assert ctx.stats().bands == len(seen) > 3
    assert sorted(order.tolist()) == list(range(n))
    assert not np.isnan(stored).any()
    # stored packed index (p, q) -> caller's pair (order[p], order[q])
    pi, qi = np.triu_indices(n, 1)
    a, b = order[pi], order[qi]
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    np.testing.assert_array_equal(stored.view(np.uint32), ref[cms.packed_index(lo, hi, n)].view(np.uint32))


def test_binned_tile_and_pairs(cms, data):
    print("binned tile/pairs", flush=True)
    tree, abd = data
    ref = _plain(cms, tree, abd)
    n = abd.n
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        t = ctx.tile(5, 40, 17, 50)
        r, c = np.meshgrid(np.arange(5, 45), np.arange(17, 67), indexing="ij")
        exp = np.ones((40, 50), np.float32)
        m = r != c
        exp[m] = ref[cms.packed_index(np.minimum(r, c)[m], np.maximum(r, c)[m], n)]
        np.testing.assert_array_equal(t.view(np.uint32), exp.view(np.uint32))
        rows = np.array([0, 7, 150, 202, 3], np.int32)
        cols = np.array([1, 200, 9, 0, 3], np.int32)
        got = ctx.sim_pairs(rows, cols)
        for k in range(4):
            a, b = sorted((int(rows[k]), int(cols[k])))
            assert got[k] == ref[cms.packed_index(a, b, n)]


@pytest.mark.parametrize("metric", [0, 1])
def test_sliced_embedding_identical(cms, data, metric):
    # the embedding built in slices of 64 stored positions (as for sample
    # sets whose walk layout exceeds the slice budget), and from the F32 walk
    # layout instead of the leaf lists (as for leaf streams too long for the
    # list builder), gives the same integers
    print(f"sliced embedding metric={metric}", flush=True)
    tree, abd = data
    ref = _plain(cms, tree, abd, metric)
    for bin_ in (0, 1):
        for slice_, input_ in ((64, 0), (0, 1), (64, 1)):
            with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, metric=metric, bin=bin_, embed_slice=slice_,
                             embed_input=input_) as ctx:
                ctx.set_tree(tree)
                ctx.set_samples_csr(abd)
                got = ctx.all_pairs()
            np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_unstored_entries_never_read(cms, data):
    """The embedding stores a row's entries of a 32-sample block only where
    the block has mass in that row; the rest of Y keeps what the memory held.
    Here that is a far denser sample set embedded first by the same context
    (same sizes, so the layout's memory is reused): the all-pairs tiles, the
    pair list and rectangles at unaligned offsets (rows from the union of
    two blocks' bitmaps, lanes masked by their own block's) must still give
    the bits of a fresh context."""
    print("unstored entries", flush=True)
    tree, abd = data
    n = abd.n
    ref = _plain(cms, tree, abd)
    cp = cms.synth_copy_numbers(3003, tree.leaf_n, 15)
    dense = cms.counts_to_abundance(cms.synth_counts(2903, n, tree.leaf_n, 15000, 19000, 1.0, 30, 0.3), cp)
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=0) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(dense)
        ctx.all_pairs()
        ctx.set_samples_csr(abd)
        got = ctx.all_pairs()
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
        for r0, nr, c0, nc in ((5, 40, 17, 50), (33, 70, 1, 90), (100, 103, 61, 142)):
            t = ctx.tile(r0, nr, c0, nc)
            r, c = np.meshgrid(np.arange(r0, r0 + nr), np.arange(c0, c0 + nc), indexing="ij")
            exp = np.ones((nr, nc), np.float32)
            m = r != c
            exp[m] = ref[cms.packed_index(np.minimum(r, c)[m], np.maximum(r, c)[m], n)]
            np.testing.assert_array_equal(t.view(np.uint32), exp.view(np.uint32))
        rows = np.arange(0, n, 3, dtype=np.int32)
        cols = ((rows * 7 + 11) % n).astype(np.int32)
        keep = rows != cols
        got = ctx.sim_pairs(rows[keep], cols[keep])
        lo, hi = np.minimum(rows[keep], cols[keep]), np.maximum(rows[keep], cols[keep])
        np.testing.assert_array_equal(got.view(np.uint32), ref[cms.packed_index(lo, hi, n)].view(np.uint32))


@pytest.fixture(scope="module")
def wide(cms, gpu_ok):
    # 20,500 samples: 205,761 block tiles, enough for the tuner to time its
    # 12 one-wave-per-tile variants on real tiles (24 segments of 4,096
    # spread over the triangle) and keep their results
    tree = cms.synth_tree(1077, 2000, 0.05)
    cp = cms.synth_copy_numbers(3077, 2000, 15)
    counts = cms.synth_counts(2077, 20500, 2000, 5, 60, 1.1, 30, 0.5)
    return tree, cms.counts_to_abundance(counts, cp)


def test_autotune_on_real_tiles(cms, wide):
    """The tuned segments' results stay in the output and the tiles between
    them are computed once: the same bits as a forced variant, for the
    all-pairs and the streamed entry points (whose tuning runs are
    discarded)."""
    print("autotune on real tiles", flush=True)
    tree, abd = wide
    n = abd.n
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1, variant=2) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        ref = ctx.all_pairs()
        rows_ref = ctx.stats().entries_visited
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1, autotune=1) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        got = ctx.all_pairs()
        st = ctx.stats()
        # 205,761 tiles: the variants with one wave per tile are timed
        assert all(t > 0 for t in st.tune_ms[:12]) and all(t == 0 for t in st.tune_ms[12:])
        assert 1 <= st.variant <= 12
        assert st.entries_visited == rows_ref  # tuned tiles counted once
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    for use_graph in (0, 1):
        with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1, autotune=1, use_graph=use_graph) as ctx:
            ctx.set_tree(tree)
            ctx.set_samples_csr(abd)
            stored = np.full(n * (n - 1) // 2, np.nan, np.float32)

            def fn(first, count, sims):
                stored[first:first + count] = sims

            ctx.all_pairs_stream(fn, band_pairs=190_000_000)  # first band: about 440 of 641 row blocks, ~198,000 tiles
            order = ctx.sample_order()
            assert ctx.stats().entries_visited == rows_ref
        assert not np.isnan(stored).any()
        pi, qi = np.triu_indices(n, 1)
        a, b = order[pi], order[qi]
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        np.testing.assert_array_equal(stored.view(np.uint32), ref[cms.packed_index(lo, hi, n)].view(np.uint32))
        del pi, qi, a, b, lo, hi


def test_default_shares_few_tiles(cms, data):
    """Without autotuning, a triangle of fewer than 4,096 block tiles runs the
    split tiles shared by four waves (variant 14); the result is the integers
    of one wave per tile (variant 8), bit for bit."""
    print("default variant with few tiles", flush=True)
    tree, abd = data
    outs = {}
    for variant in (0, 8):
        with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1, variant=variant) as ctx:
            ctx.set_tree(tree)
            ctx.set_samples_csr(abd)
            outs[variant] = ctx.all_pairs()
            st = ctx.stats()
            assert st.variant == (14 if variant == 0 else 8)
            assert st.entries_sparse > 0 and st.entries_half > 0
    np.testing.assert_array_equal(outs[0].view(np.uint32), outs[8].view(np.uint32))


def test_light_records_recounted(cms, data):
    """The light records are packed block after block in an allocation sized
    by their count, which a context counts once per sample set, metric and
    binning and reuses while those stay: a sparse set, then a denser one
    (more records: the allocation grows), then the sparse one again, the
    metric switched and the binning switched, all in one context, each give
    the bits of a fresh context."""
    print("light records recounted", flush=True)
    tree, abd = data
    n = abd.n
    cp = cms.synth_copy_numbers(3003, tree.leaf_n, 15)
    dense = cms.counts_to_abundance(cms.synth_counts(2904, n, tree.leaf_n, 3000, 8000, 1.0, 30, 0.3), cp)
    fresh = {}
    for name, s in (("sparse", abd), ("dense", dense)):
        for metric in (0, 1):
            for bin_ in (0, 1):
                with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, metric=metric, bin=bin_) as ctx:
                    ctx.set_tree(tree)
                    ctx.set_samples_csr(s)
                    fresh[name, metric, bin_] = ctx.all_pairs()
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, metric=0, bin=0) as ctx:
        ctx.set_tree(tree)
        for name, s, metric, bin_ in (("sparse", abd, 0, 0), ("dense", dense, 0, 0), ("sparse", abd, 0, 0),
                                      ("sparse", abd, 1, 0), ("sparse", abd, 1, 1), ("dense", dense, 1, 1),
                                      ("dense", dense, 0, 1)):
            ctx.set_opts(cms.KERNEL_SHAREDMIN, metric=metric, bin=bin_)
            ctx.set_samples_csr(s)
            for _ in range(2):  # the second preparation reuses the count
                ctx.prepare(force=True)
                got = ctx.all_pairs()
                np.testing.assert_array_equal(got.view(np.uint32), fresh[name, metric, bin_].view(np.uint32))
