"""Serving form on the GPU: resident references, continuously batched query
sets (cms_server_*).

- Every query-reference similarity is within 1e-5 of the FP64 oracle
  (weighted: the walk of comp.h:332-375 in FP64; unweighted: the presence-set
  ratio of comp.h:378-446), in the caller's reference order, with binned and
  unbinned references.
- Alone vs in the mix: a query set's output is bit-identical whether it runs
  alone or joins mid-cycle among other sets of different sizes, with query
  slots and device memory tight enough that admission waits.
- Errors are reported at submit (sizes, values, fixed-point range) and the
  context is read-only while it serves.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def data(cms, gpu_ok):
    tree = cms.synth_tree(1004, 20000, 0.02)
    cp = cms.synth_copy_numbers(3004, 20000, 15)
    refs = cms.counts_to_abundance(cms.synth_counts(2004, 300, 20000, 20, 4000, 1.0, 30, 0.3, log_uniform=True), cp)
    qs = cms.counts_to_abundance(cms.synth_counts(5004, 160, 20000, 20, 4000, 1.0, 30, 0.3, log_uniform=True), cp)
    return tree, refs, qs


def _oracle(cms, orc, tree, refs, q, metric):
    """FP64 similarity of every query of q to every reference: (nq, n_ref)."""
    both = np.vstack([refs.dense(tree.leaf_n), q.dense(tree.leaf_n)])
    ii, jj = np.meshgrid(np.arange(q.n) + refs.n, np.arange(refs.n), indexing="ij")
    ii, jj = ii.ravel().astype(np.int32), jj.ravel().astype(np.int32)
    if metric == cms.METRIC_WEIGHTED:
        s = orc.pairs_f64(tree, both, ii, jj)
    else:
        _, s = orc.pairs_unweight(tree, both, ii, jj)
    return s.reshape(q.n, refs.n)


def _ctx(cms, tree, refs, metric=0, bin_=1):
    ctx = cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, metric=metric, bin=bin_)
    ctx.set_tree(tree)
    ctx.set_samples_csr(refs)
    return ctx


@pytest.mark.parametrize("metric,bin_", [(0, 1), (0, 0), (1, 1)])
def test_serve_matches_fp64_oracle(cms, orc, data, metric, bin_):
    print(f"serve vs oracle metric={metric} bin={bin_}", flush=True)
    tree, refs, qs = data
    q = qs.subset(np.arange(40))
    exp = _oracle(cms, orc, tree, refs, q, metric)
    with _ctx(cms, tree, refs, metric, bin_) as ctx, cms.Server(ctx, slots=64, bands=3) as srv:
        got = srv.submit(q).wait(timeout=60)
        st = srv.stats()
    assert got.shape == (40, refs.n)
    assert np.isfinite(got).all()
    err = np.abs(got.astype(np.float64) - exp)
    assert err.max() <= TOL, err.max()
    assert st.completed == 1 and st.steps >= 3


@pytest.mark.parametrize("metric", [0, 1])
def test_cross_variants_identical(cms, data, metric):
    """Every cross-kernel variant (plain tiles 1..6, split tiles 7..8, which
    read the light entries of the query blocks and the references from
    records) gives the same integers: bit-identical outputs. Two query sets
    share the slots, so some query blocks hold inactive slots with stale
    data; bands=3 and 64 slots leave few tiles per step, so tile rows are
    split over several waves."""
    print(f"serve cross variants metric={metric}", flush=True)
    tree, refs, qs = data
    outs = {}
    with _ctx(cms, tree, refs, metric) as ctx:
        for variant in range(1, 9):
            with cms.Server(ctx, slots=64, bands=3, variant=variant) as srv:
                a = srv.submit(qs.subset(np.arange(45)))
                b = srv.submit(qs.subset(np.arange(45, 70)))
                outs[variant] = (a.wait(timeout=60).copy(), b.wait(timeout=60).copy())
    for variant in range(2, 9):
        for k in range(2):
            np.testing.assert_array_equal(outs[variant][k].view(np.uint32), outs[1][k].view(np.uint32))


def test_binned_and_unbinned_references_identical(cms, data):
    print("serve binned vs unbinned", flush=True)
    tree, refs, qs = data
    q = qs.subset(np.arange(17))
    outs = []
    for bin_ in (0, 1):
        with _ctx(cms, tree, refs, 0, bin_) as ctx, cms.Server(ctx, slots=32, bands=4) as srv:
            outs.append(srv.submit(q).wait(timeout=60).copy())
    np.testing.assert_array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("metric", [0, 1])
def test_alone_vs_in_the_mix(cms, data, metric):
    print(f"alone vs mix metric={metric}", flush=True)
    tree, refs, qs = data
    rng = np.random.default_rng(11 + metric)
    target = qs.subset(np.arange(100, 123))  # 23 queries
    with _ctx(cms, tree, refs, metric) as ctx:
        with cms.Server(ctx, slots=64, bands=5) as srv:
            alone = srv.submit(target).wait(timeout=60).copy()
        # the mix: sets of 1..40 queries arriving over time; 64 slots and a
        # budget of two to three sets' memory (about 14 MB each here) make
        # admission wait for slots and memory; the target arrives while
        # others are mid-cycle
        with cms.Server(ctx, slots=64, bands=5, memory_budget=40 << 20) as srv:
            reqs = []
            for k in range(12):
                n = int(rng.integers(1, 41))
                idx = rng.choice(np.arange(100), n, replace=False)
                reqs.append((idx, srv.submit(qs.subset(idx))))
                if k == 5:
                    tgt = srv.submit(target)
                time.sleep(0.002)
            mixed = tgt.wait(timeout=120)
            for idx, r in reqs:
                r.wait(timeout=120)
            st = srv.stats()
    np.testing.assert_array_equal(mixed.view(np.uint32), alone.view(np.uint32))
    assert st.completed == 13 and st.failed == 0
    assert st.arena_peak <= 40 << 20
    # every set in the mix also equals its rows of one big set run alone
    with _ctx(cms, tree, refs, metric) as ctx, cms.Server(ctx, slots=128, bands=2) as srv:
        full = srv.submit(qs.subset(np.arange(100))).wait(timeout=60)
        for idx, r in reqs:
            np.testing.assert_array_equal(r.out.view(np.uint32), full[idx].view(np.uint32))


def test_join_mid_cycle_overlaps(cms, data):
    print("join mid cycle", flush=True)
    tree, refs, qs = data
    with _ctx(cms, tree, refs) as ctx, cms.Server(ctx, slots=128, bands=8) as srv:
        a = srv.submit(qs.subset(np.arange(0, 60)))
        time.sleep(0.001)
        b = srv.submit(qs.subset(np.arange(60, 70)))
        a.wait(timeout=60)
        b.wait(timeout=60)
        ia, ib = a.info, b.info
    assert ia.join_step >= 0 and ib.join_step >= 0
    # both ran a full cycle of 8 bands; the later set joined at a later step
    assert ib.join_step >= ia.join_step
    assert ia.t_submit <= ia.t_admit <= ia.t_join <= ia.t_done
    assert ib.t_submit <= ib.t_admit <= ib.t_join <= ib.t_done
    for i in (ia, ib):  # device completion, when the server saw it before delivery
        assert i.t_computed == -1.0 or i.t_join <= i.t_computed <= i.t_done


def test_submit_errors_and_readonly_context(cms, data):
    print("serve errors", flush=True)
    tree, refs, qs = data
    with _ctx(cms, tree, refs) as ctx, cms.Server(ctx, slots=32, bands=2, memory_budget=64 << 20) as srv:
        with pytest.raises(cms.CmsError) as e:
            srv.submit(qs.subset(np.arange(33)))  # more queries than slots
        assert e.value.code == cms.E_INVAL
        bad = qs.subset(np.arange(2))
        bad.val = bad.val.copy()
        bad.val[0] = -1.0
        with pytest.raises(cms.CmsError) as e:
            srv.submit(bad)
        assert e.value.code == cms.E_UNSUPPORTED
        oob = qs.subset(np.arange(2))
        oob.col = oob.col.copy()
        oob.col[0] = tree.leaf_n
        with pytest.raises(cms.CmsError) as e:
            srv.submit(oob)
        assert e.value.code == cms.E_INVAL
        huge = qs.subset(np.arange(1))
        huge.val = huge.val * np.float32(1e6)  # mass far beyond the references' fixed-point range
        with pytest.raises(cms.CmsError) as e:
            srv.submit(huge)
        assert e.value.code == cms.E_RANGE
        with pytest.raises(cms.CmsError) as e:
            ctx.set_samples_csr(refs)
        assert e.value.code == cms.E_STATE
        # the server still works after rejected submissions
        out = srv.submit(qs.subset(np.arange(3))).wait(timeout=60)
        assert np.isfinite(out).all()
    # after the server is gone the context accepts changes again
    ctx2 = _ctx(cms, tree, refs)
    with cms.Server(ctx2, slots=32, bands=2):
        pass
    ctx2.set_samples_csr(refs)
    ctx2.close()
