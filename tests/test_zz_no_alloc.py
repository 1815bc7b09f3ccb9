"""cms_sim_pairs_noalloc: the pair list through the device's static scratch,
with no device allocation. Its output must equal cms_sim_pairs' bit for bit
(the same kernels), on two calls in a row, the second with fewer pairs, for
the walk (FP32, bit-exact with comp.h:332-375) and the shared-minimum
kernel (binned layout); and it refuses unprepared layouts and lists beyond
its scratch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data(cms, gpu_ok):
    tree, cp, counts, abd = cms.workload(1, samples=300)
    return tree, abd


def _pairs(n, k, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, k).astype(np.int32), rng.integers(0, n, k).astype(np.int32)


@pytest.mark.parametrize("kernel,bin_", [(0, 0), (1, 1), (1, 0)])
def test_noalloc_matches_sim_pairs(cms, data, kernel, bin_):
    print(f"noalloc kernel={kernel} bin={bin_}", flush=True)
    tree, abd = data
    with cms.Context(0, kernel=kernel, bin=bin_) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        ctx.prepare()
        for k, seed in ((20000, 1), (777, 2)):  # second call smaller: stale scratch must not leak in
            r, c = _pairs(abd.n, k, seed)
            exp = ctx.sim_pairs(r, c)
            got = ctx.sim_pairs_noalloc(r, c)
            assert got.shape == (k,)
            np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_noalloc_refusals(cms, data):
    print("noalloc refusals", flush=True)
    tree, abd = data
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        with pytest.raises(cms.CmsError) as e:  # layouts not built yet
            ctx.sim_pairs_noalloc([0], [1])
        assert e.value.code == cms.E_STATE
        ctx.prepare()
        big = np.zeros(cms.NOALLOC_MAX_PAIRS + 1, np.int32)
        with pytest.raises(cms.CmsError) as e:
            ctx.sim_pairs_noalloc(big, big)
        assert e.value.code == cms.E_INVAL
        with pytest.raises(cms.CmsError) as e:
            ctx.sim_pairs_noalloc([0], [abd.n])
        assert e.value.code == cms.E_INVAL
        # the largest list the scratch holds
        r, c = _pairs(abd.n, cms.NOALLOC_MAX_PAIRS, 3)
        np.testing.assert_array_equal(ctx.sim_pairs_noalloc(r, c).view(np.uint32), ctx.sim_pairs(r, c).view(np.uint32))
