"""Non-tree metrics of comp -M 2..5: cosine, Euclidean, Jensen-Shannon and
Bray-Curtis over the leaf abundances (dist.h:13-159, dispatched by
_Comp_Tree::Calc_sim(a, b, mode), comp.h:296-330; similarity = 1 - dist).

CPU: the oracle's FP32 restatement equals the compiled reference bit for bit
on every golden pair (hex fixtures), its FP64 form agrees with it and with
closed forms (identical, disjoint and degenerate samples).
GPU, two kernels:
- the merge kernels (CMS_KERNEL_WALK, cuda-comp's default) follow the
  reference's FP32 operation order and reproduce the host C library's logf
  for JSD: all four equal the reference (golden hex) and the oracle's FP32
  restatement bit for bit, and `cuda-comp -M 2..5` writes the reference's
  .dist bytes;
- the leaf tiles (CMS_KERNEL_SHAREDMIN) are within 1e-5 of the FP64 oracle,
  with the reference's degenerate-sample semantics.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN_CASES

MODES = [2, 3, 4, 5]
TOL = 1e-5
KERNELS = ["merge", "tiles"]


def _hex(g, m):
    rows = [ln.split() for ln in g.read(f"hex_M{m}.txt").decode().splitlines() if ln.strip()]
    i = np.array([int(r[0]) for r in rows], np.int32)
    j = np.array([int(r[1]) for r in rows], np.int32)
    s = np.array([float.fromhex(r[2]) for r in rows], np.float32)
    return i, j, s


def _abd(cms, g):
    db = cms.Database(g.root, "G")
    abd, _ = cms.Table(g.table).abundance(db, True)
    return abd


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("mode", MODES)
def test_oracle_bit_exact_vs_reference(cms, orc, golden, case, mode):
    g = golden(case)
    abd = _abd(cms, g)
    i, j, s = _hex(g, mode)
    s32, s64 = orc.pairs_leaf(mode, abd, i, j)
    np.testing.assert_array_equal(s32.view(np.uint32), s.view(np.uint32))
    ok = np.isfinite(s64)
    assert np.abs(s64[ok] - s[ok]).max() < 5e-5  # FP32 rounding of the reference


def test_closed_forms(orc):
    a = np.array([[1, 0, 3, 0, 0],    # 0
                  [2, 0, 6, 0, 0],    # 1: a multiple of 0
                  [0, 5, 0, 1, 0],    # 2: disjoint from 0
                  [0, 0, 0, 0, 0],    # 3: all zero
                  [1, 1, 1, 1, 0]], np.float32)
    pairs = np.array([[0, 1], [0, 2], [0, 3], [3, 3], [0, 4]], np.int32)
    r = {m: orc.pairs_leaf(m, a, pairs[:, 0], pairs[:, 1])[1] for m in MODES}
    # identical distributions: similarity 1 for every metric
    for m in MODES:
        assert abs(r[m][0] - 1.0) < 1e-12
    # disjoint supports: cosine 0, Bray-Curtis 0, JSD 1 - sqrt(ln 2),
    # Euclidean 1 - sqrt(|p|^2 + |q|^2)
    assert abs(r[2][1]) < 1e-12 and abs(r[5][1]) < 1e-12
    assert abs(r[4][1] - (1 - math.sqrt(math.log(2)))) < 1e-12
    assert abs(r[3][1] - (1 - math.sqrt((1 / 16 + 9 / 16) + (25 / 36 + 1 / 36)))) < 1e-12
    # an all-zero sample: 0 for cosine, Euclidean and Bray-Curtis; JSD NaN
    # against a non-empty sample and 1 against another empty one
    for m in (2, 3, 5):
        assert r[m][2] == 0.0
    assert math.isnan(r[4][2]) and r[4][3] == 1.0
    # p = (1/4, 0, 3/4, 0), q = 1/4 each: Bray-Curtis = sum min = 1/4 + 1/4
    assert abs(r[5][4] - 0.5) < 1e-12
    assert abs(r[2][4] - 4 / (math.sqrt(10) * 2)) < 1e-12


def test_planted_mistake_moves_jsd(orc):
    # the FP64 JSD against a version with the mixture taken as p + q (no 1/2):
    # the comparison above would catch a dropped factor
    rng = np.random.default_rng(3)
    a = rng.gamma(0.5, 1.0, (4, 50)).astype(np.float32) * (rng.random((4, 50)) < 0.4)
    _, s64 = orc.pairs_leaf(4, a, np.array([0, 1], np.int32), np.array([2, 3], np.int32))
    p = a[0] / a[0].sum()
    q = a[2] / a[2].sum()
    m = p + q
    wrong = 0.0
    for x, y, z in zip(p, q, m):
        if x > 0:
            wrong += x * math.log(x / z)
        if y > 0:
            wrong += y * math.log(y / z)
    assert abs((1 - math.sqrt(max(wrong, 0) / 2)) - s64[0]) > 1e-3


# ------------------------------------------------------------------ GPU

def _leaf_tree(cms, leaf_n):
    z = np.zeros(0, np.int32)
    f = np.zeros(0, np.float32)
    return cms.Tree(leaf_n, z, f, z, f, z)


def _same(got, exp, tol=TOL):
    """|got - exp| <= tol where exp is finite; NaN exactly where exp is NaN."""
    got = np.asarray(got, np.float64)
    nan = np.isnan(exp)
    np.testing.assert_array_equal(np.isnan(got), nan)
    err = np.abs(got[~nan] - exp[~nan])
    return float(err.max()) if err.size else 0.0


def _kernel(cms, name):
    return cms.KERNEL_WALK if name == "merge" else cms.KERNEL_SHAREDMIN


def _check_reference_order(got, s32, mode):
    """Merge kernels against the reference's FP32 values, bit for bit (NaN
    bits included)."""
    got = np.asarray(got, np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), s32.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kernel", KERNELS)
def test_gpu_golden(cms, orc, gpu_ok, golden, case, mode, kernel):
    """Golden inputs through cms_all_pairs: the merge kernels reproduce the
    reference's FP32 values; the leaf tiles are within 1e-5 of them and of the
    FP64 definitions."""
    print(f"leaf metric -M {mode} golden {case} {kernel}", flush=True)
    g = golden(case)
    db = cms.Database(g.root, "G")
    abd, _ = cms.Table(g.table).abundance(db, True)
    i, j, ref = _hex(g, mode)
    with cms.Context(0, kernel=_kernel(cms, kernel), metric=mode) as c:
        c.set_tree(db.tree())
        c.set_samples(abd)
        got = c.all_pairs()
    n = abd.shape[0]
    p = cms.packed_index(i, j, n)
    if kernel == "merge":
        _check_reference_order(got[p], ref, mode)
        return
    _, s64 = orc.pairs_leaf(mode, abd, i, j)
    assert _same(got[p], ref.astype(np.float64)) <= TOL
    assert _same(got[p], s64) <= TOL


def _synthetic(seed, n, leaf_n):
    rng = np.random.default_rng(seed)
    a = (rng.gamma(0.3, 10.0, (n, leaf_n)) * (rng.random((n, leaf_n)) < 0.05)).astype(np.float32)
    a[3] = 0.0                      # an all-zero sample
    a[7] = 0.0                      # and another
    a[11] = a[12] * 3.0             # a multiple of another sample
    a[20] = 0.0
    a[20, :5] = [1e-30, 1.0, 2.0, 0.0, 5.0]   # a mass far below 2^-24 of the rest
    a[21] = 0.0
    a[21, :5] = [4.0, 1e-30, 0.0, 7.0, 0.0]
    a[30] = 0.0
    a[30, leaf_n - 1] = 2.5         # one leaf, the last one
    return a


@pytest.mark.gpu
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,leaf_n", [(2, 40), (45, 33), (203, 3000), (513, 20000)])
def test_gpu_synthetic_vs_fp64(cms, orc, gpu_ok, mode, n, leaf_n):
    """Ragged sample counts (tiles past n), leaf counts across bitmap windows,
    degenerate samples; no tree (order_n 0) is accepted for these metrics."""
    print(f"leaf metric -M {mode} synthetic n={n} leaves={leaf_n} tiles", flush=True)
    a = _synthetic(5 + n, max(n, 32), leaf_n)[:n]
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, metric=mode) as c:
        c.set_tree(_leaf_tree(cms, leaf_n))
        c.set_samples(a)
        got = c.all_pairs()
        got2 = c.all_pairs()        # a second call reuses the prepared layout
    i, j = np.triu_indices(n, 1)
    _, s64 = orc.pairs_leaf(mode, a, i.astype(np.int32), j.astype(np.int32))
    assert _same(got, s64) <= TOL
    np.testing.assert_array_equal(got.view(np.uint32), got2.view(np.uint32))
    if mode == 4 and n > 7:
        # the reference's NaN for one empty sample, bit for bit (0/0 on the host)
        k = cms.packed_index(np.array([0]), np.array([3]), n)[0]
        assert got.view(np.uint32)[k] == 0xFFC00000
        k = cms.packed_index(np.array([3]), np.array([7]), n)[0]
        assert got[k] == 1.0


@pytest.mark.gpu
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,leaf_n", [(2, 40), (45, 33), (203, 3000), (513, 20000)])
def test_gpu_merge_synthetic_vs_fp32_oracle(cms, orc, gpu_ok, mode, n, leaf_n):
    """The merge kernels on the synthetic sets (degenerate samples, a mass far
    below the rest, a single last leaf) against the oracle's FP32
    restatement of dist.h; a second call and a kernel switch in between give
    the same bits."""
    print(f"leaf metric -M {mode} synthetic n={n} leaves={leaf_n} merge", flush=True)
    a = _synthetic(5 + n, max(n, 32), leaf_n)[:n]
    with cms.Context(0, metric=mode) as c:
        c.set_tree(_leaf_tree(cms, leaf_n))
        c.set_samples(a)
        got = c.all_pairs()
        c.set_opts(kernel=cms.KERNEL_SHAREDMIN, metric=mode)
        c.all_pairs()
        c.set_opts(kernel=cms.KERNEL_WALK, metric=mode)
        got2 = c.all_pairs()
    i, j = np.triu_indices(n, 1)
    s32, _ = orc.pairs_leaf(mode, a, i.astype(np.int32), j.astype(np.int32))
    _check_reference_order(got, s32, mode)
    np.testing.assert_array_equal(got.view(np.uint32), got2.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("wl", [1, 3])
def test_gpu_merge_workload_trees(cms, orc, gpu_ok, wl):
    """The merge kernels at the workloads' shape: 99,322 leaves, 512 samples
    of workload 1 (200-2,000 entries) and workload 3 (log-uniform 50-20,000
    entries, so the visiting order and ragged list lengths matter); 3,000
    random pairs of the triangle against the oracle's FP32 restatement."""
    print(f"leaf merge workload {wl} 512 samples", flush=True)
    tree, _, _, abd = cms.workload(wl, samples=512)
    dense = abd.dense(tree.leaf_n)
    n = dense.shape[0]
    rng = np.random.default_rng(40 + wl)
    i = rng.integers(0, n - 1, 3000)
    j = (i + 1 + rng.integers(0, n - 1 - i)).astype(np.int64)
    p = cms.packed_index(i, j, n)
    with cms.Context(0, metric=2) as c:
        c.set_tree(_leaf_tree(cms, tree.leaf_n))
        c.set_samples_csr(abd)
        for mode in MODES:
            c.set_opts(kernel=cms.KERNEL_WALK, metric=mode)
            got = c.all_pairs()
            s32, _ = orc.pairs_leaf(mode, dense, i.astype(np.int32), j.astype(np.int32))
            _check_reference_order(got[p], s32, mode)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [2, 5])
def test_gpu_merge_switches_layout(cms, orc, gpu_ok, mode):
    """Cosine merges the values and the others a / sum a: switching between
    them rebuilds the entry lists."""
    a = _synthetic(17, 64, 900)
    i, j = np.triu_indices(64, 1)
    with cms.Context(0, metric=mode) as c:
        c.set_tree(_leaf_tree(cms, 900))
        c.set_samples(a)
        c.all_pairs()
        other = 5 if mode == 2 else 2
        c.set_opts(kernel=cms.KERNEL_WALK, metric=other)
        got = c.all_pairs()
    s32, _ = orc.pairs_leaf(other, a, i.astype(np.int32), j.astype(np.int32))
    np.testing.assert_array_equal(got.view(np.uint32), s32.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", KERNELS)
def test_gpu_leaf_metrics_csr_and_guards(cms, orc, gpu_ok, kernel):
    """CSR input equals dense input bit for bit; entry points other than
    cms_all_pairs(_device) reject -M 2..5 with an error; negative abundances
    are rejected."""
    a = _synthetic(9, 96, 500)
    rowptr = np.zeros(a.shape[0] + 1, np.int64)
    nz = [np.nonzero(r)[0] for r in a]
    rowptr[1:] = np.cumsum([len(x) for x in nz])
    col = np.concatenate(nz).astype(np.int32)
    val = np.concatenate([a[k, x] for k, x in enumerate(nz)]).astype(np.float32)
    with cms.Context(0, kernel=_kernel(cms, kernel), metric=cms.METRIC_JSD) as c:
        c.set_tree(_leaf_tree(cms, 500))
        c.set_samples(a)
        dense = c.all_pairs()
        c.set_samples_csr(cms.Samples(rowptr, col, val))
        csr = c.all_pairs()
        np.testing.assert_array_equal(dense.view(np.uint32), csr.view(np.uint32))
        with pytest.raises(cms.CmsError):
            c.sim_pairs(np.array([0], np.int32), np.array([1], np.int32))
        with pytest.raises(cms.CmsError):
            c.tile(0, 4, 0, 4)
        b = a.copy()
        b[5, 3] = -1.0
        c.set_samples(b)
        with pytest.raises(cms.CmsError):
            c.all_pairs()


def _dist_values(path):
    rows = [ln.split("\t") for ln in open(path).read().splitlines()]
    return rows[0], [r[0] for r in rows[1:]], np.array([[float(x) for x in r[1:]] for r in rows[1:]])


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("mode", MODES)
def test_cuda_comp_leaf_metric_bytes(cms, gpu_ok, golden, tmp_path, case, mode):
    """cuda-comp -M 2..5 (default kernel) writes the .dist file that the
    reference's values give through the .dist writer, byte for byte."""
    print(f"cuda-comp -M {mode} {case} bytes", flush=True)
    g = golden(case)
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    out = str(tmp_path / "o.dist")
    env = dict(os.environ, ParallelMETA=g.root)
    r = subprocess.run([exe, "-T", g.table, "-D", "G", "-o", out, "-M", str(mode)], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    names = cms.Table(g.table).sample_names()
    n = len(names)
    i, j, s = _hex(g, mode)
    packed = np.zeros(n * (n - 1) // 2, np.float32)
    packed[cms.packed_index(i, j, n)] = s
    assert len(i) == packed.size
    exp = str(tmp_path / "exp.dist")
    cms.write_dist(exp, names, packed, False)
    assert open(out, "rb").read() == open(exp, "rb").read()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kernel", ["walk", "sharedmin"])
def test_cuda_comp_leaf_metric(cms, gpu_ok, golden, tmp_path, mode, kernel):
    """cuda-comp -M m prints 1 - sim of the reference's values (6 digits),
    with either kernel."""
    print(f"cuda-comp -M {mode} --kernel {kernel}", flush=True)
    g = golden("core")
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    out = str(tmp_path / "o.dist")
    env = dict(os.environ, ParallelMETA=g.root)
    r = subprocess.run([exe, "-T", g.table, "-D", "G", "-o", out, "-M", str(mode), "--kernel", kernel], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    hdr, names, d = _dist_values(out)
    ref_hdr = g.read("dist_rT.dist").decode().splitlines()[0].split("\t")
    assert hdr == ref_hdr
    i, j, s = _hex(g, mode)
    exp = 1.0 - s.astype(np.float64)
    assert _same(d[i, j], exp, 1e-5 + 5e-6) <= 1.5e-5
    assert _same(d[j, i], exp, 1e-5 + 5e-6) <= 1.5e-5
    assert np.all(np.diag(d) == 0.0)


@pytest.mark.gpu
def test_cuda_comp_no_tree_falls_back_to_jsd(cms, gpu_ok, golden, tmp_path):
    """A database without a tree runs -M 0 as -M 4 (comp.h:299-302)."""
    import shutil
    g = golden("config0")
    root = str(tmp_path / "root")
    shutil.copytree(g.root, root)
    cfg = os.path.join(root, "databases", "db.config")
    lines = open(cfg).read().splitlines()
    k = lines.index("#tree")
    assert lines[k + 1] == "Y"
    lines[k + 1] = "N"
    open(cfg, "w").write("\n".join(lines) + "\n")
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    env = dict(os.environ, ParallelMETA=root)
    outs = []
    for m in ("0", "4"):
        out = str(tmp_path / f"o{m}.dist")
        r = subprocess.run([exe, "-T", os.path.join(root, "table.txt"), "-D", "G", "-o", out, "-M", m], env=env,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(open(out, "rb").read())
    assert outs[0] == outs[1]
