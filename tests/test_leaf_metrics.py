"""Non-tree metrics of comp -M 2..5: cosine, Euclidean, Jensen-Shannon and
Bray-Curtis over the leaf abundances (dist.h:13-159, dispatched by
_Comp_Tree::Calc_sim(a, b, mode), comp.h:296-330; similarity = 1 - dist).

CPU: the oracle's FP32 restatement equals the compiled reference bit for bit
on every golden pair (hex fixtures), its FP64 form agrees with it and with
closed forms (identical, disjoint and degenerate samples).
GPU: the leaf-tile kernels are within 1e-5 of the FP64 oracle (with the
reference's degenerate-sample semantics) and `cuda-comp -M 2..5` prints the
reference's .dist values to the 6-digit print.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN_CASES

MODES = [2, 3, 4, 5]
TOL = 1e-5


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
