"""ctypes wrapper of liboracle.so (cms_oracle.c) — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")


def build(ref: bool = True) -> None:
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


if not os.path.exists(LIB):
    build(ref=False)
_lib = C.CDLL(LIB)

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

_lib.orc_calc_sim_f32.restype = C.c_float
_lib.orc_calc_sim_f32.argtypes = [C.c_int, _i32p, _f32p, _i32p, _f32p, _i32p, _f32p, _f32p]
_lib.orc_calc_sim_f64.restype = C.c_double
_lib.orc_calc_sim_f64.argtypes = [C.c_int, _i32p, _f32p, _i32p, _f32p, _i32p, _f32p, _f32p]
_lib.orc_sharedmin_f64.restype = C.c_double
_lib.orc_sharedmin_f64.argtypes = [C.c_int, _i32p, _f32p, _i32p, _f32p, _i32p, _f32p, _f32p]
_lib.orc_load_abd.restype = C.c_int
_lib.orc_load_abd.argtypes = [C.c_int, _f32p, _i32p, _f32p, C.c_int, _f32p]
_lib.orc_all_pairs_f32.restype = C.c_long
_lib.orc_all_pairs_f32.argtypes = [C.c_int, C.c_int, _f32p, C.c_int, _i32p, _f32p, _i32p, _f32p, _i32p, _f32p,
                                   C.c_int, C.c_long]
_lib.orc_pairs_f32.restype = None
_lib.orc_pairs_f32.argtypes = [C.c_int, _f32p, C.c_int, _i32p, _f32p, _i32p, _f32p, _i32p, C.c_long, _i32p, _i32p,
                               _f32p, C.c_int]
_lib.orc_pairs_f64.restype = None
_lib.orc_pairs_f64.argtypes = [C.c_int, _f32p, C.c_int, _i32p, _f32p, _i32p, _f32p, _i32p, C.c_long, _i32p, _i32p,
                               _f64p, C.c_int, C.c_int]


_lib.orc_pairs_unweight.restype = None
_lib.orc_pairs_unweight.argtypes = [C.c_int, _f32p, C.c_int, _i32p, _f32p, _i32p, _f32p, _i32p, C.c_long, _i32p,
                                    _i32p, C.c_void_p, C.c_void_p, C.c_int]
_lib.orc_pairs_leaf.restype = None
_lib.orc_pairs_leaf.argtypes = [C.c_int, C.c_int, _f32p, C.c_long, _i32p, _i32p, C.c_void_p, C.c_void_p, C.c_int]
_lib.orc_leaf_sim_f32.restype = C.c_float
_lib.orc_leaf_sim_f32.argtypes = [C.c_int, C.c_int, _f32p, _f32p]


def _tree_args(tree):
    return (tree.order_n,) + tuple(np.ascontiguousarray(a) for a in tree.arrays())


def calc_sim_f32(tree, a, b) -> float:
    """comp.h:332-375 in FP32, same operation order."""
    return float(_lib.orc_calc_sim_f32(*_tree_args(tree), np.ascontiguousarray(a, np.float32),
                                       np.ascontiguousarray(b, np.float32)))


def calc_sim_f64(tree, a, b) -> float:
    return float(_lib.orc_calc_sim_f64(*_tree_args(tree), np.ascontiguousarray(a, np.float32),
                                       np.ascontiguousarray(b, np.float32)))


def sharedmin_f64(tree, a, b) -> float:
    return float(_lib.orc_sharedmin_f64(*_tree_args(tree), np.ascontiguousarray(a, np.float32),
                                        np.ascontiguousarray(b, np.float32)))


def load_abd(row, col_leaf, col_cp, leaf_n):
    out = np.zeros(leaf_n, np.float32)
    m = _lib.orc_load_abd(len(row), np.ascontiguousarray(row, np.float32), np.ascontiguousarray(col_leaf, np.int32),
                          np.ascontiguousarray(col_cp, np.float32), leaf_n, out)
    return out, m


def all_pairs_f32(tree, abd: np.ndarray, threads: int = 0, pairs_limit: int = 0):
    abd = np.ascontiguousarray(abd, np.float32)
    n = abd.shape[0]
    np_ = n * (n - 1) // 2
    out = np.zeros(max(np_, 1), np.float32)
    cnt = _lib.orc_all_pairs_f32(n, abd.shape[1], abd.reshape(-1), *_tree_args(tree), out, threads, pairs_limit)
    return out[:cnt]


def pairs_f32(tree, abd, row, col, threads: int = 0):
    abd = np.ascontiguousarray(abd, np.float32)
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    out = np.zeros(max(len(row), 1), np.float32)
    _lib.orc_pairs_f32(abd.shape[1], abd.reshape(-1), *_tree_args(tree), len(row), row, col, out, threads)
    return out[: len(row)]


def pairs_f64(tree, abd, row, col, closed_form: bool = False, threads: int = 0):
    abd = np.ascontiguousarray(abd, np.float32)
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    out = np.zeros(max(len(row), 1), np.float64)
    _lib.orc_pairs_f64(abd.shape[1], abd.reshape(-1), *_tree_args(tree), len(row), row, col, out, int(closed_form),
                       threads)
    return out[: len(row)]


def ref_binary(name: str) -> str | None:
    """Path of a compiled reference tool under oracle/_ref, or None."""
    p = os.path.join(REF_DIR, name)
    return p if os.path.exists(p) else None


def pairs_unweight(tree, abd, row, col, threads: int = 0):
    """comp.h:378-446 in FP32 (same order) and the FP64 presence-set ratio."""
    abd = np.ascontiguousarray(abd, np.float32)
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    s32 = np.zeros(max(len(row), 1), np.float32)
    s64 = np.zeros(max(len(row), 1), np.float64)
    _lib.orc_pairs_unweight(abd.shape[1], abd.reshape(-1), *_tree_args(tree), len(row), row, col,
                            s32.ctypes.data, s64.ctypes.data, threads)
    return s32[: len(row)], s64[: len(row)]


def pairs_leaf(mode, abd, row, col, threads: int = 0):
    """comp -M 2..5 (dist.h:13-159 via comp.h:296-330): FP32 as the reference
    and FP64 from the definitions."""
    abd = np.ascontiguousarray(abd, np.float32)
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    s32 = np.zeros(max(len(row), 1), np.float32)
    s64 = np.zeros(max(len(row), 1), np.float64)
    _lib.orc_pairs_leaf(int(mode), abd.shape[1], abd.reshape(-1), len(row), row, col, s32.ctypes.data,
                        s64.ctypes.data, threads)
    return s32[: len(row)], s64[: len(row)]
