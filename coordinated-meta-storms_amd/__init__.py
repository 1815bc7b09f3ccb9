"""Python binding of libcms, the MI355X-native all-pairs Meta-Storms library.

The package mirrors the reference's host interface for the `-T` path
(qdu-bioinfo/Coordinated-Meta-Storms):

* ``Database`` - `_PMDB` + `_Comp_Tree` tree loading (db.h:21-358,
  comp.h:122-192): db.config, tree/id.txt, tree/order.txt, copy_number.txt.
* ``Table`` - `_Table_Format::Load_Table` (table_format.h:110-165).
* ``Table.abundance`` - `_Comp_Tree::Load_abd` (comp.h:243-279).
* ``Context.all_pairs`` - the all-pairs loop of `Multi_GPU_split_MetaStorm`
  (cms_cuda.cu:1470-2235) / `Multi_Comp_Table` (comp_sam.cpp:303-369).
* ``Context.sim_pairs`` - the reference kernels' pair-list boundary
  (`CalcSimOfDiagonalSquare`, cms_cuda.cu:52-54).
* ``write_dist`` - `Output_Matrix` (cms_cuda.cu:1305-1342).

Everything is a thin ctypes layer over the C ABI in include/cms.h; the
distance computation runs only in the gfx950 kernels of libcms.so. If the
library is missing the import raises (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcms.so")
BIN_DIR = os.path.join(_HERE, "bin")

OK = 0
E_INVAL, E_NOMEM, E_HIP, E_SCHEDULE, E_STATE, E_IO, E_UNSUPPORTED = -1, -2, -3, -4, -5, -6, -7
#: pairs per cms_sim_pairs_noalloc call (CMS_NOALLOC_MAX_PAIRS)
NOALLOC_MAX_PAIRS = 1 << 20
E_RANGE, E_TIMEOUT = -8, -9
KERNEL_WALK, KERNEL_SHAREDMIN = 0, 1
METRIC_WEIGHTED, METRIC_UNWEIGHTED = 0, 1
# comp -M 2..5 (dist.h): leaf-abundance metrics, no tree needed
METRIC_COS, METRIC_EU, METRIC_JSD, METRIC_BC = 2, 3, 4, 5
STORE_F32, STORE_BF16, STORE_F16 = 0, 1, 2
REG_SIZE = 150


class CmsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libcms error {code}: {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(f"libcms.so not built at {LIB_PATH}; run __graft_entry__.build() "
                      "(make -C coordinated-meta-storms_amd/csrc)")
lib = C.CDLL(LIB_PATH)

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class Opts(C.Structure):
    _fields_ = [("kernel", C.c_int), ("storage", C.c_int), ("tile_a", C.c_int), ("tile_b", C.c_int),
                ("chunk", C.c_int), ("use_graph", C.c_int), ("metric", C.c_int), ("bin", C.c_int),
                ("autotune", C.c_int), ("variant", C.c_int), ("embed_slice", C.c_int),
                ("embed_input", C.c_int), ("wide_rows", C.c_int), ("reserved", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("prep_ms", C.c_double), ("pair_node_updates", C.c_double),
                ("dense_updates", C.c_double), ("valu_ops", C.c_double), ("bytes_moved", C.c_double),
                ("tiles", C.c_int64), ("nslots", C.c_int), ("scale_bits", C.c_int), ("entries", C.c_int),
                ("variant", C.c_int), ("tune_ms", C.c_double * 16), ("bands", C.c_int64),
                ("entries_visited", C.c_int64), ("entries_staged", C.c_int64), ("entries_sparse", C.c_int64),
                ("sparse_pairs", C.c_int64), ("entries_half", C.c_int64), ("half_pairs", C.c_int64)]


class ServerOpts(C.Structure):
    _fields_ = [("slots", C.c_int), ("bands", C.c_int), ("memory_budget", C.c_int64), ("variant", C.c_int),
                ("reserved", C.c_int * 3)]


class RequestInfo(C.Structure):
    _fields_ = [("rc", C.c_int), ("nq", C.c_int), ("join_step", C.c_int64), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_join", C.c_double), ("t_done", C.c_double),
                ("t_computed", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64), ("queued", C.c_int64),
                ("active", C.c_int64), ("steps", C.c_int64), ("kernel_ms", C.c_double), ("pairs", C.c_double),
                ("pair_entries", C.c_double), ("arena_bytes", C.c_int64), ("arena_peak", C.c_int64),
                ("slots", C.c_int), ("slots_peak", C.c_int), ("bands", C.c_int), ("uptime_s", C.c_double),
                ("entries_staged", C.c_double), ("entries_sparse", C.c_double), ("sparse_pairs", C.c_double),
                ("entries_half", C.c_double), ("half_pairs", C.c_double)]


class SynthOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_samples", C.c_int), ("leaf_n", C.c_int), ("nnz_lo", C.c_int),
                ("nnz_hi", C.c_int), ("nnz_log_uniform", C.c_int), ("zipf_s", C.c_double),
                ("nb_mean", C.c_double), ("nb_dispersion", C.c_double)]


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("cms_opts_default", None, [C.POINTER(Opts)])
_sig("cms_create", C.c_int, [C.c_int, C.POINTER(Opts), C.POINTER(_vp)])
_sig("cms_destroy", None, [_vp])
_sig("cms_last_error", C.c_char_p, [_vp])
_sig("cms_set_opts", C.c_int, [_vp, C.POINTER(Opts)])
_sig("cms_set_tree", C.c_int, [_vp, C.c_int, C.c_int, _i32p, _f32p, _i32p, _f32p, _i32p])
_sig("cms_set_samples", C.c_int, [_vp, C.c_int, _f32p])
_sig("cms_set_samples_csr", C.c_int, [_vp, C.c_int, _i64p, _i32p, _f32p])
_sig("cms_all_pairs", C.c_int, [_vp, _f32p])
_sig("cms_all_pairs_device", C.c_int, [_vp, _vp, _vp])
_sig("cms_copy_result", C.c_int, [_vp, _f32p, C.c_int64])
_sig("cms_synchronize", C.c_int, [_vp])
_sig("cms_device_count", C.c_int, [])
_sig("cms_calibrate_copy", C.c_int, [_vp, C.c_int64, C.c_int, C.POINTER(C.c_double)])
_BAND_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_float))
_sig("cms_all_pairs_stream", C.c_int, [_vp, C.c_int64, _BAND_FN, _vp])
_sig("cms_sample_order", C.c_int, [_vp, _i32p])
_sig("cms_sim_pairs", C.c_int, [_vp, C.c_int64, _i32p, _i32p, _f32p])
_sig("cms_sim_pairs_noalloc", C.c_int, [_vp, C.c_int64, _i32p, _i32p, _f32p])
_sig("cms_tile", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _f32p])
_sig("cms_get_stats", C.c_int, [_vp, C.POINTER(Stats)])
_sig("cms_prepare", C.c_int, [_vp, C.c_int])
_sig("cms_validate_tree", C.c_int, [C.c_int, C.c_int, _i32p, _f32p, _i32p, _f32p, _i32p, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)])
_sig("cms_db_open", C.c_int, [C.c_char_p, C.c_char, C.POINTER(_vp)])
_sig("cms_db_close", None, [_vp])
_sig("cms_db_leaf_n", C.c_int, [_vp])
_sig("cms_db_order_n", C.c_int, [_vp])
_sig("cms_db_has_tree", C.c_int, [_vp])
_sig("cms_db_has_cp", C.c_int, [_vp])
_sig("cms_db_order", C.c_int, [_vp] + [C.POINTER(_vp)] * 5)
_sig("cms_db_leaf_name", C.c_char_p, [_vp, C.c_int])
_sig("cms_table_load", C.c_int, [C.c_char_p, C.POINTER(_vp)])
_sig("cms_table_from_dense", C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), _f32p,
                                      C.POINTER(_vp)])
_sig("cms_table_free", None, [_vp])
_sig("cms_table_n", C.c_int, [_vp])
_sig("cms_table_f", C.c_int, [_vp])
_sig("cms_table_sample_name", C.c_char_p, [_vp, C.c_int])
_sig("cms_table_feature_name", C.c_char_p, [_vp, C.c_int])
_sig("cms_table_skipped", C.c_int, [_vp])
_sig("cms_table_save_csr", C.c_int, [_vp, C.c_char_p])
_sig("cms_table_is_sparse", C.c_int, [_vp])
_sig("cms_table_from_csr", C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), _i64p,
                                    _i32p, _f32p, C.POINTER(_vp)])
_sig("cms_table_abundance", C.c_int, [_vp, _vp, C.c_int, _f32p, _vp])
_sig("cms_table_abundance_csr", C.c_int, [_vp, _vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)])
_sig("cms_write_dist", C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), _f32p, C.c_int])
_sig("cms_sample_file_abundance", C.c_int, [_vp, C.c_char_p, C.c_int, _f32p, C.POINTER(C.c_int)])
_sig("cms_sample_files_csr", C.c_int, [_vp, C.c_int, C.POINTER(C.c_char_p), C.c_int, C.POINTER(_vp),
                                       C.POINTER(_vp), C.POINTER(_vp), _vp])
_sig("cms_list_load", C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(_vp)])
_sig("cms_list_n", C.c_int, [_vp])
_sig("cms_list_path", C.c_char_p, [_vp, C.c_int])
_sig("cms_list_name", C.c_char_p, [_vp, C.c_int])
_sig("cms_list_free", None, [_vp])
_sig("cms_synth_tree", C.c_int, [C.c_uint64, C.c_int, C.c_double, C.c_double, _i32p, _f32p, _i32p, _f32p, _i32p,
                                 C.POINTER(C.c_int)])
_sig("cms_synth_samples", C.c_int, [C.POINTER(SynthOpts), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)])
_sig("cms_synth_copy_numbers", C.c_int, [C.c_uint64, C.c_int, C.c_int, _f32p])
_sig("cms_counts_to_abundance_csr", C.c_int, [C.c_int, _i64p, _i32p, _f32p, _vp, _f32p])
_sig("cms_write_db", C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, _i32p, _f32p, _i32p, _f32p, _i32p, _vp])
_sig("cms_write_table", C.c_int, [C.c_char_p, C.c_int, C.c_int, _i64p, _i32p, _f32p])
_sig("cms_write_table_csr", C.c_int, [C.c_char_p, C.c_int, C.c_int, _i64p, _i32p, _f32p])
_sig("cms_free", None, [_vp])
_sig("cms_version", C.c_char_p, [])
_sig("cms_newick_parse", C.c_int, [C.c_char_p, C.c_int64, C.POINTER(_vp)])
_sig("cms_newick_load", C.c_int, [C.c_char_p, C.POINTER(_vp)])
_sig("cms_newick_free", None, [_vp])
_sig("cms_newick_leaf_n", C.c_int, [_vp])
_sig("cms_newick_order_n", C.c_int, [_vp])
_sig("cms_newick_leaf_name", C.c_char_p, [_vp, C.c_int])
_sig("cms_newick_order", C.c_int, [_vp] + [C.POINTER(_vp)] * 5)
_sig("cms_newick_info", C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)])
_sig("cms_newick_write_db", C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_char_p])
_sig("cms_server_opts_default", None, [C.POINTER(ServerOpts)])
_sig("cms_server_create", C.c_int, [_vp, C.POINTER(ServerOpts), C.POINTER(_vp)])
_sig("cms_server_submit", C.c_int, [_vp, C.c_int, _i64p, _i32p, _f32p, _vp, C.POINTER(C.c_int64)])
_sig("cms_server_wait", C.c_int, [_vp, C.c_int64, C.c_double, C.POINTER(RequestInfo)])
_sig("cms_server_drain", C.c_int, [_vp, C.c_double])
_sig("cms_server_get_stats", C.c_int, [_vp, C.POINTER(ServerStats)])
_sig("cms_server_destroy", None, [_vp])

#: every symbol declared in include/cms.h (checked by the CPU test suite)
EXPORTED = [
    "cms_opts_default", "cms_create", "cms_destroy", "cms_last_error", "cms_set_opts", "cms_set_tree",
    "cms_set_samples", "cms_set_samples_csr", "cms_all_pairs", "cms_all_pairs_device", "cms_copy_result",
    "cms_synchronize", "cms_device_count", "cms_calibrate_copy", "cms_all_pairs_stream",
    "cms_sample_order", "cms_sim_pairs", "cms_sim_pairs_noalloc",
    "cms_tile", "cms_get_stats", "cms_prepare", "cms_validate_tree", "cms_db_open", "cms_db_close", "cms_db_leaf_n", "cms_db_order_n",
    "cms_db_has_tree", "cms_db_has_cp", "cms_db_order", "cms_db_leaf_name", "cms_table_load",
    "cms_table_from_dense", "cms_table_free", "cms_table_n", "cms_table_f", "cms_table_sample_name",
    "cms_table_feature_name", "cms_table_skipped", "cms_table_save_csr", "cms_table_is_sparse",
    "cms_table_from_csr", "cms_table_abundance", "cms_table_abundance_csr", "cms_write_table_csr",
    "cms_write_dist", "cms_sample_file_abundance", "cms_sample_files_csr", "cms_list_load", "cms_list_n",
    "cms_list_path", "cms_list_name", "cms_list_free",
    "cms_synth_tree", "cms_synth_samples", "cms_synth_copy_numbers", "cms_counts_to_abundance_csr",
    "cms_write_db", "cms_write_table", "cms_free", "cms_version", "cms_newick_parse", "cms_newick_load",
    "cms_newick_free", "cms_newick_leaf_n", "cms_newick_order_n", "cms_newick_leaf_name", "cms_newick_order",
    "cms_newick_info", "cms_newick_write_db", "cms_server_opts_default", "cms_server_create",
    "cms_server_submit", "cms_server_wait", "cms_server_drain", "cms_server_get_stats", "cms_server_destroy",
]


def last_error(ctx=None) -> str:
    m = lib.cms_last_error(ctx)
    return m.decode(errors="replace") if m else ""


def _check(rc: int, ctx=None):
    if rc != OK:
        raise CmsError(rc, last_error(ctx))


# ------------------------------------------------------------------ tree / db

@dataclass
class Tree:
    """Schedule arrays exactly as in tree/order.txt (comp.h:161-192)."""
    leaf_n: int
    order_1: np.ndarray
    dist_1: np.ndarray
    order_2: np.ndarray
    dist_2: np.ndarray
    order_d: np.ndarray
    leaf_names: list | None = None
    max_slots: int = 0

    @property
    def order_n(self) -> int:
        return int(self.order_1.shape[0])

    def arrays(self):
        return (self.order_1, self.dist_1, self.order_2, self.dist_2, self.order_d)


def synth_tree(seed: int, leaf_n: int, bl_mean: float, bl_max: float = 0.99) -> Tree:
    """Seeded Yule tree in Sethi-Ullman order.txt form (cms_synth_tree)."""
    on = leaf_n - 1
    o1 = np.zeros(on, np.int32)
    o2 = np.zeros(on, np.int32)
    od = np.zeros(on, np.int32)
    d1 = np.zeros(on, np.float32)
    d2 = np.zeros(on, np.float32)
    slots = C.c_int(0)
    _check(lib.cms_synth_tree(seed, leaf_n, bl_mean, bl_max, o1, d1, o2, d2, od, C.byref(slots)))
    return Tree(leaf_n, o1, d1, o2, d2, od, [f"T{i}" for i in range(leaf_n)], slots.value)


def validate_tree(tree: Tree):
    """(nslots, is_tree, nonneg_beta) or CmsError(E_SCHEDULE) (cms_validate_tree)."""
    ns, it, nb = C.c_int(0), C.c_int(0), C.c_int(0)
    _check(lib.cms_validate_tree(tree.leaf_n, tree.order_n, *[np.ascontiguousarray(a) for a in tree.arrays()],
                                 C.byref(ns), C.byref(it), C.byref(nb)))
    return ns.value, bool(it.value), bool(nb.value)


class Database:
    """A ParallelMETA database directory (db.config + tree + copy numbers)."""

    def __init__(self, root: str | None = None, letter: str = "G"):
        h = _vp()
        _check(lib.cms_db_open(root.encode() if root else None, letter.encode()[:1], C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib.cms_db_close(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def leaf_n(self) -> int:
        return lib.cms_db_leaf_n(self._h)

    @property
    def order_n(self) -> int:
        return lib.cms_db_order_n(self._h)

    @property
    def has_tree(self) -> bool:
        return bool(lib.cms_db_has_tree(self._h))

    @property
    def has_cp(self) -> bool:
        return bool(lib.cms_db_has_cp(self._h))

    def leaf_name(self, i: int) -> str:
        return lib.cms_db_leaf_name(self._h, i).decode()

    def tree(self) -> Tree:
        ptrs = [_vp() for _ in range(5)]
        _check(lib.cms_db_order(self._h, *[C.byref(p) for p in ptrs]))
        n = self.order_n

        def arr(p, dt):
            if n == 0:
                return np.zeros(0, dt)
            ct = C.c_int32 if dt == np.int32 else C.c_float
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,)).astype(dt, copy=True)

        return Tree(self.leaf_n, arr(ptrs[0], np.int32), arr(ptrs[1], np.float32), arr(ptrs[2], np.int32),
                    arr(ptrs[3], np.float32), arr(ptrs[4], np.int32))


class Newick:
    """A Newick tree as a reference schedule (cms_newick_*): leaves in order of
    appearance, binarised, Sethi-Ullman order with lowest-free registers."""

    def __init__(self, path: str | None = None, text: str | None = None):
        h = _vp()
        if (path is None) == (text is None):
            raise ValueError("give exactly one of path or text")
        if path is not None:
            _check(lib.cms_newick_load(path.encode(), C.byref(h)))
        else:
            b = text.encode()
            _check(lib.cms_newick_parse(b, len(b), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib.cms_newick_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def leaf_n(self) -> int:
        return lib.cms_newick_leaf_n(self._h)

    @property
    def order_n(self) -> int:
        return lib.cms_newick_order_n(self._h)

    @property
    def names(self) -> list:
        return [lib.cms_newick_leaf_name(self._h, i).decode() for i in range(self.leaf_n)]

    def info(self) -> dict:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib.cms_newick_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"nslots": a.value, "multifurcations": b.value, "unary_merged": c.value}

    def tree(self) -> Tree:
        ptrs = [_vp() for _ in range(5)]
        _check(lib.cms_newick_order(self._h, *[C.byref(p) for p in ptrs]))
        n = self.order_n

        def arr(p, dt):
            if n == 0:
                return np.zeros(0, dt)
            ct = C.c_int32 if dt == np.int32 else C.c_float
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,)).astype(dt, copy=True)

        return Tree(self.leaf_n, arr(ptrs[0], np.int32), arr(ptrs[1], np.float32), arr(ptrs[2], np.int32),
                    arr(ptrs[3], np.float32), arr(ptrs[4], np.int32))

    def write_db(self, root: str, name: str, cp_path: str | None = None):
        _check(lib.cms_newick_write_db(self._h, root.encode(), name.encode(),
                                       cp_path.encode() if cp_path else None))


class Table:
    """OTU count table, rows = samples (table_format.h:110-165)."""

    def __init__(self, path: str | None = None, *, counts=None, samples=None, features=None, csr=None):
        """From a file (text table or sparse table file), from dense counts
        (n x f), or from csr=(rowptr, col, counts) over f = len(features)."""
        h = _vp()
        if path is not None:
            _check(lib.cms_table_load(path.encode(), C.byref(h)))
        elif csr is not None:
            rp, col, cnt = csr
            rp = np.ascontiguousarray(rp, np.int64)
            col = np.ascontiguousarray(col, np.int32)
            cnt = np.ascontiguousarray(cnt, np.float32)
            n, f = len(rp) - 1, len(features)
            sn = (C.c_char_p * max(n, 1))(*[s.encode() for s in samples])
            fn = (C.c_char_p * max(f, 1))(*[s.encode() for s in features])
            _check(lib.cms_table_from_csr(n, f, sn, fn, rp, col if col.size else np.zeros(1, np.int32),
                                          cnt if cnt.size else np.zeros(1, np.float32), C.byref(h)))
        else:
            counts = np.ascontiguousarray(counts, np.float32)
            n, f = counts.shape
            sn = (C.c_char_p * n)(*[s.encode() for s in samples])
            fn = (C.c_char_p * f)(*[s.encode() for s in features])
            _check(lib.cms_table_from_dense(n, f, sn, fn, counts, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib.cms_table_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def n(self) -> int:
        return lib.cms_table_n(self._h)

    @property
    def f(self) -> int:
        return lib.cms_table_f(self._h)

    @property
    def skipped(self) -> int:
        return lib.cms_table_skipped(self._h)

    @property
    def is_sparse(self) -> bool:
        return bool(lib.cms_table_is_sparse(self._h))

    def save_csr(self, path: str):
        """Write the table as a sparse table file (cms_table_save_csr)."""
        _check(lib.cms_table_save_csr(self._h, path.encode()))

    def sample_names(self) -> list:
        return [lib.cms_table_sample_name(self._h, i).decode() for i in range(self.n)]

    def feature_names(self) -> list:
        return [lib.cms_table_feature_name(self._h, j).decode() for j in range(self.f)]

    def abundance(self, db: Database, cp_correct: bool = True):
        """Dense n x leaf_n percent-scale abundances and mapped counts (comp.h:243-279)."""
        out = np.zeros((self.n, db.leaf_n), np.float32)
        mapped = np.zeros(self.n, np.int32)
        _check(lib.cms_table_abundance(db._h, self._h, int(cp_correct), out.reshape(-1),
                                       mapped.ctypes.data_as(_vp)))
        return out, mapped

    def abundance_csr(self, db: Database, cp_correct: bool = True) -> "Samples":
        """The same abundances as CSR samples (cms_table_abundance_csr)."""
        rp, col, val = _vp(), _vp(), _vp()
        _check(lib.cms_table_abundance_csr(db._h, self._h, int(cp_correct), C.byref(rp), C.byref(col), C.byref(val)))
        try:
            n = self.n
            rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int64)), shape=(n + 1,)).copy()
            nnz = int(rowptr[-1])
            cols = np.ctypeslib.as_array(C.cast(col, C.POINTER(C.c_int32)), shape=(max(nnz, 1),))[:nnz].copy()
            vals = np.ctypeslib.as_array(C.cast(val, C.POINTER(C.c_float)), shape=(max(nnz, 1),))[:nnz].copy()
        finally:
            for p in (rp, col, val):
                lib.cms_free(p)
        return Samples(rowptr, cols, vals)


def sample_file_abundance(db: Database, path: str, cp_correct: bool = True):
    """One classification file's leaf abundances and its number of OTUs in
    the tree (_Comp_Tree::Load_abd(const char *, float *, bool), comp.h:194-236)."""
    out = np.zeros(max(db.leaf_n, 1), np.float32)
    m = C.c_int(0)
    _check(lib.cms_sample_file_abundance(db._h, path.encode(), int(cp_correct), out, C.byref(m)))
    return out[: db.leaf_n], m.value


def sample_files_csr(db: Database, paths, cp_correct: bool = True):
    """Several files as CSR samples (zero abundances dropped) and the mapped
    OTU count of each (cms_sample_files_csr)."""
    n = len(paths)
    arr = (C.c_char_p * max(n, 1))(*[p.encode() for p in paths])
    mapped = np.zeros(max(n, 1), np.int32)
    rp, col, val = _vp(), _vp(), _vp()
    _check(lib.cms_sample_files_csr(db._h, n, arr, int(cp_correct), C.byref(rp), C.byref(col), C.byref(val),
                                    mapped.ctypes.data_as(_vp)))
    try:
        rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int64)), shape=(n + 1,)).copy()
        nnz = int(rowptr[-1])
        cols = np.ctypeslib.as_array(C.cast(col, C.POINTER(C.c_int32)), shape=(max(nnz, 1),))[:nnz].copy()
        vals = np.ctypeslib.as_array(C.cast(val, C.POINTER(C.c_float)), shape=(max(nnz, 1),))[:nnz].copy()
    finally:
        for p in (rp, col, val):
            lib.cms_free(p)
    return Samples(rowptr, cols, vals), mapped[:n]


class SampleList:
    """Sample list of `-l list [-p prefix]` (Load_List with ids and prefix,
    utility.h:344-391): .paths (files to load) and .names (sample ids)."""

    def __init__(self, path: str, prefix: str = ""):
        h = _vp()
        _check(lib.cms_list_load(path.encode(), prefix.encode(), C.byref(h)))
        try:
            n = lib.cms_list_n(h)
            self.paths = [lib.cms_list_path(h, i).decode() for i in range(n)]
            self.names = [lib.cms_list_name(h, i).decode() for i in range(n)]
        finally:
            lib.cms_list_free(h)

    def __len__(self):
        return len(self.paths)


def write_dist(path: str, names, sim_packed: np.ndarray, is_sim: bool = False):
    """Output_Matrix (cms_cuda.cu:1305-1342)."""
    n = len(names)
    arr = (C.c_char_p * max(n, 1))(*[s.encode() for s in names])
    sim = np.ascontiguousarray(sim_packed, np.float32)
    if sim.size == 0:
        sim = np.zeros(1, np.float32)
    _check(lib.cms_write_dist(path.encode(), n, arr, sim, int(is_sim)))


# ------------------------------------------------------------ synthetic data

@dataclass
class Samples:
    """CSR samples: row s has (leaf, value) for its nonzero leaves."""
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def n(self) -> int:
        return int(self.rowptr.shape[0] - 1)

    def dense(self, leaf_n: int) -> np.ndarray:
        out = np.zeros((self.n, leaf_n), np.float32)
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        out[rows, self.col] = self.val
        return out

    def subset(self, idx) -> "Samples":
        idx = np.asarray(idx, dtype=np.int64)
        lens = np.diff(self.rowptr)[idx]
        rp = np.zeros(len(idx) + 1, np.int64)
        rp[1:] = np.cumsum(lens)
        col = np.concatenate([self.col[self.rowptr[i]:self.rowptr[i + 1]] for i in idx]) if len(idx) else \
            np.zeros(0, np.int32)
        val = np.concatenate([self.val[self.rowptr[i]:self.rowptr[i + 1]] for i in idx]) if len(idx) else \
            np.zeros(0, np.float32)
        return Samples(rp, col.astype(np.int32), val.astype(np.float32))


def synth_counts(seed, n_samples, leaf_n, nnz_lo, nnz_hi, zipf_s, nb_mean, nb_dispersion,
                 log_uniform=False) -> Samples:
    o = SynthOpts(seed, n_samples, leaf_n, nnz_lo, nnz_hi, int(log_uniform), zipf_s, nb_mean, nb_dispersion)
    rp, col, cnt = _vp(), _vp(), _vp()
    _check(lib.cms_synth_samples(C.byref(o), C.byref(rp), C.byref(col), C.byref(cnt)))
    try:
        rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int64)), shape=(n_samples + 1,)).copy()
        nnz = int(rowptr[-1])
        if nnz:
            c = np.ctypeslib.as_array(C.cast(col, C.POINTER(C.c_int32)), shape=(nnz,)).copy()
            v = np.ctypeslib.as_array(C.cast(cnt, C.POINTER(C.c_float)), shape=(nnz,)).copy()
        else:
            c, v = np.zeros(0, np.int32), np.zeros(0, np.float32)
    finally:
        lib.cms_free(rp)
        lib.cms_free(col)
        lib.cms_free(cnt)
    return Samples(rowptr, c, v)


def synth_copy_numbers(seed: int, leaf_n: int, cp_max: int = 15) -> np.ndarray:
    cp = np.zeros(leaf_n, np.float32)
    _check(lib.cms_synth_copy_numbers(seed, leaf_n, cp_max, cp))
    return cp


def counts_to_abundance(s: Samples, cp: np.ndarray | None) -> Samples:
    """Load_abd arithmetic (comp.h:252-275) on CSR counts whose columns are leaves."""
    out = np.zeros_like(s.val)
    cpp = None
    if cp is not None:
        cp = np.ascontiguousarray(cp, np.float32)
        cpp = cp.ctypes.data_as(_vp)
    _check(lib.cms_counts_to_abundance_csr(s.n, s.rowptr, s.col, s.val if s.val.size else np.zeros(1, np.float32),
                                           cpp, out if out.size else np.zeros(1, np.float32)))
    return Samples(s.rowptr, s.col, out)


def write_db(root: str, name: str, tree: Tree, cp: np.ndarray | None):
    cpp = None
    if cp is not None:
        cp = np.ascontiguousarray(cp, np.float32)
        cpp = cp.ctypes.data_as(_vp)
    _check(lib.cms_write_db(root.encode(), name.encode(), tree.leaf_n, tree.order_n, *tree.arrays(), cpp))


def write_table_csr(path: str, counts: Samples, leaf_n: int):
    """The table of write_table as a sparse table file (cms_write_table_csr)."""
    _check(lib.cms_write_table_csr(path.encode(), counts.n, leaf_n, counts.rowptr, counts.col,
                                   counts.val if counts.val.size else np.zeros(1, np.float32)))


def write_table(path: str, counts: Samples, leaf_n: int):
    _check(lib.cms_write_table(path.encode(), counts.n, leaf_n, counts.rowptr, counts.col,
                               counts.val if counts.val.size else np.zeros(1, np.float32)))


#: The workloads of the benchmark plan (tree seed 1000+k, samples 2000+k,
#: copy numbers 3000+k for configuration k).
WORKLOADS = {
    0: dict(leaves=5000, bl_mean=0.03, samples=64, nnz=(100, 800), log_uniform=False, zipf=1.1, nb=(50, 0.5)),
    1: dict(leaves=99322, bl_mean=0.02, samples=10000, nnz=(200, 2000), log_uniform=False, zipf=1.1, nb=(40, 0.4)),
    2: dict(leaves=99322, bl_mean=0.02, samples=50000, nnz=(200, 2000), log_uniform=False, zipf=1.1, nb=(40, 0.4)),
    3: dict(leaves=99322, bl_mean=0.02, samples=100000, nnz=(50, 20000), log_uniform=True, zipf=1.0, nb=(30, 0.3)),
    4: dict(leaves=99322, bl_mean=0.02, samples=200000, nnz=(50, 20000), log_uniform=True, zipf=1.0, nb=(30, 0.3)),
}


def workload(k: int, samples: int | None = None, cp_correct: bool = True):
    """Tree, copy numbers, counts and abundances of workload k (optionally fewer samples)."""
    w = WORKLOADS[k]
    tree = synth_tree(1000 + k, w["leaves"], w["bl_mean"], 0.99)
    cp = synth_copy_numbers(3000 + k, w["leaves"], 15)
    n = w["samples"] if samples is None else samples
    counts = synth_counts(2000 + k, n, w["leaves"], w["nnz"][0], w["nnz"][1], w["zipf"], w["nb"][0], w["nb"][1],
                          w["log_uniform"])
    abd = counts_to_abundance(counts, cp if cp_correct else None)
    return tree, cp, counts, abd


# ------------------------------------------------------------------ context

class Context:
    """One GPU context (cms_create). Not thread-safe."""

    def __init__(self, device: int = 0, kernel: int = KERNEL_WALK, storage: int = STORE_F32,
                 metric: int = METRIC_WEIGHTED, **opts):
        o = Opts()
        lib.cms_opts_default(C.byref(o))
        o.kernel, o.storage, o.metric = kernel, storage, metric
        for k, v in opts.items():  # bin, autotune, variant, use_graph
            setattr(o, k, int(v))
        h = _vp()
        _check(lib.cms_create(device, C.byref(o), C.byref(h)))
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            lib.cms_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def error(self) -> str:
        return last_error(self._h)

    def set_opts(self, kernel: int = KERNEL_WALK, storage: int = STORE_F32, metric: int = METRIC_WEIGHTED, **opts):
        o = Opts()
        lib.cms_opts_default(C.byref(o))
        o.kernel, o.storage, o.metric = kernel, storage, metric
        for k, v in opts.items():
            setattr(o, k, int(v))
        _check(lib.cms_set_opts(self._h, C.byref(o)), self._h)

    def set_tree(self, tree: Tree):
        _check(lib.cms_set_tree(self._h, tree.leaf_n, tree.order_n, *[np.ascontiguousarray(a) for a in tree.arrays()]),
               self._h)

    def set_samples(self, abd: np.ndarray):
        abd = np.ascontiguousarray(abd, np.float32)
        _check(lib.cms_set_samples(self._h, abd.shape[0], abd.reshape(-1) if abd.size else np.zeros(1, np.float32)),
               self._h)
        self.n = abd.shape[0]

    def set_samples_csr(self, s: Samples):
        val = s.val if s.val.size else np.zeros(1, np.float32)
        col = s.col if s.col.size else np.zeros(1, np.int32)
        _check(lib.cms_set_samples_csr(self._h, s.n, np.ascontiguousarray(s.rowptr, np.int64), col, val), self._h)
        self.n = s.n

    def all_pairs(self) -> np.ndarray:
        out = np.zeros(max(self.n * (self.n - 1) // 2, 1), np.float32)
        _check(lib.cms_all_pairs(self._h, out), self._h)
        return out[: self.n * (self.n - 1) // 2]

    def all_pairs_device(self, d_ptr: int | None = None, stream: int | None = None):
        """Enqueue all pairs into a device buffer (None: the context's own)."""
        _check(lib.cms_all_pairs_device(self._h, _vp(d_ptr) if d_ptr else None, _vp(stream) if stream else None),
               self._h)

    def all_pairs_stream(self, fn, band_pairs: int = 0):
        """Streamed all-pairs: fn(first, count, sims) per band, sims a float32
        view valid during the call, indices in the stored-position packed
        triangle (see sample_order)."""
        def cb(_user, first, count, ptr):
            fn(int(first), int(count), np.ctypeslib.as_array(ptr, shape=(int(count),)))
        cfn = _BAND_FN(cb)
        _check(lib.cms_all_pairs_stream(self._h, band_pairs, cfn, None), self._h)

    def sample_order(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1), np.int32)
        _check(lib.cms_sample_order(self._h, out), self._h)
        return out[: self.n]

    def copy_result(self, count: int | None = None) -> np.ndarray:
        """Read back the context-owned result buffer (after all_pairs_device)."""
        if count is None:
            count = self.n * (self.n - 1) // 2
        out = np.zeros(max(count, 1), np.float32)
        _check(lib.cms_copy_result(self._h, out, count), self._h)
        return out[:count]

    def synchronize(self):
        _check(lib.cms_synchronize(self._h), self._h)

    def calibrate_copy(self, nbytes: int, reps: int = 5) -> float:
        """Average ms of a plain device copy of nbytes (traffic-counter reference)."""
        ms = C.c_double(0)
        _check(lib.cms_calibrate_copy(self._h, nbytes, reps, C.byref(ms)), self._h)
        return ms.value

    def sim_pairs(self, row, col) -> np.ndarray:
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        out = np.zeros(max(len(row), 1), np.float32)
        _check(lib.cms_sim_pairs(self._h, len(row), row if len(row) else np.zeros(1, np.int32),
                                 col if len(col) else np.zeros(1, np.int32), out), self._h)
        return out[: len(row)]

    def sim_pairs_noalloc(self, row, col) -> np.ndarray:
        """cms_sim_pairs_noalloc: the pair list with no device allocation
        (layouts must be prepared; at most NOALLOC_MAX_PAIRS pairs)."""
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        out = np.zeros(max(len(row), 1), np.float32)
        _check(lib.cms_sim_pairs_noalloc(self._h, len(row), row if len(row) else np.zeros(1, np.int32),
                                         col if len(col) else np.zeros(1, np.int32), out), self._h)
        return out[: len(row)]

    def tile(self, r0: int, nr: int, c0: int, nc: int) -> np.ndarray:
        out = np.zeros(max(nr * nc, 1), np.float32)
        _check(lib.cms_tile(self._h, r0, nr, c0, nc, out), self._h)
        return out[: nr * nc].reshape(nr, nc)

    def prepare(self, force: bool = False):
        _check(lib.cms_prepare(self._h, int(force)), self._h)

    def stats(self) -> Stats:
        s = Stats()
        _check(lib.cms_get_stats(self._h, C.byref(s)), self._h)
        return s


# ------------------------------------------------------------------ serving

class Request:
    """A submitted query set: `out` (nq x n_ref) is filled when it finishes."""

    def __init__(self, server, rid: int, out: np.ndarray, keep):
        self.server, self.id, self.out, self._keep = server, rid, out, keep
        self.info = None

    def wait(self, timeout: float | None = None) -> np.ndarray:
        """Block until finished (timeout in seconds; None = forever); returns out."""
        if self.info is None:
            info = RequestInfo()
            rc = lib.cms_server_wait(self.server._h, self.id, -1.0 if timeout is None else float(timeout),
                                     C.byref(info))
            if rc == E_TIMEOUT:
                raise TimeoutError(last_error())
            self.info = info
            self._keep = None
            self.server._inflight.pop(self.id, None)
            _check(rc)
        return self.out


class Server:
    """Serving loop over a context's samples as resident references
    (cms_server_create). The context stays read-only while the server lives."""

    def __init__(self, ctx: "Context", slots: int = 1024, bands: int = 16, memory_budget: int = 0, variant: int = 0):
        o = ServerOpts()
        lib.cms_server_opts_default(C.byref(o))
        o.slots, o.bands, o.memory_budget, o.variant = slots, bands, memory_budget, variant
        h = _vp()
        _check(lib.cms_server_create(ctx._h, C.byref(o), C.byref(h)), ctx._h)
        self._h = h
        self.ctx = ctx
        self.n_ref = ctx.n
        self._inflight = {}  # id -> Request: output buffers the server still writes

    def submit(self, q: "Samples") -> Request:
        rowptr = np.ascontiguousarray(q.rowptr, np.int64)
        col = np.ascontiguousarray(q.col, np.int32)
        val = np.ascontiguousarray(q.val, np.float32)
        out = np.empty((q.n, self.n_ref), np.float32)
        rid = C.c_int64()
        _check(lib.cms_server_submit(self._h, q.n, rowptr, col if len(col) else np.zeros(1, np.int32),
                                     val if len(val) else np.zeros(1, np.float32),
                                     out.ctypes.data_as(C.c_void_p), C.byref(rid)))
        req = Request(self, rid.value, out, (rowptr, col, val))
        self._inflight[req.id] = req
        return req

    def drain(self, timeout: float | None = None):
        rc = lib.cms_server_drain(self._h, -1.0 if timeout is None else float(timeout))
        if rc == E_TIMEOUT:
            raise TimeoutError(last_error())
        _check(rc)

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(lib.cms_server_get_stats(self._h, C.byref(st)))
        return st

    def close(self):
        """Finish every submitted set (their outputs are filled), then free."""
        if getattr(self, "_h", None):
            lib.cms_server_destroy(self._h)
            self._h = None
            self._inflight.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def packed_index(m, n, N):
    """Reference packed upper-triangle index p = m*N + n - (m+1)(m+2)/2, m < n."""
    m = np.asarray(m, np.int64)
    n = np.asarray(n, np.int64)
    return m * N + n - (m + 1) * (m + 2) // 2


def device_count() -> int:
    """HIP devices visible to libcms (0 when none)."""
    n = lib.cms_device_count()
    return max(n, 0)


def version() -> str:
    return lib.cms_version().decode()
