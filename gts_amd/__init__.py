"""GTS pivot-tree index for MI355X (gfx950): Python front end over libgts.so.

The native library (``gts_amd/libgts.so``, built by ``python -m gts_amd.build``)
does all the work on the GPU; this module only marshals arguments through the C
ABI declared in ``include/gts.h``.  There is no CPU fallback: if the library is
missing, every call raises.

Entry points mirror the reference (Source Code/GTS/):

=====================  ===========================================================
``load_dataset``       ``load`` (include/file.cuh:28)
``load_queries``       ``loadQuery`` (include/file.cuh:112)
``load_updates``       ``loadUpdate`` (include/update.cuh:43)
``GTSIndex``           ``indexConstru`` (include/tree.cuh:332)
``.range_query``       ``searchIndexRnnV2`` (include/search_v2.cuh:780)
``.knn_query``         ``searchIndexKnnV2`` (include/search_v2.cuh:975)
``.update``            ``updateIndexRnn`` (include/update.cuh:449)
=====================  ===========================================================
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgts.so")

GTS_OK, GTS_E_IO, GTS_E_FORMAT, GTS_E_ARG, GTS_E_OOM, GTS_E_DEVICE, GTS_E_CAPACITY = range(7)
VERIFY_FP32, VERIFY_BF16X3 = 0, 1
MAX_K = 4096                  # GTS_MAX_K
MLP_NOALLOC_MAX_NB = 2048     # GTS_MLP_NOALLOC_MAX_NB
PLAN_AUTO, PLAN_PRUNED, PLAN_EXHAUSTIVE, PLAN_CELLS, PLAN_PERSISTENT = 0, 1, 2, 3, 4
RADIUS_NETWORK, RADIUS_PIVOTS = 0, 1


class GTSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("GTS error %d: %s" % (code, msg))
        self.code = code


class _BuildParams(C.Structure):
    _fields_ = [("order", C.c_int), ("max_size", C.c_int), ("max_h", C.c_int),
                ("device", C.c_int)]


class _UpdateOp(C.Structure):
    _fields_ = [("flag", C.c_int32), ("id", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("frontier_pairs", C.c_int64), ("leaf_records", C.c_int64),
                ("verified_pairs", C.c_int64), ("candidates", C.c_int64),
                ("band_pairs", C.c_int64), ("retries", C.c_int64), ("ms_total", C.c_double),
                ("ms_verify", C.c_double), ("verify_launches", C.c_int64),
                ("exhaustive_passes", C.c_int64), ("dense_rows", C.c_int64),
                ("dense_persistent", C.c_int64), ("graph_replays", C.c_int64),
                ("cell_scan", C.c_int64), ("live_chunks", C.c_int64),
                ("ms_update_search", C.c_double), ("ms_update_apply", C.c_double),
                ("update_queries", C.c_int64), ("update_changes", C.c_int64),
                ("ms_cells_build", C.c_double), ("budget_batches", C.c_int64),
                ("ms_verify_batch", C.c_double), ("verify_launches_batch", C.c_int64),
                ("live_chunks_batch", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def lib():
    """The loaded native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GTSError(GTS_E_IO, "libgts.so not built (run: python -m gts_amd.build)")
        L = C.CDLL(LIB_PATH)
        P, I64, I32 = C.c_void_p, C.c_int64, C.c_int
        L.gts_last_error.restype = C.c_char_p
        L.gts_version.restype = C.c_int
        sig = {
            "gts_dataset_load": [C.c_char_p, C.POINTER(P)],
            "gts_dataset_from_vectors": [P, I64, I32, I32, C.POINTER(P)],
            "gts_dataset_from_strings": [P, P, I64, I32, C.POINTER(P)],
            "gts_dataset_info": [P, P, P, P],
            "gts_dataset_rows": [P, P, I64],
            "gts_load_queries": [C.c_char_p, P, I64, P],
            "gts_load_updates": [C.c_char_p, P, I64, P],
            "gts_build": [P, C.POINTER(_BuildParams), C.POINTER(P)],
            "gts_index_info": [P, P, P, P],
            "gts_index_export": [P, P, P, P, P, P, P],
            "gts_range_query": [P, P, I64, C.c_float, P, P, P, P, I64],
            "gts_knn_query": [P, P, I64, I32, P, P, P],
            "gts_knn_query_var": [P, P, I64, P, I32, P, P, P],
            "gts_knn_query_dev": [P, P, I64, P, I32, I32, P, P, P, I32, I32],
            "gts_update": [P, P, I64, C.c_float, I32, P, I64, P, I64, P],
            "gts_learned_train": [P, P, I64, I32, I32, C.c_float, C.c_uint64, P],
            "gts_last_stats": [P, C.POINTER(Stats)],
            "gts_mlp_loss_grad": [P, P, P, I64, P, P, P],
            "gts_mlp_loss_grad_noalloc": [P, P, P, I64, P, P, P],
            "gts_save_knn": [C.c_char_p, C.c_char_p, P, P, I32, I64],
            "gts_save_range": [C.c_char_p, C.c_char_p, P, P, P, P, I64],
            "gts_knn_query_ex": [P, P, I64, P, I32, I32, P, P, P, I32, I32],
            "gts_device_alloc": [P, I64, C.POINTER(P)],
            "gts_device_free": [P, P],
            "gts_copy_to_device": [P, P, P, I64],
            "gts_copy_to_host": [P, P, P, I64],
            "gts_copy_device": [P, P, P, I64],
            "gts_set_knn_plan": [P, I32],
            "gts_set_radius_source": [P, I32],
            "gts_set_memory_budget": [P, I64],
            "gts_set_rebuild_period": [P, I64],
            "gts_set_fp32_mfma": [P, I32],
            "gts_set_autotune": [P, I32],
            "gts_set_dense_schedule": [P, I32, I32],
            "gts_set_graph": [P, I32],
            "gts_set_timing": [P, I32],
            "gts_set_cell_count": [P, I32],
            "gts_set_radius_quantile": [P, C.c_double],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.gts_mlp_num_params.argtypes = []
        L.gts_mlp_num_params.restype = C.c_int
        L.gts_dataset_free.argtypes = [P]
        L.gts_dataset_free.restype = None
        L.gts_index_free.argtypes = [P]
        L.gts_index_free.restype = None
        _lib = L
    return _lib


def _check(rc):
    if rc != GTS_OK:
        raise GTSError(rc, lib().gts_last_error().decode(errors="replace"))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class NativeDataset:
    """A dataset held by the native library (host memory)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def from_python(cls, ds):
        h = C.c_void_p()
        if ds.strings is not None:
            mat, lens = ds.string_matrix()
            _check(lib().gts_dataset_from_strings(_p(mat), _p(lens), len(lens), mat.shape[1],
                                                  C.byref(h)))
        else:
            v = np.ascontiguousarray(ds.vectors, dtype=np.float32)
            _check(lib().gts_dataset_from_vectors(_p(v), v.shape[0], v.shape[1], ds.metric,
                                                  C.byref(h)))
        return cls(h)

    def info(self):
        n, dim, metric = C.c_int64(), C.c_int(), C.c_int()
        _check(lib().gts_dataset_info(self.h, C.byref(n), C.byref(dim), C.byref(metric)))
        return n.value, dim.value, metric.value

    def __del__(self):
        try:
            if self.h:
                lib().gts_dataset_free(self.h)
        except Exception:
            pass


def load_dataset(path):
    """``load`` (file.cuh:28): the reference text data format."""
    h = C.c_void_p()
    _check(lib().gts_dataset_load(path.encode(), C.byref(h)))
    return NativeDataset(h)


def load_queries(path):
    """``loadQuery`` (file.cuh:112): count line, then one ID per line."""
    n = C.c_int64()
    _check(lib().gts_load_queries(path.encode(), None, 0, C.byref(n)))
    ids = np.zeros(max(1, n.value), np.int32)
    _check(lib().gts_load_queries(path.encode(), _p(ids), len(ids), C.byref(n)))
    return ids[:n.value]


def load_updates(path):
    """``loadUpdate`` (update.cuh:43): count line, then ``flag id`` per line."""
    n = C.c_int64()
    _check(lib().gts_load_updates(path.encode(), None, 0, C.byref(n)))
    ops = np.zeros((max(1, n.value), 2), np.int32)
    _check(lib().gts_load_updates(path.encode(), _p(ops), len(ops), C.byref(n)))
    return ops[:n.value]


class DeviceArray:
    """Device memory owned through the C ABI (gts_device_alloc / gts_device_free)."""

    def __init__(self, index, shape, dtype):
        self.index = index
        self.shape = tuple(shape) if not np.isscalar(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape))
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        _check(lib().gts_device_alloc(index.h, self.nbytes, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, index, a):
        a = np.ascontiguousarray(a)
        d = cls(index, a.shape, a.dtype)
        _check(lib().gts_copy_to_device(index.h, d.ptr, _p(a), a.nbytes))
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().gts_copy_to_host(self.index.h, _p(out), self.ptr, self.nbytes))
        return out

    def copy_from(self, other):
        """Device-to-device copy of ``other`` (same byte size) into this array."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(lib().gts_copy_device(self.index.h, self.ptr, other.ptr, self.nbytes))

    def free(self):
        if self.ptr and getattr(self.index, "h", None):
            lib().gts_device_free(self.index.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GTSIndex:
    """The pivot tree on one GPU (``indexConstru``, tree.cuh:332)."""

    def __init__(self, data, order=10, max_size=20, max_h=0, device=0):
        nd = data if isinstance(data, NativeDataset) else NativeDataset.from_python(data)
        self._ds = nd
        p = _BuildParams(order, max_size, max_h, device)
        h = C.c_void_p()
        _check(lib().gts_build(nd.h, C.byref(p), C.byref(h)))
        self.h = h
        self.n, self.dim, self.metric = nd.info()
        self.order, self.max_size = order, max_size

    def close(self):
        if getattr(self, "h", None):
            lib().gts_index_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        h, nn, n = C.c_int(), C.c_int64(), C.c_int64()
        _check(lib().gts_index_info(self.h, C.byref(h), C.byref(nn), C.byref(n)))
        return {"height": h.value, "num_nodes": nn.value, "n": n.value}

    def export_tree(self):
        inf = self.info()
        nn, n = inf["num_nodes"], inf["n"]
        out = {k: np.zeros(nn, np.int32) for k in ("lid", "size", "pivot", "flags")}
        out["perm"] = np.zeros(n, np.int32)
        out["min_dis"] = np.zeros(nn, np.float32)
        _check(lib().gts_index_export(self.h, _p(out["perm"]), _p(out["lid"]), _p(out["size"]),
                                      _p(out["pivot"]), _p(out["min_dis"]), _p(out["flags"])))
        out.update(inf)
        return out

    def stats(self):
        s = Stats()
        _check(lib().gts_last_stats(self.h, C.byref(s)))
        return s.as_dict()

    def range_query(self, qids, r, with_ids=False):
        """``searchIndexRnnV2``: counts per query (and CSR IDs ascending, distances)."""
        q = np.ascontiguousarray(qids, np.int32)
        counts = np.zeros(len(q), np.int64)
        if not with_ids:
            _check(lib().gts_range_query(self.h, _p(q), len(q), float(r), _p(counts), None, None,
                                         None, 0))
            return counts
        cap = max(1024, 64 * len(q))
        while True:
            offs = np.zeros(len(q) + 1, np.int64)
            ids = np.zeros(cap, np.int32)
            dists = np.zeros(cap, np.float32)
            rc = lib().gts_range_query(self.h, _p(q), len(q), float(r), _p(counts), _p(offs),
                                       _p(ids), _p(dists), cap)
            if rc == GTS_E_CAPACITY:
                cap = int(offs[-1]) + 1
                continue
            _check(rc)
            return counts, offs, ids[:offs[-1]], dists[:offs[-1]]

    def knn_query(self, qids, k, learned=False, verify=VERIFY_FP32):
        """``searchIndexKnnV2``: ids, dists [nq, kmax] ascending by (distance, ID), kth[nq].
        ``k`` is an int or one k per query; ``learned`` uses the radius network and
        ``verify`` selects FP32 or BF16 hi+lo MFMA verification (L2)."""
        q = np.ascontiguousarray(qids, np.int32)
        if np.isscalar(k):
            ks = np.full(len(q), int(k), np.int32)
        else:
            ks = np.ascontiguousarray(k, np.int32)
        kmax = int(ks.max()) if len(ks) else 1
        ids = np.zeros((len(q), kmax), np.int32)
        dists = np.zeros((len(q), kmax), np.float32)
        kth = np.zeros(len(q), np.float32)
        if not learned and verify == VERIFY_FP32:
            _check(lib().gts_knn_query_var(self.h, _p(q), len(q), _p(ks), kmax, _p(ids),
                                           _p(dists), _p(kth)))
        else:
            _check(lib().gts_knn_query_ex(self.h, _p(q), len(q), _p(ks), 0, kmax, _p(ids),
                                          _p(dists), _p(kth), int(bool(learned)), int(verify)))
        return ids, dists, kth

    def device_array(self, shape, dtype):
        """A device buffer on the index's GPU (see DeviceArray)."""
        return DeviceArray(self, shape, dtype)

    def knn_query_dev(self, qids_dev, k, ids_dev, dists_dev, kth_dev, ks_dev=None,
                      learned=False, verify=VERIFY_FP32):
        """kNN on DeviceArray buffers resident on the index's GPU; returns when done."""
        kstride = ids_dev.shape[1]
        _check(lib().gts_knn_query_dev(
            self.h, qids_dev.ptr, qids_dev.size, None if ks_dev is None else ks_dev.ptr,
            int(k), kstride, ids_dev.ptr, dists_dev.ptr, kth_dev.ptr, int(bool(learned)),
            int(verify)))

    def update(self, ops, r, cache_cap=10, fresh=None, period=None):
        """``updateIndexRnn``: apply ops in order; one count per query op.
        ``period``: also rebuild after every period-th op (0 turns it off,
        None keeps the index's setting)."""
        if period is not None:
            _check(lib().gts_set_rebuild_period(self.h, int(period)))
        o = np.ascontiguousarray(np.asarray(ops, np.int32).reshape(-1, 2))
        nq = int((o[:, 0] == 2).sum())
        counts = np.zeros(max(1, nq), np.int64)
        got = C.c_int64()
        fr = None if fresh is None else np.ascontiguousarray(fresh, np.float32)
        _check(lib().gts_update(self.h, _p(o), len(o), float(r), int(cache_cap), _p(fr),
                                0 if fr is None else fr.shape[0], _p(counts), len(counts),
                                C.byref(got)))
        return counts[:got.value]

    def set_knn_plan(self, plan):
        """Plan of later learned BF16X3 kNN calls (PLAN_AUTO, PLAN_PRUNED, PLAN_EXHAUSTIVE, PLAN_CELLS)."""
        _check(lib().gts_set_knn_plan(self.h, int(plan)))

    def set_autotune(self, on):
        """Dense-kernel schedule autotuner on (default) or off."""
        _check(lib().gts_set_autotune(self.h, int(bool(on))))

    def set_dense_schedule(self, persistent, rows):
        """Fix the dense kernel's schedule (rows per range, persistent or grid); rows=0: automatic."""
        _check(lib().gts_set_dense_schedule(self.h, int(bool(persistent)), int(rows)))

    def set_radius_quantile(self, q):
        """Safety factor of the learned radius: cover quantile q of the training ratios."""
        _check(lib().gts_set_radius_quantile(self.h, float(q)))

    def set_fp32_mfma(self, on):
        """FP32-form L2 counts and kNN candidates on the FP32 matrix cores: True, False,
        or None for the default (on for rows of 32 dimensions or more)."""
        _check(lib().gts_set_fp32_mfma(self.h, -1 if on is None else int(bool(on))))

    def set_memory_budget(self, nbytes):
        """Memory budget of later query calls in bytes (0: half of the free device memory)."""
        _check(lib().gts_set_memory_budget(self.h, int(nbytes)))

    def set_radius_source(self, source):
        """Radius of later learned kNN calls: RADIUS_NETWORK (trained network)
        or RADIUS_PIVOTS (k-th distance to one tree level's pivots, no training)."""
        _check(lib().gts_set_radius_source(self.h, int(source)))

    def set_cell_count(self, ncell):
        """Cells of the cell scan (0: automatic); rebuilt at the next learned kNN call."""
        _check(lib().gts_set_cell_count(self.h, int(ncell)))

    def set_graph(self, on):
        """Graph replay of repeated learned kNN calls on (default) or off."""
        _check(lib().gts_set_graph(self.h, int(bool(on))))

    def set_timing(self, on):
        """HIP-event timing of kNN calls (ms_total, ms_verify) on (default) or off."""
        _check(lib().gts_set_timing(self.h, int(bool(on))))

    def learned_train(self, train_qids, kmax, epochs=30, lr=3e-3, seed=0):
        """Trains the radius network on exact kNN radii of ``train_qids``."""
        q = np.ascontiguousarray(train_qids, np.int32)
        loss = C.c_float()
        _check(lib().gts_learned_train(self.h, _p(q), len(q), int(kmax), int(epochs), float(lr),
                                       int(seed), C.byref(loss)))
        return loss.value

    def mlp_loss_grad_noalloc(self, X_dev, t_dev, w_dev, grad_dev, loss_dev):
        """The same on DeviceArray buffers (X [nb, 16], t [nb], w, grad, loss [1]),
        allocating nothing: a fixed per-GPU array of the library holds the
        activations (nb <= MLP_NOALLOC_MAX_NB)."""
        _check(lib().gts_mlp_loss_grad_noalloc(self.h, X_dev.ptr, t_dev.ptr, X_dev.shape[0], w_dev.ptr,
                                               grad_dev.ptr, loss_dev.ptr))

    def mlp_loss_grad(self, X, t, w):
        """Loss and gradients of the radius network on the GPU (gradient check)."""
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        w = np.ascontiguousarray(w, np.float32)
        g = np.zeros_like(w)
        loss = C.c_float()
        _check(lib().gts_mlp_loss_grad(self.h, _p(X), _p(t), X.shape[0], _p(w), _p(g),
                                       C.byref(loss)))
        return loss.value, g
