"""CPU oracle for the GTS hot path -- TEST INFRASTRUCTURE ONLY.

A ctypes front end over ``liboracle.so`` (built from ``gts_oracle.c`` by the
Makefile next to it).  Only the tests, the smoke check and the CPU-baseline leg
of ``bench.py`` import this package; the GPU library never does.

Each function states the reference routine it restates; the C file carries the
file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force=False):
    """Compile liboracle.so with the Makefile (gcc -fopenmp)."""
    if force and os.path.exists(_LIB):
        os.remove(_LIB)
    subprocess.check_call(["make", "-s", "-C", _HERE])


class _Data(C.Structure):
    _fields_ = [("n", C.c_int64), ("dim", C.c_int), ("metric", C.c_int),
                ("vec", C.c_void_p), ("str", C.c_void_p), ("len", C.c_void_p),
                ("stride", C.c_int)]


class _Tree(C.Structure):
    _fields_ = [("order", C.c_int), ("max_size", C.c_int), ("max_h", C.c_int),
                ("height", C.c_int), ("n", C.c_int64), ("num_nodes", C.c_int64),
                ("level_start", C.POINTER(C.c_int64)), ("perm", C.POINTER(C.c_int32)),
                ("lid", C.POINTER(C.c_int32)), ("size", C.POINTER(C.c_int32)),
                ("pivot", C.POINTER(C.c_int32)), ("ppivot", C.POINTER(C.c_int32)),
                ("min_dis", C.POINTER(C.c_double)), ("exists", C.POINTER(C.c_int8)),
                ("is_leaf", C.POINTER(C.c_int8))]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.c_void_p
        L.orc_edit.restype = C.c_int
        L.orc_edit.argtypes = [P, C.c_int, P, C.c_int]
        L.orc_dist_vec.restype = C.c_double
        L.orc_dist_vec.argtypes = [P, P, C.c_int, C.c_int]
        L.orc_dist.restype = C.c_double
        L.orc_dist.argtypes = [P, C.c_int64, C.c_int64]
        L.orc_bf_range_count.argtypes = [P, P, C.c_int, C.c_double, P, C.c_int]
        L.orc_bf_range_ids.argtypes = [P, P, C.c_int, C.c_double, P, P, P, C.c_int]
        L.orc_bf_knn.argtypes = [P, P, C.c_int, C.c_int, P, P, C.c_int]
        L.orc_tree_build.restype = C.c_int
        L.orc_tree_build.argtypes = [P, C.c_int, C.c_int, C.c_int, C.POINTER(_Tree), C.c_int]
        L.orc_tree_free.argtypes = [C.POINTER(_Tree)]
        L.orc_tree_range.argtypes = [P, C.POINTER(_Tree), P, C.c_int, C.c_double, P, P, P, C.c_int]
        L.orc_tree_knn.argtypes = [P, C.POINTER(_Tree), P, C.c_int, P, C.c_int, P, P, C.c_int,
                                   P, C.c_int]
        L.orc_update.restype = C.c_int64
        L.orc_update.argtypes = [P, P, C.c_int64, P, P, C.c_int64, C.c_double, C.c_int, P, C.c_int]
        L.orc_update2.restype = C.c_int64
        L.orc_update2.argtypes = [P, P, C.c_int64, P, P, C.c_int64, C.c_double, C.c_int, P, P, C.c_int]
        L.orc_update_stream.restype = C.c_int64
        L.orc_update_stream.argtypes = [C.c_int64, C.c_int64, P, P, C.c_int64, C.c_int, P, P]
        L.orc_update3.restype = C.c_int64
        L.orc_update3.argtypes = [P, P, C.c_int64, P, P, C.c_int64, C.c_double, C.c_int, C.c_int64, P, P, C.c_int]
        L.orc_update_stream2.restype = C.c_int64
        L.orc_update_stream2.argtypes = [C.c_int64, C.c_int64, P, P, C.c_int64, C.c_int, C.c_int64, P, P]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_mlp_loss_grad.argtypes = [P, P, C.c_int64, C.c_int, C.c_int] + [P] * 13
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleData:
    """Keeps the numpy buffers alive behind the C view."""

    def __init__(self, ds):
        self.metric = ds.metric
        if ds.strings is not None:
            self.mat, self.lens = ds.string_matrix()
            self.vec = None
            self.c = _Data(len(ds.strings), int(self.mat.shape[1]), ds.metric, None,
                           _ptr(self.mat), _ptr(self.lens), int(self.mat.shape[1]))
        else:
            self.vec = np.ascontiguousarray(ds.vectors, dtype=np.float32)
            self.mat = self.lens = None
            self.c = _Data(self.vec.shape[0], self.vec.shape[1], ds.metric, _ptr(self.vec),
                           None, None, 0)

    @property
    def ref(self):
        return C.byref(self.c)


def edit(a: bytes, b: bytes) -> int:
    """Levenshtein distance (tree.cuh:128-154)."""
    A = np.frombuffer(a, dtype=np.uint8) if a else np.zeros(1, np.uint8)
    B = np.frombuffer(b, dtype=np.uint8) if b else np.zeros(1, np.uint8)
    return lib().orc_edit(_ptr(A), len(a), _ptr(B), len(b))


def dist(od: OracleData, i: int, j: int) -> float:
    return lib().orc_dist(od.ref, int(i), int(j))


def bf_range(od, qids, r, with_ids=False, threads=0):
    """Brute-force range: counts (and CSR ids/dists ascending by ID)."""
    q = np.ascontiguousarray(qids, dtype=np.int32)
    counts = np.zeros(len(q), dtype=np.int64)
    lib().orc_bf_range_count(od.ref, _ptr(q), len(q), float(r), _ptr(counts), threads)
    if not with_ids:
        return counts
    offs = np.zeros(len(q) + 1, dtype=np.int64)
    offs[1:] = np.cumsum(counts)
    ids = np.zeros(max(1, offs[-1]), dtype=np.int32)
    dists = np.zeros(max(1, offs[-1]), dtype=np.float64)
    lib().orc_bf_range_ids(od.ref, _ptr(q), len(q), float(r), _ptr(offs), _ptr(ids),
                           _ptr(dists), threads)
    return counts, offs, ids[:offs[-1]], dists[:offs[-1]]


def bf_knn(od, qids, k, threads=0):
    """Brute-force kNN: ids, dists [nq, k] ascending by (distance, ID)."""
    q = np.ascontiguousarray(qids, dtype=np.int32)
    ids = np.zeros((len(q), k), dtype=np.int32)
    dists = np.zeros((len(q), k), dtype=np.float64)
    lib().orc_bf_knn(od.ref, _ptr(q), len(q), int(k), _ptr(ids), _ptr(dists), threads)
    return ids, dists


class OracleTree:
    """indexConstru restated (tree.cuh:332-393); arrays exposed as numpy views."""

    def __init__(self, od, order=10, max_size=20, max_h=0, threads=0):
        self.od = od
        self.t = _Tree()
        rc = lib().orc_tree_build(od.ref, order, max_size, max_h, C.byref(self.t), threads)
        if rc != 0:
            raise ValueError("bad tree parameters")
        t = self.t
        nn = t.num_nodes

        def arr(p, n):
            return np.ctypeslib.as_array(p, shape=(n,)).copy()
        self.height = t.height
        self.order = t.order
        self.level_start = arr(t.level_start, t.height + 1)
        self.perm = arr(t.perm, t.n)
        self.lid = arr(t.lid, nn)
        self.size = arr(t.size, nn)
        self.pivot = arr(t.pivot, nn)
        self.ppivot = arr(t.ppivot, nn)
        self.min_dis = arr(t.min_dis, nn)
        self.exists = arr(t.exists, nn)
        self.is_leaf = arr(t.is_leaf, nn)

    def range(self, qids, r, threads=0):
        """Two-stage range search (searchIndexRnnV2): counts, verified, tested."""
        q = np.ascontiguousarray(qids, dtype=np.int32)
        c = np.zeros(len(q), np.int64)
        v = np.zeros(len(q), np.int64)
        te = np.zeros(len(q), np.int64)
        lib().orc_tree_range(self.od.ref, C.byref(self.t), _ptr(q), len(q), float(r), _ptr(c),
                             _ptr(v), _ptr(te), threads)
        return c, v, te

    def knn(self, qids, k, threads=0):
        """Two-stage kNN (searchIndexKnnV2); ``k`` an int or one k per query.
        Returns ids, dists [nq, kmax] (padded with -1 / inf) and verified counts."""
        q = np.ascontiguousarray(qids, dtype=np.int32)
        if np.isscalar(k):
            ks = None
            kmax = int(k)
        else:
            ks = np.ascontiguousarray(k, dtype=np.int32)
            kmax = int(ks.max())
        ids = np.full((len(q), kmax), -1, dtype=np.int32)
        dists = np.full((len(q), kmax), np.inf, dtype=np.float64)
        v = np.zeros(len(q), np.int64)
        lib().orc_tree_knn(self.od.ref, C.byref(self.t), _ptr(q), len(q), _ptr(ks),
                           kmax if ks is None else 0, _ptr(ids), _ptr(dists), kmax, _ptr(v), threads)
        return ids, dists, v

    def __del__(self):
        try:
            lib().orc_tree_free(C.byref(self.t))
        except Exception:
            pass


def update_counts(od, flags, oids, r, cache_cap=10, fresh=None, threads=0, evaluate=None, period=0):
    """updateIndexRnn restated: one count per query op (update.cuh:449-638).
    ``evaluate``: optional per-op mask; query ops outside it get count -1.
    ``period`` > 0 also rebuilds after every period-th op."""
    fl = np.ascontiguousarray(flags, dtype=np.int32)
    ids = np.ascontiguousarray(oids, dtype=np.int32)
    nq = int((fl == 2).sum())
    out = np.zeros(max(1, nq), dtype=np.int64)
    fr = None if fresh is None else np.ascontiguousarray(fresh, dtype=np.float32)
    ev = None if evaluate is None else np.ascontiguousarray(evaluate, dtype=np.int8)
    got = lib().orc_update3(od.ref, _ptr(fr), 0 if fr is None else fr.shape[0], _ptr(fl),
                            _ptr(ids), len(fl), float(r), int(cache_cap), int(period), _ptr(ev), _ptr(out),
                            threads)
    if got < 0:
        raise ValueError("invalid update op at index %d" % (-got - 1))
    return out[:got]


def update_stream(n0, nfresh, kinds, u, cache_cap, period=0):
    """Concrete (flag, id) ops from drawn kinds (0 insert, 1 delete, 2 query)
    and uniforms: fresh-row inserts (flag 3), uniform live logical deletes,
    queries of uniform live tree rows (orc_update_stream2), under the rebuild
    rule of update_counts with the same cache_cap and period."""
    k = np.ascontiguousarray(kinds, dtype=np.int8)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    fl = np.zeros(len(k), np.int32)
    ids = np.zeros(len(k), np.int32)
    lib().orc_update_stream2(int(n0), int(nfresh), _ptr(k), _ptr(uu), len(k), int(cache_cap), int(period),
                             _ptr(fl), _ptr(ids))
    keep = fl >= 0
    return np.stack([fl[keep], ids[keep]], axis=1)


def mlp_loss_grad(X, t, W1, b1, W2, b2, W3, b3):
    """FP64 forward/backward of the 16-64-64-1 GELU radius network."""
    X = np.ascontiguousarray(X, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    ws = [np.ascontiguousarray(w, np.float32) for w in (W1, b1, W2, b2, W3, b3)]
    din, dh = X.shape[1], ws[0].shape[0]
    g = [np.zeros(w.shape, np.float64) for w in ws]
    y = np.zeros(X.shape[0], np.float64)
    loss = lib().orc_mlp_loss_grad(_ptr(X), _ptr(t), X.shape[0], din, dh,
                                   *[_ptr(w) for w in ws], *[_ptr(x) for x in g], _ptr(y))
    return loss, g, y
