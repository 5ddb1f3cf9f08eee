"""GPU tests at the limits of the kNN path, against the FP64 brute-force oracle:

* more exact ties at the k-th distance than the top-k's per-query candidate
  capacity (GTS_MAX_K = 4096): 5,000 copies of one row (L2) and 5,000 copies
  of one string (edit distance) -- every pass overflows whatever the bound,
  and the exact fallback (all objects' FP64 or edit distances, stable radix
  sort, ties by ID) answers those queries;
* 2-d points in tight clusters away from the origin: the plain form
  converges with few reruns (no exact fallback) and equals brute force;
* device inputs of gts_knn_query_dev out of range (query IDs, k beyond the
  row stride): the call fails with GTS_E_ARG and the next valid call on the
  same index returns exact results;
* host-side limits: k above GTS_MAX_K, and a radius network asked for more
  neighbours than there are objects.
"""
import numpy as np
import pytest

import oracle
from gts_amd import (GTS_E_ARG, VERIFY_BF16X3, VERIFY_FP32, DeviceArray, GTSError, GTSIndex,
                     datagen)

pytestmark = pytest.mark.gpu

GTS_MAX_K = 4096


def _close(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    return np.all(np.abs(got - want) <= 1e-5 * np.maximum(np.abs(want), 1.0))


def _check(idx, od, qids, k, **kw):
    ids, dists, kth = idx.knn_query(qids, k, **kw)
    bids, bd = oracle.bf_knn(od, qids, k)
    assert (ids == bids).all(), "kNN ID sets differ"
    assert _close(dists, bd)
    assert _close(kth, bd[:, -1])


def test_ties_beyond_candidate_capacity_l2():
    rng = np.random.default_rng(11)
    x = rng.random((12000, 16)).astype(np.float32)
    x[2000:7000] = x[3]                       # 5,001 identical rows (3 and 2000..6999)
    ds = datagen.Dataset(2, vectors=x)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    qids = np.array([3, 2000, 6999, 4500, 0, 11999], np.int32)
    for k in (1, 16, 64):
        _check(idx, od, qids, k)
        # the tied queries overflowed on every pass until the exact fallback
        assert idx.stats()["retries"] >= 12, idx.stats()
    _check(idx, od, qids, 16, verify=VERIFY_BF16X3)


def test_ties_beyond_candidate_capacity_edit():
    rng = np.random.default_rng(12)
    n = 9000
    letters = np.frombuffer(b"abcdefgh", np.uint8)
    strs = []
    for i in range(n):
        if 1000 <= i < 6000:
            strs.append(b"abcdabcd")              # 5,000 identical strings
        else:
            ln = int(rng.integers(4, 12))
            strs.append(bytes(letters[rng.integers(0, 8, ln)]))
    ds = datagen.Dataset(6, strings=strs)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    qids = np.array([1000, 5999, 3000, 17, 8999], np.int32)
    for k in (1, 16, 64):
        _check(idx, od, qids, k)


def test_dense_low_dimensional_clusters_converge_without_the_fallback():
    """2-d points in tight clusters away from the origin (sigma 1e-3, k-th
    distances ~1e-4 at norms ~0.7), k = 64, the plain FP32 form: by default
    the direct-difference tiles verify them (their error band is relative;
    the norm expansion's absolute band held tens of thousands of rows per
    query and sent most queries through the exact fallback), results equal
    brute force and the reruns stay few: about one per query at k = 64 (the
    expansion's band needed eleven and 60 times the time)."""
    rng = np.random.default_rng(31)
    centres = 0.2 + 0.6 * rng.random((20, 2))
    x = (centres[rng.integers(0, 20, 200_000)] + 1e-3 * rng.standard_normal((200_000, 2))).astype(np.float32)
    ds = datagen.Dataset(2, vectors=x)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    qids = rng.integers(0, len(x), 2000).astype(np.int32)
    ids, dists, _ = idx.knn_query(qids, 64, verify=VERIFY_FP32)
    st = idx.stats()
    assert st["retries"] < 4 * len(qids), st
    sel = np.arange(0, len(qids), 20)
    bids, bd = oracle.bf_knn(od, qids[sel], 64)
    assert (ids[sel] == bids).all()
    assert _close(dists[sel], bd)


def test_device_inputs_out_of_range_fail_cleanly():
    rng = np.random.default_rng(5)
    ds = datagen.Dataset(2, vectors=rng.random((3000, 12)).astype(np.float32))
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    nq, kstride = 40, 8
    qids = rng.integers(0, 3000, nq).astype(np.int32)
    ids = idx.device_array((nq, kstride), np.int32)
    dists = idx.device_array((nq, kstride), np.float32)
    kth = idx.device_array(nq, np.float32)
    bad_q = qids.copy()
    bad_q[7] = 3000
    bad_q[9] = -5
    for q, ks in ((bad_q, None), (qids, np.full(nq, kstride + 1, np.int32)),
                  (qids, np.where(np.arange(nq) == 3, 0, 4).astype(np.int32))):
        dq = DeviceArray.from_host(idx, q)
        dk = None if ks is None else DeviceArray.from_host(idx, ks)
        with pytest.raises(GTSError) as e:
            idx.knn_query_dev(dq, 4, ids, dists, kth, ks_dev=dk, verify=VERIFY_FP32)
        assert e.value.code == GTS_E_ARG
    # a valid call afterwards: the error flag was cleared, results exact
    idx.knn_query_dev(DeviceArray.from_host(idx, qids), 4, ids, dists, kth, verify=VERIFY_FP32)
    bids, bd = oracle.bf_knn(od, qids, 4)
    got = ids.to_host()
    assert (got[:, :4] == bids).all() and (got[:, 4:] == -1).all()
    assert _close(dists.to_host()[:, :4], bd)


def test_k_and_kmax_limits():
    rng = np.random.default_rng(6)
    ds = datagen.Dataset(2, vectors=rng.random((5000, 8)).astype(np.float32))
    idx = GTSIndex(ds)
    with pytest.raises(GTSError) as e:
        idx.knn_query(np.array([1, 2], np.int32), GTS_MAX_K + 1)
    assert e.value.code == GTS_E_ARG
    small = GTSIndex(datagen.Dataset(2, vectors=rng.random((600, 8)).astype(np.float32)))
    with pytest.raises(GTSError) as e:
        small.learned_train(np.arange(100, dtype=np.int32), kmax=601, epochs=2)
    assert e.value.code == GTS_E_ARG
    # k = GTS_MAX_K itself is served
    ids, _, _ = idx.knn_query(np.array([4], np.int32), GTS_MAX_K)
    bids, _ = oracle.bf_knn(oracle.OracleData(ds), np.array([4], np.int32), GTS_MAX_K)
    assert (ids == bids).all()
