"""GPU tests of the persistent low-dimensional kNN pass (learned calls over L2,
L1 and Linf rows of at most 8 dimensions run as one launch, a wave per query
walking the tree depth-first): results equal the level-by-level search
(PLAN_PRUNED), the plain search and the FP64 brute-force oracle, for uneven k,
exact ties, k above n, radii that hold fewer than k objects (the walk without
the radius), device-side argument errors, and repeated batches that replay a
captured graph with per-query k in device memory (checked on every replay,
with other work between the replays).
"""
import numpy as np
import pytest

import oracle
from gts_amd import (GTS_E_ARG, PLAN_AUTO, PLAN_PERSISTENT, PLAN_PRUNED, RADIUS_PIVOTS, DeviceArray, GTSError,
                     GTSIndex, datagen)

pytestmark = pytest.mark.gpu


def _points(n, dim, seed, clusters=40):
    rng = np.random.default_rng(seed)
    cen = rng.random((clusters, dim))
    lab = rng.integers(0, clusters, n)
    x = cen[lab] + rng.normal(0, 0.02, (n, dim))
    return x.astype(np.float32)


def _check_rows(ids, dists, kth, bids, bd, ks):
    for i, k in enumerate(ks):
        k = int(k)
        kk = min(k, bids.shape[1])
        assert (ids[i, :kk] == bids[i, :kk]).all(), i
        assert (ids[i, kk:] == -1).all()
        assert np.allclose(dists[i, :kk], bd[i, :kk], rtol=1e-5, atol=1e-6)
        if k <= bids.shape[1]:
            assert abs(kth[i] - bd[i, k - 1]) <= 1e-5 * max(1.0, bd[i, k - 1])
        else:
            assert np.isinf(kth[i])


@pytest.mark.parametrize("metric,dim", [(2, 2), (1, 3), (0, 8), (2, 6)])
def test_persistent_equals_pruned_plain_and_oracle(metric, dim):
    x = _points(60000, dim, seed=metric * 10 + dim)
    ds = datagen.Dataset(metric, vectors=x)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(3)
    qids = rng.integers(0, len(x), 3000).astype(np.int32)
    ks = rng.integers(1, 65, len(qids)).astype(np.int32)
    idx.learned_train(rng.integers(0, len(x), 3000).astype(np.int32), kmax=64, epochs=40)
    got = {}
    for plan in (PLAN_PRUNED, PLAN_PERSISTENT, PLAN_AUTO):
        idx.set_knn_plan(plan)
        got[plan] = idx.knn_query(qids, ks, learned=True)
    plain = idx.knn_query(qids, ks)
    bids, bd = oracle.bf_knn(od, qids, 64)
    for plan in got:
        assert (got[plan][0] == plain[0]).all() and (got[plan][1] == plain[1]).all(), plan
        assert (got[plan][2] == plain[2]).all()
    _check_rows(*got[PLAN_PERSISTENT], bids, bd, ks)


def test_ties_small_n_and_short_radii():
    rng = np.random.default_rng(7)
    x = rng.random((5000, 2)).astype(np.float32)
    x[1000:1300] = x[5]                       # 301 identical points
    ds = datagen.Dataset(2, vectors=x)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    idx.learned_train(rng.integers(0, 5000, 2000).astype(np.int32), kmax=64, epochs=30)
    idx.set_knn_plan(PLAN_PERSISTENT)
    qids = np.array([5, 1000, 1299, 0, 4999, 2500], np.int32)
    ks = np.array([64, 1, 17, 64, 33, 2], np.int32)
    bids, bd = oracle.bf_knn(od, qids, 64)
    _check_rows(*idx.knn_query(qids, ks, learned=True), bids, bd, ks)
    # radii covering half the training ratios: many queries hold fewer than k
    # objects within theirs and walk again without it
    idx.set_radius_quantile(0.3)
    many = rng.integers(0, 5000, 2000).astype(np.int32)
    mk = rng.integers(1, 65, len(many)).astype(np.int32)
    bids, bd = oracle.bf_knn(od, many, 64)
    _check_rows(*idx.knn_query(many, mk, learned=True), bids, bd, mk)
    # the pivot radius source, no network
    idx.set_radius_source(RADIUS_PIVOTS)
    _check_rows(*idx.knn_query(many, mk, learned=True), bids, bd, mk)
    # fewer objects than k
    tiny = GTSIndex(datagen.Dataset(2, vectors=x[:40]))
    tiny.set_radius_source(RADIUS_PIVOTS)
    tiny.set_knn_plan(PLAN_PERSISTENT)
    tq = np.arange(40, dtype=np.int32)
    tb, tbd = oracle.bf_knn(oracle.OracleData(datagen.Dataset(2, vectors=x[:40])), tq, 40)
    _check_rows(*tiny.knn_query(tq, np.full(40, 64, np.int32), learned=True), tb, tbd, np.full(40, 64))


def test_graph_replays_with_device_ks():
    x = _points(200000, 2, seed=5, clusters=300)
    ds = datagen.Dataset(2, vectors=x)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(11)
    idx.learned_train(rng.integers(0, len(x), 4000).astype(np.int32), kmax=64, epochs=40)
    qids = rng.integers(0, len(x), 20000).astype(np.int32)
    ks = rng.integers(1, 65, len(qids)).astype(np.int32)
    d_q = DeviceArray.from_host(idx, qids)
    d_k = DeviceArray.from_host(idx, ks)
    ids = idx.device_array((len(qids), 64), np.int32)
    dists = idx.device_array((len(qids), 64), np.float32)
    kth = idx.device_array(len(qids), np.float32)
    sel = np.arange(0, len(qids), 97)
    bids, bd = oracle.bf_knn(od, qids[sel], 64)
    other = rng.integers(0, len(x), 500).astype(np.int32)
    ob, obd = oracle.bf_knn(od, other, 8)
    replays = 0
    for call in range(6):
        idx.knn_query_dev(d_q, 0, ids, dists, kth, ks_dev=d_k, learned=True)
        replays += idx.stats()["graph_replays"]
        hi, hd, hk = ids.to_host(), dists.to_host(), kth.to_host()
        _check_rows(hi[sel], hd[sel], hk[sel], bids, bd, ks[sel])
        # eager work on the same index between the replays
        oi, od8, _ = idx.knn_query(other, 8, learned=True)
        assert (oi == ob).all() and np.allclose(od8, obd, rtol=1e-5, atol=1e-6)
    assert replays >= 3, replays
    # a device-side argument error through the persistent pass, then exact again
    bad = qids.copy()
    bad[17] = len(x)
    with pytest.raises(GTSError) as e:
        idx.knn_query_dev(DeviceArray.from_host(idx, bad), 0, ids, dists, kth, ks_dev=d_k, learned=True)
    assert e.value.code == GTS_E_ARG
    idx.knn_query_dev(d_q, 0, ids, dists, kth, ks_dev=d_k, learned=True)
    _check_rows(ids.to_host()[sel], dists.to_host()[sel], kth.to_host()[sel], bids, bd, ks[sel])
