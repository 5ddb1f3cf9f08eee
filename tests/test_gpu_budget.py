"""GPU tests of the memory-budgeted split of query calls (searchIndexRnnV2 /
searchIndexKnnV2 size their scratch from the free device memory and split the
query range depth-first, search_v2.cuh:791-828, 921-966, 1177-1221).  With a
budget forced far below one batch's working set, each call runs in several
sub-batches (stats budget_batches), and every result equals the unsplit call's
and the FP64 brute-force oracle's: range counts and ID lists, plain and
learned kNN (tree and cell-scan plans), the update stream's counts; a budget
below one query's search fails with GTS_E_OOM.
"""
import numpy as np
import pytest

import oracle
from gts_amd import (GTS_E_ARG, GTS_E_OOM, PLAN_CELLS, PLAN_PRUNED, VERIFY_BF16X3, GTSError,
                     GTSIndex, datagen)

pytestmark = pytest.mark.gpu


def _close(a, b):
    return np.allclose(np.asarray(a, np.float64), np.asarray(b, np.float64), rtol=1e-5, atol=1e-6)


def test_range_split_equals_unsplit_and_oracle():
    ds, qids, r = datagen.gaussian_mixture(n=30011, dim=32, clusters=40, nq=600, seed=4)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    c0, o0, i0, d0 = idx.range_query(qids, r, with_ids=True)
    assert idx.stats()["budget_batches"] == 1
    n0 = idx.range_query(qids, r)
    idx.set_memory_budget(1 << 20)
    c1, o1, i1, d1 = idx.range_query(qids, r, with_ids=True)
    s1 = idx.stats()
    print("range with IDs: %d sub-batches, %d results" % (s1["budget_batches"], len(i1)), flush=True)
    assert s1["budget_batches"] > 1
    assert (c1 == c0).all() and (o1 == o0).all() and (i1 == i0).all() and (d1 == d0).all()
    n1 = idx.range_query(qids, r)
    assert idx.stats()["budget_batches"] > 1
    assert (n1 == n0).all()
    bc, bo, bi, bd = oracle.bf_range(od, qids, r, with_ids=True)
    assert (c1 == bc).all() and (o1 == bo).all() and (i1 == bi).all()
    assert _close(d1, bd)


def test_knn_split_equals_unsplit_and_oracle():
    ds, qids, _ = datagen.gaussian_mixture(n=30011, dim=64, clusters=60, nq=1500, seed=6)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(2)
    ks = rng.integers(1, 33, len(qids)).astype(np.int32)
    want = {}
    for name, k, kw in (("plain", 16, {}), ("uneven", ks, {}), ("bf16", 16, {"verify": VERIFY_BF16X3})):
        want[name] = idx.knn_query(qids, k, **kw)
        assert idx.stats()["budget_batches"] == 1
    # at least 256 per-query candidate lists of 32 KB: 12 MB holds sub-batches of about 300
    idx.set_memory_budget(12 << 20)
    bids, bd = oracle.bf_knn(od, qids, 32)
    for name, k, kw in (("plain", 16, {}), ("uneven", ks, {}), ("bf16", 16, {"verify": VERIFY_BF16X3})):
        ids, d, kth = idx.knn_query(qids, k, **kw)
        st = idx.stats()
        print(name, "sub-batches", st["budget_batches"], flush=True)
        assert st["budget_batches"] >= 4
        assert (ids == want[name][0]).all() and _close(d, want[name][1]) and _close(kth, want[name][2])
        kk = np.full(len(qids), 16) if np.isscalar(k) else k
        for i in range(len(qids)):
            assert (ids[i, :kk[i]] == bids[i, :kk[i]]).all()
            assert _close(d[i, :kk[i]], bd[i, :kk[i]])


@pytest.mark.parametrize("plan", [PLAN_PRUNED, PLAN_CELLS])
def test_learned_knn_split(plan):
    ds, qids, _ = datagen.gaussian_mixture(n=40009, dim=128, clusters=80, nq=2048, seed=8)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(4)
    idx.learned_train(rng.integers(0, 40009, 1500).astype(np.int32), kmax=16, epochs=30)
    idx.set_knn_plan(plan)
    ids0, d0, _ = idx.knn_query(qids, 16, learned=True, verify=VERIFY_BF16X3)
    idx.set_memory_budget(20 << 20)
    for _ in range(2):   # the second call captures each sub-batch's first pass into a graph
        ids1, d1, _ = idx.knn_query(qids, 16, learned=True, verify=VERIFY_BF16X3)
        assert idx.stats()["budget_batches"] >= 4
        assert (ids1 == ids0).all() and _close(d1, d0)
    bids, bd = oracle.bf_knn(od, qids, 16)
    assert (ids1 == bids).all() and _close(d1, bd)


def test_update_split_equals_oracle():
    rng = np.random.default_rng(9)
    x = rng.random((6000, 8)).astype(np.float32)
    ds = datagen.Dataset(1, vectors=x)
    od = oracle.OracleData(ds)
    kinds = rng.choice(3, size=3000, p=[0.3, 0.2, 0.5]).astype(np.int8)
    fresh = rng.random((1200, 8)).astype(np.float32)
    ops = oracle.update_stream(6000, len(fresh), kinds, rng.random(3000), 200)
    want = oracle.update_counts(od, ops[:, 0], ops[:, 1], 0.9, cache_cap=200, fresh=fresh)
    idx = GTSIndex(datagen.Dataset(1, vectors=x))
    idx.set_memory_budget(64 << 10)
    got = idx.update(ops, 0.9, cache_cap=200, fresh=fresh)
    assert idx.stats()["budget_batches"] > 1
    assert list(got) == list(want)


def test_budget_limits():
    ds, qids, r = datagen.gaussian_mixture(n=5000, dim=16, clusters=10, nq=50, seed=1)
    idx = GTSIndex(ds)
    with pytest.raises(GTSError) as e:
        idx.set_memory_budget(-1)
    assert e.value.code == GTS_E_ARG
    idx.set_memory_budget(64)   # below one query's first level
    with pytest.raises(GTSError) as e:
        idx.range_query(qids, r)
    assert e.value.code == GTS_E_OOM
    idx.set_memory_budget(1 << 20)   # below 256 kNN candidate lists
    with pytest.raises(GTSError) as e:
        idx.knn_query(qids, 4)
    assert e.value.code == GTS_E_OOM
    idx.set_memory_budget(0)   # automatic again: the same index answers exactly
    ids, _, _ = idx.knn_query(qids, 4)
    assert (ids == oracle.bf_knn(oracle.OracleData(ds), qids, 4)[0]).all()
