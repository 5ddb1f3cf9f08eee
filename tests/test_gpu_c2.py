"""The headline workload at full size on the GPU: 1,000,000 x 300 FP32 rows
(1,000-cluster Gaussian mixture, sigma 0.3, rows L2-normalised, seed 0), L2,
10,000 concurrent kNN queries, k in {8, 16, 32}.

For each k:
* the plain path (the reference algorithm: ring pruning with the k-th pivot
  bound, FP32 verification, search_v2.cuh:975-1228) returns IDs equal to the
  FP64 brute-force oracle on a fixed sample of 64 queries over all N rows,
  distances within 1e-5 relative;
* the learned form (radius network trained on plain-kNN radii, BF16 hi+lo
  MFMA verification, the automatic plan -- the cell scan at this size) called
  through the device entry point three times (plan and tune, graph capture,
  graph replay) returns IDs identical to the plain path on every query, and
  distances and k-th distances within 1e-5 relative of the plain path's.
"""
import time

import numpy as np
import pytest

import oracle
from gts_amd import VERIFY_BF16X3, VERIFY_FP32, DeviceArray, GTSIndex, datagen

pytestmark = pytest.mark.gpu

N, DIM, CLUSTERS, NQ = 1_000_000, 300, 1000, 10_000
KS = (8, 16, 32)


def _log(*a):
    print(*a, flush=True)


def _close(a, b):
    return np.all(np.abs(a - b) <= 1e-5 * np.maximum(np.abs(b), 1.0))


@pytest.fixture(scope="module")
def c2():
    t = time.time()
    ds, qids, _ = datagen.gaussian_mixture(n=N, dim=DIM, clusters=CLUSTERS, nq=NQ, seed=0)
    _log("C2 generated in %.1f s" % (time.time() - t))
    t = time.time()
    idx = GTSIndex(ds, 10, 20, 0)
    _log("C2 index built in %.1f s: %s" % (time.time() - t, idx.info()))
    sel = np.arange(0, NQ, NQ // 64)[:64]
    t = time.time()
    bids, bd = oracle.bf_knn(oracle.OracleData(ds), qids[sel], max(KS))
    _log("C2 oracle sample (64 queries, k=%d) in %.1f s" % (max(KS), time.time() - t))
    return ds, qids, idx, sel, bids, bd


@pytest.mark.parametrize("k", KS)
def test_c2_plain_and_learned_match_oracle(c2, k):
    ds, qids, idx, sel, bids, bd = c2
    t = time.time()
    pid, pd, pk = idx.knn_query(qids, k, learned=False, verify=VERIFY_FP32)
    _log("k=%d plain: %.3f s, stats %s" % (k, time.time() - t, idx.stats()))
    assert (pid[sel] == bids[:, :k]).all()
    assert _close(pd[sel], bd[:, :k])
    assert _close(pk[sel], bd[:, k - 1])
    assert (pid >= 0).all()

    rng = np.random.default_rng(1)
    train = rng.integers(0, N, size=4000).astype(np.int32)
    loss = idx.learned_train(train, kmax=k, epochs=60)
    _log("k=%d network trained, loss %.4f" % (k, loss))
    assert np.isfinite(loss)
    d_q = DeviceArray.from_host(idx, qids.astype(np.int32))
    ids_d = idx.device_array((NQ, k), np.int32)
    dists_d = idx.device_array((NQ, k), np.float32)
    kth_d = idx.device_array(NQ, np.float32)
    for call in range(3):   # plan + tune, capture, replay
        ids_d.copy_from(DeviceArray.from_host(idx, np.full((NQ, k), -7, np.int32)))
        idx.knn_query_dev(d_q, k, ids_d, dists_d, kth_d, learned=True, verify=VERIFY_BF16X3)
        st = idx.stats()
        _log("k=%d learned call %d: stats %s" % (k, call, st))
        lid, ld, lk = ids_d.to_host(), dists_d.to_host(), kth_d.to_host()
        assert st["cell_scan"] > 0, "the automatic plan should pick the cell scan at this size"
        assert (lid == pid).all(), "learned IDs differ from the plain path on %d queries" % int(
            (lid != pid).any(axis=1).sum())
        assert _close(ld, pd) and _close(lk, pk)
    assert st["graph_replays"] == 1
