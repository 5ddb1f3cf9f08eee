"""GPU parity: the HIP index against the CPU oracle on the same seeded inputs.

Bar: result ID sets bit-identical to the exact (FP64) sets with ties broken by
object ID; FP32 output distances within 1e-5 relative of the FP64 distances
(1e-5 absolute below 1); edit distances exact.
"""
import numpy as np
import pytest

import oracle
from gts_amd import GTSIndex, datagen

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _close(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    return np.all(np.abs(got - want) <= RTOL * np.maximum(np.abs(want), 1.0))


def _vec_ds(n, dim, metric, seed, clusters=0):
    rng = np.random.default_rng(seed)
    if clusters:
        c = rng.standard_normal((clusters, dim))
        x = c[rng.integers(0, clusters, n)] + 0.3 * rng.standard_normal((n, dim))
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    else:
        x = rng.random((n, dim))
    return datagen.Dataset(metric, vectors=x.astype(np.float32))


def _check_knn(idx, od, qids, k):
    ids, dists, kth = idx.knn_query(qids, k)
    bids, bd = oracle.bf_knn(od, qids, k)
    assert (ids == bids).all(), "kNN ID sets differ"
    assert _close(dists, bd)
    assert _close(kth, bd[:, -1])
    return ids, dists


def _check_range(idx, od, qids, r):
    counts, offs, ids, dists = idx.range_query(qids, r, with_ids=True)
    bc, boffs, bids, bd = oracle.bf_range(od, qids, r, with_ids=True)
    assert (counts == bc).all(), "range counts differ"
    assert (offs == boffs).all()
    assert (ids == bids).all(), "range ID sets differ"
    assert _close(dists, bd)
    assert (idx.range_query(qids, r) == bc).all()


def test_c1_plain_small_l2():
    """20,000 x 300 L2, 50 clusters: range at the 1% quantile and kNN k = 8."""
    ds, qids, r = datagen.gaussian_mixture(n=20000, dim=300, clusters=50, nq=256, seed=0)
    print("c1: r =", r, flush=True)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds, 10, 20, 0)
    assert idx.info()["height"] == 4
    # the work the pruning leaves: verified (query, object) pairs of the plain
    # path against the oracle's tree search over the same T, leaf capacity and
    # height (FP64 ring tests there; FP32 here, whose error margins keep a few
    # more leaves: at least the oracle's pairs, at most 3% more; measured
    # 1.3% for range and equal for kNN)
    ot = oracle.OracleTree(od, 10, 20, 0)
    idx.range_query(qids, r)
    g_range = idx.stats()["verified_pairs"]
    _, o_range, _ = ot.range(qids, r)
    idx.knn_query(qids, 8)
    g_knn = idx.stats()["verified_pairs"]
    _, _, o_knn = ot.knn(qids, 8)
    print("verified pairs: range %d (oracle %d), knn %d (oracle %d)" % (g_range, o_range.sum(), g_knn,
          o_knn.sum()), flush=True)
    assert o_range.sum() <= g_range <= 1.03 * o_range.sum()
    assert o_knn.sum() <= g_knn <= 1.03 * o_knn.sum()
    _check_range(idx, od, qids, r)
    _check_knn(idx, od, qids, 8)
    st = idx.stats()
    assert st["verified_pairs"] <= 20000 * 256


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_low_dim_metrics_prune_and_match(metric):
    ds = _vec_ds(30000, 3, metric, seed=metric)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds, 10, 20, 0)
    qids = np.arange(0, 30000, 97, dtype=np.int32)
    d = [oracle.dist(od, int(qids[0]), j) for j in range(0, 30000, 11)]
    r = float(np.quantile(d, 0.01))
    _check_range(idx, od, qids, r)
    verified_range = idx.stats()["verified_pairs"]
    assert verified_range < 0.2 * 30000 * len(qids), "ring pruning should cut work in 3-d"
    _check_knn(idx, od, qids, 10)
    assert idx.stats()["verified_pairs"] < 0.2 * 30000 * len(qids)


def test_variable_k_and_large_k():
    ds = _vec_ds(5000, 8, 2, seed=7, clusters=20)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds, 10, 20, 0)
    qids = np.arange(0, 5000, 41, dtype=np.int32)
    ks = (np.arange(len(qids)) % 64 + 1).astype(np.int32)
    ids, dists, kth = idx.knn_query(qids, ks)
    for i, k in enumerate(ks):
        bids, bd = oracle.bf_knn(od, qids[i:i + 1], int(k))
        assert (ids[i, :k] == bids[0]).all()
        assert (ids[i, k:] == -1).all()
        assert _close(dists[i, :k], bd[0])
    _check_knn(idx, od, qids[:5], 300)


def test_ties_duplicates_and_repeated_queries():
    rng = np.random.default_rng(3)
    base = rng.random((200, 4)).astype(np.float32)
    x = np.repeat(base, 25, axis=0)          # 5000 rows, every point 25 times
    ds = datagen.Dataset(1, vectors=x)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds, 10, 20, 0)
    qids = np.array([0, 0, 24, 25, 4999, 1234, 1234], np.int32)
    _check_knn(idx, od, qids, 30)            # the 25 copies tie at distance 0
    _check_range(idx, od, qids, 0.0)
    _check_range(idx, od, qids, 0.3)


def test_knn_band_beyond_wave_capacity():
    """600 exact copies of one row: every copy is inside the error band of the
    k-th distance, more than the wave-level top-k holds, so the block-level
    selection takes those queries; ties are broken by ID."""
    rng = np.random.default_rng(8)
    x = rng.random((5000, 24)).astype(np.float32)
    x[1000:1600] = x[7]
    ds = datagen.Dataset(2, vectors=x)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    qids = np.array([7, 1000, 1599, 3, 4999], np.int32)
    for k in (1, 16, 64):
        _check_knn(idx, od, qids, k)


@pytest.mark.parametrize("n,T,ms,mh", [(15, 10, 20, 0), (205, 10, 20, 0), (2099, 10, 20, 0),
                                       (3000, 10, 20, 3), (777, 2, 3, 0), (1000, 32, 40, 0)])
def test_tree_shapes(n, T, ms, mh):
    """Root leaf, mixed leaf levels, capped height, binary and wide trees."""
    ds = _vec_ds(n, 5, 2, seed=n)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds, T, ms, mh)
    tr = idx.export_tree()
    ot = oracle.OracleTree(od, T, ms, mh)
    assert tr["height"] == ot.height
    assert sorted(tr["perm"].tolist()) == list(range(n))
    ex = (tr["flags"] & 1) != 0
    assert (ex == (ot.exists != 0)).all()
    assert (((tr["flags"] & 2) != 0)[ex] == (ot.is_leaf[ex] != 0)).all()
    assert (tr["size"][ex] == ot.size[ex]).all() and (tr["lid"][ex] == ot.lid[ex]).all()
    # children tile their parent's range in distance order from its pivot
    perm = tr["perm"]
    checked = 0
    for nd in np.nonzero(ex & ((tr["flags"] & 2) == 0))[0]:
        piv = tr["pivot"][nd]
        lid, size = tr["lid"][nd], tr["size"][nd]
        assert piv in set(perm[lid:lid + size].tolist())
        # the pivot rule id_list[(lid + rid) / 2] (tree.cuh:65-73), where the
        # node's order before its split is decided beyond FP32 error (the
        # GPU sorts FP32 keys, ties by ID): rank (size - 1) // 2 by (FP64
        # distance to the parent's pivot, ID), the root's identity order
        if nd == 0:
            assert piv == (size - 1) // 2
            checked += 1
        else:
            par = (nd - 1) // T
            members = [int(x) for x in perm[lid:lid + size]]
            keyed = sorted((oracle.dist(od, x, int(tr["pivot"][par])), x) for x in members)
            ks = np.array([k for k, _ in keyed])
            mid = (size - 1) // 2
            near = np.abs(ks - ks[mid]) <= 1e-5 * max(1.0, ks[mid])
            if near.sum() == 1:   # the middle key is distinct from every other
                assert piv == keyed[mid][1], (nd, piv, keyed[mid])
                checked += 1
        prev_hi = -np.inf
        for c in range(T):
            ch = nd * T + 1 + c
            cl, cs = tr["lid"][ch], tr["size"][ch]
            if cs == 0:
                continue
            d = np.array([oracle.dist(od, int(perm[p]), int(piv)) for p in range(cl, cl + cs)])
            assert d.min() >= prev_hi - 1e-5 * max(1, prev_hi)
            assert abs(tr["min_dis"][ch] - d.min()) <= 1e-5 * max(1.0, d.min())
            prev_hi = d.max()
    assert checked >= (ex & ((tr["flags"] & 2) == 0)).sum() // 2   # most nodes are decided
    qids = np.arange(0, n, max(1, n // 40), dtype=np.int32)
    _check_knn(idx, od, qids, min(n, 7))
    dd = [oracle.dist(od, 0, j) for j in range(n)]
    _check_range(idx, od, qids, float(np.quantile(dd, 0.05)))


def test_k_larger_than_n():
    ds = _vec_ds(12, 3, 2, seed=1)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds, 10, 20, 0)
    ids, dists, kth = idx.knn_query(np.array([0, 5], np.int32), 20)
    bids, bd = oracle.bf_knn(od, np.array([0, 5], np.int32), 12)
    assert (ids[:, :12] == bids).all() and (ids[:, 12:] == -1).all()
    assert np.isinf(dists[:, 12:]).all() and np.isinf(kth).all()


def test_edit_distance_strings():
    ds, qids = datagen.strings_edit(n=20000, roots=500, nq=200, seed=5)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds, 10, 20, 0)
    for r in (1, 2, 3):
        _check_range(idx, od, qids, r)
    ids, dists, kth = idx.knn_query(qids, 12)
    bids, bd = oracle.bf_knn(od, qids, 12)
    assert (ids == bids).all()
    assert (dists == bd).all()


def test_long_strings_fall_back_to_dp():
    rng = np.random.default_rng(9)
    strs = [rng.choice(np.frombuffer(b"ab", np.uint8), size=int(rng.integers(50, 120))).tobytes()
            for _ in range(600)]
    ds = datagen.Dataset(6, strings=strs)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds, 10, 20, 0)
    qids = np.arange(0, 600, 13, dtype=np.int32)
    _check_knn(idx, od, qids, 5)
    _check_range(idx, od, qids, 30)


def test_bad_arguments_raise():
    from gts_amd import GTSError
    ds = _vec_ds(100, 3, 2, seed=2)
    idx = GTSIndex(ds)
    with pytest.raises(GTSError):
        idx.knn_query(np.array([100], np.int32), 3)
    with pytest.raises(GTSError):
        idx.knn_query(np.array([1], np.int32), 0)
    with pytest.raises(GTSError):
        idx.range_query(np.array([-1], np.int32), 0.5)
    with pytest.raises(GTSError):
        GTSIndex(ds, order=1)


def test_high_dim_uses_staged_tiles():
    """dim 520 exceeds the resident-query LDS budget: the staged tile kernel runs."""
    ds = _vec_ds(3000, 520, 2, seed=4, clusters=10)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    qids = np.arange(0, 3000, 61, dtype=np.int32)
    _check_knn(idx, od, qids, 9)
    d = [oracle.dist(od, 0, j) for j in range(0, 3000, 3)]
    _check_range(idx, od, qids, float(np.quantile(d, 0.05)))
    _check_range(ds and GTSIndex(_vec_ds(3000, 520, 1, seed=5)), oracle.OracleData(_vec_ds(3000, 520, 1, seed=5)),
                 qids, 170.0)


def test_cli_result_writers_c1(tmp_path):
    """The CLI on the 20,000 x 300 L2 set (reference text format): kNN k = 8
    and range at the 1% quantile, with the saveK / saveR writers
    (file.cuh:154-199) -- every ID line equals brute force, distances within
    the "%f" rounding."""
    import os
    import subprocess
    ds, qids, r = datagen.gaussian_mixture(n=20000, dim=300, clusters=50, nq=256, seed=0)
    od = oracle.OracleData(ds)
    data, qf = tmp_path / "c1.txt", tmp_path / "q.txt"
    with open(data, "w") as f:
        f.write("300 20000 2\n")
        for row in ds.vectors:
            f.write(" ".join("%.9g" % v for v in row) + "\n")
    qf.write_text("%d\n" % len(qids) + "".join("%d\n" % q for q in qids))
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gts_amd", "bin", "gts")
    fid, fd = tmp_path / "ids.txt", tmp_path / "dis.txt"
    subprocess.run([exe, str(data), str(qf), "0", "8", str(tmp_path / "knn.txt"), "--ids", str(fid),
                    "--dists", str(fd)], check=True, capture_output=True, timeout=300)
    bids, bd = oracle.bf_knn(od, qids, 8)
    for i, (li, ld) in enumerate(zip(fid.read_text().split("\n")[:-1], fd.read_text().split("\n")[:-1])):
        assert li == "".join("%d " % v for v in bids[i])
        assert np.all(np.abs(np.array([float(v) for v in ld.split()]) - bd[i]) <= 1e-5 * np.maximum(bd[i], 1.0))
    rs = "%.6f" % r
    subprocess.run([exe, str(data), str(qf), "1", rs, str(tmp_path / "rng.txt"), "--ids", str(fid), "--dists",
                    str(fd)], check=True, capture_output=True, timeout=300)
    bc, boffs, bids, bdist = oracle.bf_range(od, qids, np.float32(rs), with_ids=True)
    lines = fid.read_text().split("\n")[:-1]
    assert len(lines) == len(qids) + len(qids)   # saveR appended after saveK's lines
    for i, li in enumerate(lines[len(qids):]):
        assert li == "".join("%d " % v for v in bids[boffs[i]:boffs[i + 1]])


def test_fp32_mfma_tiles_match_direct_tiles():
    """The FP32 form's L2 tiles on the matrix cores (norm expansion with its
    wider error band, re-decided in FP64) and the direct-difference tiles
    return identical kNN results and range counts, equal to the oracle, on
    rows with exact and near duplicates and norms spanning four decades."""
    rng = np.random.default_rng(21)
    x = rng.standard_normal((20000, 48)).astype(np.float32)
    x[:2000] *= 100.0
    x[2000:4000] *= 0.01
    x[5000:5100] = x[7]
    x[6000:6100] = x[9] + rng.normal(0, 1e-6, (100, 48)).astype(np.float32)
    ds = datagen.Dataset(2, vectors=x)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    qids = np.concatenate([np.array([7, 5000, 9, 6050, 0, 2500], np.int32),
                           rng.integers(0, 20000, 400).astype(np.int32)])
    r = float(np.quantile(np.linalg.norm(x[rng.integers(0, 20000, 2000)] - x[rng.integers(0, 20000, 2000)],
                                         axis=1), 0.01))
    got = {}
    for on in (True, False):
        idx.set_fp32_mfma(on)
        got[on] = (idx.knn_query(qids, 16), idx.knn_query(qids, 64), idx.range_query(qids, r))
    for a, b in zip(got[True], got[False]):
        if isinstance(a, tuple):
            assert all((u == v).all() for u, v in zip(a, b))
        else:
            assert (a == b).all()
    for kk, res in ((16, got[True][0]), (64, got[True][1])):
        bids, bd = oracle.bf_knn(od, qids, kk)
        assert (res[0] == bids).all()
        assert np.allclose(res[1], bd, rtol=1e-5, atol=1e-6)
    assert (got[True][2] == oracle.bf_range(od, qids, r)).all()
