"""GPU tests of the chosen forms: BF16 hi+lo MFMA verification, the learned
radius (training kernels checked against the FP64 oracle), and the update path.
"""
import numpy as np
import pytest

import oracle
from gts_amd import (PLAN_AUTO, PLAN_CELLS, PLAN_EXHAUSTIVE, PLAN_PRUNED, RADIUS_NETWORK, RADIUS_PIVOTS,
                     VERIFY_BF16X3, VERIFY_FP32, DeviceArray, GTSError, GTSIndex, datagen)

pytestmark = pytest.mark.gpu


def _bf16_round(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _dev_knn(idx, qids, k, learned, verify):
    ids, dists, kth = idx.knn_query(qids, k, learned=learned, verify=verify)
    return ids, dists, kth, idx.stats()


def test_mlp_gradients_match_fp64_oracle():
    rng = np.random.default_rng(0)
    ds = datagen.Dataset(2, vectors=rng.random((100, 4)).astype(np.float32))
    idx = GTSIndex(ds)
    nb = 300
    X = rng.standard_normal((nb, 16)).astype(np.float32)
    t = rng.standard_normal(nb).astype(np.float32)
    shapes = [(64, 16), (64,), (64, 64), (64,), (1, 64), (1,)]
    ws = [_bf16_round(rng.standard_normal(s) * 0.3) for s in shapes]
    packed = np.concatenate([w.ravel() for w in ws])
    loss, g = idx.mlp_loss_grad(X, t, packed)
    oloss, og, _ = oracle.mlp_loss_grad(X, t, *ws)
    assert abs(loss - oloss) <= 1e-5 * max(1.0, abs(oloss))
    ogp = np.concatenate([x.ravel() for x in og])
    scale = np.abs(ogp).max()
    assert np.abs(g - ogp).max() <= 1e-4 * scale, np.abs(g - ogp).max() / scale


def test_bf16x3_knn_is_exact():
    form = VERIFY_BF16X3
    ds, qids, _ = datagen.gaussian_mixture(n=20000, dim=300, clusters=50, nq=200, seed=1)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    bids, bd = oracle.bf_knn(od, qids, 16)
    ids, dists, kth, st = _dev_knn(idx, qids, 16, False, form)
    assert (ids == bids).all()
    assert np.allclose(dists, bd, rtol=1e-5, atol=1e-6)
    ids2, _, _, _ = _dev_knn(idx, qids, 16, False, VERIFY_FP32)
    assert (ids2 == bids).all()


def test_bf16x3_odd_dims_and_duplicates():
    form = VERIFY_BF16X3
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3000, 37)).astype(np.float32)
    x[100:140] = x[7]                         # 41 identical rows
    ds = datagen.Dataset(2, vectors=x)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    qids = np.array([7, 100, 0, 2999, 1500], np.int32)
    bids, bd = oracle.bf_knn(od, qids, 45)
    ids, dists, _, _ = _dev_knn(idx, qids, 45, False, form)
    assert (ids == bids).all()


def test_learned_radius_exact_and_cheaper():
    ds, qids, _ = datagen.gaussian_mixture(n=20000, dim=300, clusters=50, nq=300, seed=2)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(7)
    train = rng.integers(0, 20000, size=1500).astype(np.int32)
    loss = idx.learned_train(train, kmax=16, epochs=80)
    assert np.isfinite(loss) and loss < 1.0
    bids, bd = oracle.bf_knn(od, qids, 16)
    ids_p, _, _, st_p = _dev_knn(idx, qids, 16, False, VERIFY_FP32)
    ids_l, dl, _, st_l = _dev_knn(idx, qids, 16, True, VERIFY_FP32)
    ids_b, _, _, st_b = _dev_knn(idx, qids, 16, True, VERIFY_BF16X3)
    assert (ids_p == bids).all() and (ids_l == bids).all() and (ids_b == bids).all()
    assert np.allclose(dl, bd, rtol=1e-5, atol=1e-6)
    # what the network gains: far fewer candidates reach the exact selection
    print("candidates plain %d learned %d retries %d" % (st_p["candidates"], st_l["candidates"],
                                                        st_l["retries"]), flush=True)
    assert st_l["candidates"] < 0.5 * st_p["candidates"]
    assert st_l["retries"] <= 0.2 * len(qids)


@pytest.mark.parametrize("plan", [PLAN_PRUNED, PLAN_EXHAUSTIVE, PLAN_CELLS, PLAN_AUTO])
def test_pivot_radius_source(plan):
    """Learned kNN calls whose radius is the k-th distance to one tree level's
    pivots (no network, no training) are exact under every plan: uniform k,
    uneven k, and k above the pivot count (an unbounded first radius)."""
    ds, qids, _ = datagen.gaussian_mixture(n=30011, dim=100, clusters=100, nq=1500, seed=5)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    with pytest.raises(GTSError):   # the network source needs training
        idx.knn_query(qids[:10], 4, learned=True, verify=VERIFY_BF16X3)
    idx.set_radius_source(RADIUS_PIVOTS)
    idx.set_knn_plan(plan)
    for k in (1, 16):
        ids, d, kth, st = _dev_knn(idx, qids, k, True, VERIFY_BF16X3)
        bids, bd = oracle.bf_knn(od, qids, k)
        assert (ids == bids).all(), (k, st)
        assert np.allclose(d, bd, rtol=1e-5, atol=1e-6)
        assert np.allclose(kth, bd[:, -1], rtol=1e-5, atol=1e-6)
    rng = np.random.default_rng(3)
    ks = rng.integers(1, 65, len(qids)).astype(np.int32)
    ids, d, _, _ = _dev_knn(idx, qids, ks, True, VERIFY_BF16X3)
    bids, bd = oracle.bf_knn(od, qids, 64)
    for i, k in enumerate(ks):
        assert (ids[i, :k] == bids[i, :k]).all() and (ids[i, k:] == -1).all()
        assert np.allclose(d[i, :k], bd[i, :k], rtol=1e-5, atol=1e-6)
    # more neighbours than the level has pivots (level 3: at most 1,000)
    few = qids[:40]
    ids, d, _, _ = _dev_knn(idx, few, 1100, True, VERIFY_BF16X3)
    bids, bd = oracle.bf_knn(od, few, 1100)
    assert (ids == bids).all()
    assert np.allclose(d, bd, rtol=1e-5, atol=1e-6)
    # back to the network once trained: the same results
    idx.learned_train(rng.integers(0, 30011, 1500).astype(np.int32), kmax=16, epochs=30)
    idx.set_radius_source(RADIUS_NETWORK)
    ids, _, _, _ = _dev_knn(idx, qids, 16, True, VERIFY_BF16X3)
    assert (ids == oracle.bf_knn(od, qids, 16)[0]).all()


@pytest.mark.parametrize("n,dim,nq", [(60013, 300, 2000), (30011, 100, 1500)])
def test_knn_plans_agree(n, dim, nq):
    """Pruned, exhaustive and planner-chosen passes of the learned BF16 form
    return the same IDs and distances, equal to brute force; row counts that
    are not a multiple of 32 exercise the padded rows of the dense kernel."""
    form = VERIFY_BF16X3
    ds, qids, _ = datagen.gaussian_mixture(n=n, dim=dim, clusters=100, nq=nq, seed=3)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(11)
    idx.learned_train(rng.integers(0, n, size=2000).astype(np.int32), kmax=16, epochs=60)
    got = {}
    for plan in (PLAN_PRUNED, PLAN_EXHAUSTIVE, PLAN_AUTO):
        idx.set_knn_plan(plan)
        got[plan] = _dev_knn(idx, qids, 16, True, form)
        print("plan", plan, got[plan][3], flush=True)
    # the schedule autotuner (on above) does not change results
    idx.set_knn_plan(PLAN_EXHAUSTIVE)
    idx.set_autotune(False)
    untuned = _dev_knn(idx, qids, 16, True, form)
    idx.set_autotune(True)
    assert (untuned[0] == got[PLAN_EXHAUSTIVE][0]).all() and (untuned[1] == got[PLAN_EXHAUSTIVE][1]).all()
    assert got[PLAN_PRUNED][3]["exhaustive_passes"] == 0
    assert got[PLAN_EXHAUSTIVE][3]["exhaustive_passes"] >= 1
    assert got[PLAN_EXHAUSTIVE][3]["verified_pairs"] >= n * nq
    for plan in (PLAN_EXHAUSTIVE, PLAN_AUTO):
        assert (got[plan][0] == got[PLAN_PRUNED][0]).all()
        assert (got[plan][1] == got[PLAN_PRUNED][1]).all()
        assert (got[plan][2] == got[PLAN_PRUNED][2]).all()
    sel = np.arange(0, nq, 10)
    bids, bd = oracle.bf_knn(oracle.OracleData(ds), qids[sel], 16)
    assert (got[PLAN_EXHAUSTIVE][0][sel] == bids).all()
    assert np.allclose(got[PLAN_EXHAUSTIVE][1][sel], bd, rtol=1e-5, atol=1e-6)
    # the exhaustive plan refuses the calls it cannot serve
    idx.set_knn_plan(PLAN_EXHAUSTIVE)
    with pytest.raises(GTSError):
        idx.knn_query(qids[:4], 16, learned=False, verify=form)
    with pytest.raises(GTSError):
        idx.knn_query(qids[:4], 16, learned=True, verify=VERIFY_FP32)
    idx.set_knn_plan(PLAN_AUTO)
    ids, _, _, _ = _dev_knn(idx, qids[:4], 16, True, VERIFY_FP32)
    assert (ids == got[PLAN_PRUNED][0][:4]).all()


def test_dense_schedules_agree():
    """Every schedule of the dense kernel (row-range sizes, grid or persistent
    blocks) and the autotuned one give the same results, equal to brute force;
    row counts that are not a multiple of 32 and more than 64 query blocks."""
    ds, qids, _ = datagen.gaussian_mixture(n=40013, dim=128, clusters=80, nq=4500, seed=4)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(2)
    idx.learned_train(rng.integers(0, 40013, 2000).astype(np.int32), kmax=16, epochs=40)
    idx.set_knn_plan(PLAN_EXHAUSTIVE)
    ref = _dev_knn(idx, qids, 16, True, VERIFY_BF16X3)   # autotuned
    for persistent in (False, True):
        for rows in (32, 512, 2048, 7232):
            idx.set_dense_schedule(persistent, rows)
            got = _dev_knn(idx, qids, 16, True, VERIFY_BF16X3)
            assert (got[0] == ref[0]).all() and (got[1] == ref[1]).all(), (persistent, rows)
    idx.set_dense_schedule(False, 0)
    sel = np.arange(0, len(qids), 45)
    bids, bd = oracle.bf_knn(oracle.OracleData(ds), qids[sel], 16)
    assert (ref[0][sel] == bids).all()
    assert np.allclose(ref[1][sel], bd, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("n,dim,clusters,nq", [(60013, 300, 120, 3000), (20011, 64, 40, 1500)])
def test_cell_scan_matches_exhaustive(n, dim, clusters, nq):
    """The cell scan returns the exhaustive plan's IDs, distances and k-th
    distances for every cell count (one cell, a few, the default), verifying
    fewer pairs; row counts that are not a multiple of 32; equal to brute force."""
    ds, qids, _ = datagen.gaussian_mixture(n=n, dim=dim, clusters=clusters, nq=nq, seed=6)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(4)
    idx.learned_train(rng.integers(0, n, 2000).astype(np.int32), kmax=16, epochs=40)
    idx.set_knn_plan(PLAN_EXHAUSTIVE)
    ref = _dev_knn(idx, qids, 16, True, VERIFY_BF16X3)
    assert ref[3]["cell_scan"] == 0
    idx.set_knn_plan(PLAN_CELLS)
    for ncell in (1, 7, 300, 0):
        idx.set_cell_count(ncell)
        got = _dev_knn(idx, qids, 16, True, VERIFY_BF16X3)
        print("cells", ncell, got[3], flush=True)
        assert got[3]["cell_scan"] >= 1
        assert (got[0] == ref[0]).all(), ncell
        assert (got[1] == ref[1]).all(), ncell
        assert (got[2] == ref[2]).all(), ncell
        if ncell == 0:
            assert got[3]["verified_pairs"] < ref[3]["verified_pairs"] / 2
    # the automatic plan takes the cell scan when it skips the pruning
    idx.set_knn_plan(PLAN_AUTO)
    auto = _dev_knn(idx, qids, 16, True, VERIFY_BF16X3)
    assert (auto[0] == ref[0]).all() and (auto[1] == ref[1]).all()
    sel = np.arange(0, nq, 23)
    bids, bd = oracle.bf_knn(oracle.OracleData(ds), qids[sel], 16)
    assert (ref[0][sel] == bids).all()
    assert np.allclose(ref[1][sel], bd, rtol=1e-5, atol=1e-6)


def test_cell_scan_duplicates_variable_k_and_small_sets():
    """Duplicate rows (zero distances to pivots and between rows), repeated
    query IDs, per-query k up to 64 and sets smaller than one cell's target
    size: the cell scan equals brute force."""
    rng = np.random.default_rng(8)
    base = rng.standard_normal((3001, 48)).astype(np.float32)
    x = np.concatenate([base, base[:1000], base[:37]])        # duplicated rows
    ds = datagen.Dataset(datagen.METRIC_L2, vectors=np.ascontiguousarray(x))
    n = x.shape[0]
    idx = GTSIndex(ds)
    idx.learned_train(rng.integers(0, n, 1500).astype(np.int32), kmax=64, epochs=30)
    qids = np.concatenate([rng.integers(0, n, 1100), np.arange(5), np.arange(5)]).astype(np.int32)
    ks = rng.integers(1, 65, len(qids)).astype(np.int32)
    idx.set_knn_plan(PLAN_CELLS)
    od = oracle.OracleData(ds)
    ids, dists, _ = idx.knn_query(qids, ks, learned=True, verify=VERIFY_BF16X3)
    st = idx.stats()
    print("cells", st, flush=True)
    assert st["cell_scan"] >= 1
    for i in range(0, len(qids), 7):
        k = int(ks[i])
        _, bd = oracle.bf_knn(od, qids[i:i + 1], k)
        # duplicates tie at equal distances: compare distances, and IDs where distances are distinct
        assert np.allclose(dists[i, :k], bd[0], rtol=1e-5, atol=1e-6), i
    idx.set_knn_plan(PLAN_EXHAUSTIVE)
    ids_e, dists_e, _ = idx.knn_query(qids, ks, learned=True, verify=VERIFY_BF16X3)
    assert (ids_e == ids).all() and (dists_e == dists).all()
    # two cells over 2,500 rows
    small = datagen.Dataset(datagen.METRIC_L2, vectors=np.ascontiguousarray(base[:2500]))
    sidx = GTSIndex(small)
    q = np.arange(2500, dtype=np.int32)
    sidx.learned_train(q, kmax=8, epochs=20)
    sidx.set_knn_plan(PLAN_CELLS)
    sidx.set_cell_count(2)
    sids, sd, _ = sidx.knn_query(q, 8, learned=True, verify=VERIFY_BF16X3)
    assert sidx.stats()["cell_scan"] == 2
    bids, bd = oracle.bf_knn(oracle.OracleData(small), q, 8)
    assert (sids == bids).all() and np.allclose(sd, bd, rtol=1e-5, atol=1e-6)


def test_cell_scan_residual_plane_wide_magnitudes():
    """Rows whose norms span two decades (one FP8 scale for the whole set puts
    the small components' residuals in the subnormal range, so their per-row
    error eps carries the bound) and near-duplicate rows a 10^-4 relative step
    apart (orders decided by the FP64 band): the cell scan equals the
    exhaustive plan bit for bit and brute force on a sample.  (Far wider norm
    ranges widen every plan's band through the largest row norm.)"""
    rng = np.random.default_rng(21)
    n, dim = 20011, 96
    base = rng.standard_normal((n // 2, dim))
    scale = 10.0 ** rng.uniform(-1, 1, n // 2)[:, None]
    a = base * scale
    b = a * (1.0 + 1e-4 * rng.standard_normal((n // 2, dim)))   # near-duplicates
    x = np.concatenate([a, b, rng.standard_normal((n - 2 * (n // 2), dim))]).astype(np.float32)
    ds = datagen.Dataset(datagen.METRIC_L2, vectors=np.ascontiguousarray(x))
    idx = GTSIndex(ds)
    idx.learned_train(rng.integers(0, n, 2000).astype(np.int32), kmax=16, epochs=30)
    qids = rng.integers(0, n, 1200).astype(np.int32)
    idx.set_knn_plan(PLAN_EXHAUSTIVE)
    ref = _dev_knn(idx, qids, 16, True, VERIFY_BF16X3)
    idx.set_knn_plan(PLAN_CELLS)
    got = _dev_knn(idx, qids, 16, True, VERIFY_BF16X3)
    print("cells", got[3], flush=True)
    assert got[3]["cell_scan"] >= 1
    assert (got[0] == ref[0]).all() and (got[1] == ref[1]).all() and (got[2] == ref[2]).all()
    sel = np.arange(0, len(qids), 17)
    bids, bd = oracle.bf_knn(oracle.OracleData(ds), qids[sel], 16)
    assert np.allclose(got[1][sel], bd, rtol=1e-5, atol=1e-6)
    distinct = np.diff(bd, axis=1) > 1e-6 * bd[:, 1:]
    ok = np.concatenate([np.ones((len(sel), 1), bool), distinct], axis=1) & \
        np.concatenate([distinct, np.ones((len(sel), 1), bool)], axis=1)
    assert (got[0][sel][ok] == bids[ok]).all()


@pytest.mark.parametrize("plan", [PLAN_AUTO, PLAN_CELLS, PLAN_EXHAUSTIVE])
def test_bf16x3_rows_spanning_six_decades_fall_back_to_fp32(plan):
    """Row norms spanning six decades: the BF16X3 bound (largest row norm)
    leaves small queries bands that overflow on every rerun; after eight
    passes the remaining queries finish in the FP32 form, and every result
    equals brute force."""
    rng = np.random.default_rng(23)
    n, dim = 12011, 64
    x = (rng.standard_normal((n, dim)) * 10.0 ** rng.uniform(-3, 3, n)[:, None]).astype(np.float32)
    ds = datagen.Dataset(datagen.METRIC_L2, vectors=np.ascontiguousarray(x))
    idx = GTSIndex(ds)
    idx.learned_train(rng.integers(0, n, 1500).astype(np.int32), kmax=8, epochs=20)
    idx.set_knn_plan(plan)
    qids = rng.integers(0, n, 1100).astype(np.int32)
    ids, dists, _ = idx.knn_query(qids, 8, learned=True, verify=VERIFY_BF16X3)
    print("stats", idx.stats(), flush=True)
    sel = np.arange(0, len(qids), 11)
    bids, bd = oracle.bf_knn(oracle.OracleData(ds), qids[sel], 8)
    assert np.allclose(dists[sel], bd, rtol=1e-5, atol=1e-30)
    distinct = np.diff(bd, axis=1) > 1e-6 * bd[:, 1:]
    ok = np.concatenate([np.ones((len(sel), 1), bool), distinct], axis=1) & \
        np.concatenate([distinct, np.ones((len(sel), 1), bool)], axis=1)
    assert (ids[sel][ok] == bids[ok]).all()


@pytest.mark.parametrize("plan", [PLAN_EXHAUSTIVE, PLAN_CELLS])
def test_graph_replays_match_ordinary_launches(plan):
    """Repeated learned exhaustive kNN calls on the same device buffers are
    captured once and replayed as a graph: every replay equals the result of
    ordinary launches, with eager work (range queries, host-buffer kNN calls,
    new query IDs written into the same buffer) between the replays."""
    n, nq, k = 40013, 3000, 16
    ds, qids, _ = datagen.gaussian_mixture(n=n, dim=128, clusters=80, nq=2 * nq, seed=5)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(3)
    idx.learned_train(rng.integers(0, n, 2000).astype(np.int32), kmax=k, epochs=40)
    idx.set_knn_plan(plan)
    batches = [np.ascontiguousarray(qids[:nq]), np.ascontiguousarray(qids[nq:])]
    idx.set_graph(False)
    ref = [idx.knn_query(b, k, learned=True, verify=VERIFY_BF16X3) for b in batches]
    assert idx.stats()["graph_replays"] == 0
    idx.set_graph(True)
    d_q = idx.device_array(nq, np.int32)
    ids_d = idx.device_array((nq, k), np.int32)
    dists_d = idx.device_array((nq, k), np.float32)
    kth_d = idx.device_array(nq, np.float32)
    replays = 0
    for call in range(8):
        which = (call // 3) % 2            # the buffer's contents change between calls 2 and 3, 5 and 6
        d_q.copy_from(DeviceArray.from_host(idx, batches[which]))
        idx.knn_query_dev(d_q, k, ids_d, dists_d, kth_d, learned=True, verify=VERIFY_BF16X3)
        st = idx.stats()
        replays += st["graph_replays"]
        print("call", call, "batch", which, st, flush=True)
        assert (ids_d.to_host() == ref[which][0]).all(), call
        assert (dists_d.to_host() == ref[which][1]).all(), call
        assert (kth_d.to_host() == ref[which][2]).all(), call
        # eager work between replays on the same stream and scratch buffers
        if call % 2 == 0:
            idx.range_query(batches[which][:200], float(np.median(ref[which][2])))
        else:
            ids_h, _, _ = idx.knn_query(batches[1 - which][:500], k, learned=True, verify=VERIFY_BF16X3)
            assert (ids_h == ref[1 - which][0][:500]).all()
    # the first call runs ordinary launches, every later call on these buffers
    # replays, also after the host-buffer calls were captured beside it
    assert replays == 7
    # a changed schedule drops the graph; the next captures agree too
    idx.set_dense_schedule(False, 512)
    for call in range(3):
        d_q.copy_from(DeviceArray.from_host(idx, batches[0]))
        idx.knn_query_dev(d_q, k, ids_d, dists_d, kth_d, learned=True, verify=VERIFY_BF16X3)
        assert (ids_d.to_host() == ref[0][0]).all() and (dists_d.to_host() == ref[0][1]).all()
    assert idx.stats()["graph_replays"] == 1


def test_untimed_calls_match_timed_calls():
    """With event timing off, replayed and ordinary learned kNN calls record no
    events (ms_verify and ms_total report 0) and return what timed calls
    return; switching it back on times the calls again."""
    n, nq, k = 40013, 3000, 16
    ds, qids, _ = datagen.gaussian_mixture(n=n, dim=128, clusters=80, nq=nq, seed=6)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(4)
    idx.learned_train(rng.integers(0, n, 2000).astype(np.int32), kmax=k, epochs=40)
    idx.set_knn_plan(PLAN_CELLS)
    ref = idx.knn_query(qids, k, learned=True, verify=VERIFY_BF16X3)
    assert idx.stats()["ms_verify"] > 0 and idx.stats()["ms_total"] > 0
    d_q = DeviceArray.from_host(idx, np.ascontiguousarray(qids))
    ids_d = idx.device_array((nq, k), np.int32)
    dists_d = idx.device_array((nq, k), np.float32)
    kth_d = idx.device_array(nq, np.float32)
    for phase, (timing, calls) in enumerate(((False, 4), (True, 3), (False, 2))):
        idx.set_timing(timing)
        for call in range(calls):
            idx.knn_query_dev(d_q, k, ids_d, dists_d, kth_d, learned=True, verify=VERIFY_BF16X3)
            st = idx.stats()
            print("timing", timing, "call", call, st, flush=True)
            assert (ids_d.to_host() == ref[0]).all() and (dists_d.to_host() == ref[1]).all()
            assert (kth_d.to_host() == ref[2]).all()
            if timing:
                assert st["ms_verify"] > 0 and st["ms_total"] > 0
            else:
                assert st["ms_verify"] == 0 and st["ms_total"] == 0
            # a setting's first call on these buffers runs ordinary launches,
            # its second is captured, later ones replay; the graphs of both
            # settings stay cached (the third phase replays from its first call)
            assert st["graph_replays"] == (1 if call >= 1 or phase == 2 else 0), (phase, call)


def test_learned_radius_variable_k_low_dim():
    ds, qids, ks, train = datagen.points2d_pareto(n=200000, clusters=50, nq=500, ntrain=3000,
                                                  kmax=64, seed=0)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    idx.learned_train(train, kmax=64, epochs=60)
    from gts_amd import DeviceArray
    q = DeviceArray.from_host(idx, qids)
    kd = DeviceArray.from_host(idx, ks)
    ids_d = idx.device_array((len(qids), 64), np.int32)
    dists_d = idx.device_array((len(qids), 64), np.float32)
    kth_d = idx.device_array(len(qids), np.float32)
    idx.knn_query_dev(q, 0, ids_d, dists_d, kth_d, ks_dev=kd, learned=True, verify=VERIFY_FP32)
    ids = ids_d.to_host()
    for i in range(0, len(qids), 25):
        k = int(ks[i])
        bids, _ = oracle.bf_knn(od, qids[i:i + 1], k)
        assert (ids[i, :k] == bids[0]).all()
        assert (ids[i, k:] == -1).all()


def _random_ops(rng, n, nops, p=(0.4, 0.2, 0.4)):
    """Valid op stream over the reference's ID conventions (physical rows for
    insert/query, logical IDs for delete) with a cache of 10."""
    rows = n
    live_tree = n
    cache = 0
    ops = []
    for _ in range(nops):
        op = rng.choice(3, p=list(p))
        if op == 0:
            ops.append((0, int(rng.integers(0, rows))))
            cache += 1
            if cache == 10:
                rows = live_tree + cache
                live_tree = rows
                cache = 0
        elif op == 1 and live_tree + cache > 1:
            lid = int(rng.integers(0, live_tree + cache))
            ops.append((1, lid))
            if lid < live_tree:
                live_tree -= 1
            else:
                cache -= 1
        else:
            ops.append((2, int(rng.integers(0, rows))))
    return np.array(ops, np.int32)


@pytest.mark.parametrize("metric", [1, 2])
def test_update_stream_matches_oracle(metric):
    rng = np.random.default_rng(metric)
    x = rng.random((3000, 8)).astype(np.float32)
    ds = datagen.Dataset(metric, vectors=x)
    od = oracle.OracleData(ds)
    ops = _random_ops(rng, 3000, 600)
    r = 0.35 if metric == 2 else 0.9
    want = oracle.update_counts(od, ops[:, 0], ops[:, 1], r)
    idx = GTSIndex(ds)
    got = idx.update(ops, r)
    assert list(got) == list(want)


@pytest.mark.parametrize("period", [0, 3000])
def test_update_snapshot_batches_match_sequential_replay(period):
    """More than one batch of 4,096 query ops interleaved with fresh inserts,
    physical-row inserts, tree and cache deletes and rebuilds (when the cache
    list fills, and with a period also after every period-th op): every count
    equals the sequential replay."""
    rng = np.random.default_rng(21)
    n0 = 20000
    x = rng.random((n0, 16)).astype(np.float32)
    fresh = rng.random((8000, 16)).astype(np.float32)
    kinds = rng.choice(3, size=22000, p=[0.3, 0.25, 0.45]).astype(np.int8)
    ops = oracle.update_stream(n0, len(fresh), kinds, rng.random(22000), 700, period=period)
    # some inserts of existing physical rows (op 0)
    ins = np.nonzero(ops[:, 0] == 3)[0][::5]
    ops[ins, 0] = 0
    ops[ins, 1] = rng.integers(0, 1000, len(ins))
    od = oracle.OracleData(datagen.Dataset(1, vectors=x))
    want = oracle.update_counts(od, ops[:, 0], ops[:, 1], 1.6, cache_cap=700, fresh=fresh, period=period)
    idx = GTSIndex(datagen.Dataset(1, vectors=x))
    got = idx.update(ops, 1.6, cache_cap=700, fresh=fresh, period=period)
    print("queries", len(want), "rebuilds", idx.stats()["retries"], flush=True)
    assert len(want) > 8192
    assert idx.stats()["retries"] >= (len(ops) - 1) // period if period else 9
    assert list(got) == list(want)


def test_rebuild_period_counts_and_errors():
    """A rebuild period changes the physical numbering only: the same live
    states give the same counts; the last op triggers no period rebuild; a
    negative period is an argument error."""
    rng = np.random.default_rng(8)
    x = rng.random((3000, 6)).astype(np.float32)
    fresh = rng.random((500, 6)).astype(np.float32)
    kinds = rng.choice(3, size=1200, p=[0.3, 0.2, 0.5]).astype(np.int8)
    ops = oracle.update_stream(3000, len(fresh), kinds, rng.random(1200), 10_000, period=100)
    od = oracle.OracleData(datagen.Dataset(1, vectors=x))
    want = oracle.update_counts(od, ops[:, 0], ops[:, 1], 0.8, cache_cap=10_000, fresh=fresh, period=100)
    idx = GTSIndex(datagen.Dataset(1, vectors=x))
    got = idx.update(ops, 0.8, cache_cap=10_000, fresh=fresh, period=100)
    assert list(got) == list(want)
    assert idx.stats()["retries"] == (len(ops) - 1) // 100   # rebuilds: one per period before the last op
    with pytest.raises(GTSError):
        idx.update(ops[:5], 0.8, cache_cap=10_000, fresh=fresh, period=-1)


def test_rebuild_invalidates_the_radius_network():
    """After an update stream rebuilds the tree, the network's pivots no longer
    exist: learned queries fail until it is retrained, then are exact again."""
    rng = np.random.default_rng(4)
    x = rng.random((4000, 8)).astype(np.float32)
    ds = datagen.Dataset(2, vectors=x)
    idx = GTSIndex(ds)
    idx.learned_train(rng.integers(0, 4000, 500).astype(np.int32), kmax=8, epochs=10)
    idx.update(np.array([(0, 5)] * 10, np.int32), 0.1, cache_cap=10)   # 10 inserts: one rebuild
    with pytest.raises(GTSError):
        idx.knn_query(np.array([1, 2], np.int32), 8, learned=True)
    idx.learned_train(rng.integers(0, 4010, 500).astype(np.int32), kmax=8, epochs=10)
    q = np.array([1, 2, 4005], np.int32)
    ids, _, _ = idx.knn_query(q, 8, learned=True)
    x2 = np.concatenate([x, np.repeat(x[5:6], 10, 0)])
    bids, _ = oracle.bf_knn(oracle.OracleData(datagen.Dataset(2, vectors=x2)), q, 8)
    assert (ids == bids).all()


def test_update_stream_strings():
    ds, _ = datagen.strings_edit(n=3000, roots=100, nq=10, seed=3)
    od = oracle.OracleData(ds)
    rng = np.random.default_rng(11)
    ops = _random_ops(rng, 3000, 400)
    want = oracle.update_counts(od, ops[:, 0], ops[:, 1], 2)
    got = GTSIndex(ds).update(ops, 2)
    assert list(got) == list(want)
