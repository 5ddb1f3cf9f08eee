"""The configurations' workloads, run on one GPU with their timings printed.

A driver for profiling and inspection (``python -m gts_amd.workloads NAME``):

* ``build``   -- 1,000,000 x 300 FP32 L2 rows (1,000-cluster Gaussian
  mixture, seed 0): the dataset handed to the library, then the index built
  (indexConstru, tree.cuh:332-393) ``--reps`` times;
* ``c2plain`` -- the same rows, 10,000 concurrent kNN queries at k = 16
  through the plain two-stage search (searchIndexKnnV2, search_v2.cuh:975):
  level-by-level ring pruning, FP32 verification, exact top-k;
* ``c3``      -- 1,000,000 edit-distance strings (Zipf roots, 1-3 edits),
  10,000 concurrent range queries at r = 1, 2, 3 (searchIndexRnnV2,
  search_v2.cuh:780);
* ``c4``      -- 10,000,000 2-d points (Pareto cluster sizes), 100,000
  concurrent kNN queries with k ~ U{1..64} through the plain search.

``--calib-copy-mb N`` adds one device-to-device copy of N MiB (a kernel of
known traffic that calibrates the HBM counters of a profile).  Results are
not checked here (the GPU tests check every workload against the oracle).
"""
import argparse
import sys
import time

import numpy as np

from . import VERIFY_FP32, DeviceArray, GTSIndex, NativeDataset, datagen


def _log(*a):
    print(*a, flush=True)


def _calib(idx, mb):
    if mb <= 0:
        return
    nb = mb << 20
    a = idx.device_array(nb // 4, np.float32)
    b = idx.device_array(nb // 4, np.float32)
    b.copy_from(a)
    _log("calibration copy: %d bytes device to device" % nb)
    a.free()
    b.free()


def _build(ds, reps):
    t = time.perf_counter()
    nds = NativeDataset.from_python(ds)
    _log("objects loaded in %.3f s" % (time.perf_counter() - t))
    idx = None
    for _ in range(reps):
        if idx is not None:
            idx.close()
        t = time.perf_counter()
        idx = GTSIndex(nds, 10, 20, 0)
        _log("index built in %.3f s: %s" % (time.perf_counter() - t, idx.info()))
    return idx


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("name", choices=["build", "c2plain", "c3", "c4"])
    ap.add_argument("--reps", type=int, default=1, help="index builds (build) or calls per query batch")
    ap.add_argument("--calib-copy-mb", type=int, default=0)
    a = ap.parse_args(argv)
    if a.name in ("build", "c2plain"):
        ds, qids, _ = datagen.gaussian_mixture(n=1_000_000, dim=300, clusters=1000, nq=10_000, seed=0)
        idx = _build(ds, a.reps if a.name == "build" else 1)
        _calib(idx, a.calib_copy_mb)
        if a.name == "c2plain":
            for _ in range(a.reps):
                t = time.perf_counter()
                idx.knn_query(qids, 16, learned=False, verify=VERIFY_FP32)
                _log("plain kNN k=16, %d queries: %.3f s, stats %s" % (len(qids), time.perf_counter() - t,
                                                                        idx.stats()))
    elif a.name == "c3":
        ds, qids = datagen.strings_edit(n=1_000_000, roots=20_000, nq=10_000, seed=0)
        idx = _build(ds, 1)
        _calib(idx, a.calib_copy_mb)
        for r in (1, 2, 3):
            for _ in range(a.reps):
                t = time.perf_counter()
                idx.range_query(qids, r)
                _log("range r=%d, %d queries: %.3f s, stats %s" % (r, len(qids), time.perf_counter() - t,
                                                                  idx.stats()))
    else:
        ds, qids, ks, _ = datagen.points2d_pareto(n=10_000_000, clusters=500, nq=100_000, ntrain=50_000,
                                                  kmax=64, seed=0)
        idx = _build(ds, 1)
        _calib(idx, a.calib_copy_mb)
        q = DeviceArray.from_host(idx, qids)
        kd = DeviceArray.from_host(idx, ks)
        ids = idx.device_array((len(qids), 64), np.int32)
        dists = idx.device_array((len(qids), 64), np.float32)
        kth = idx.device_array(len(qids), np.float32)
        for _ in range(a.reps):
            t = time.perf_counter()
            idx.knn_query_dev(q, 0, ids, dists, kth, ks_dev=kd, learned=False, verify=VERIFY_FP32)
            _log("plain kNN k~U{1..64}, %d queries: %.3f s, stats %s" % (len(qids), time.perf_counter() - t,
                                                                         idx.stats()))
    return 0


if __name__ == "__main__":
    sys.exit(main())
