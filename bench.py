#!/usr/bin/env python3
"""Benchmark: exact kNN queries/sec on one MI355X.

Workload (one "step"): 10,000 concurrent kNN queries (k = 16) over 1,000,000
FP32 vectors of dimension 300 (1,000-cluster Gaussian mixture, sigma 0.3, rows
L2-normalised, seed 0), L2 metric, through the pivot tree (T = 10, leaf
capacity 20).  Query IDs and result buffers are resident in device memory when
the timed region starts; index construction, the cell partition and
radius-network training are outside it (their times are reported), as the
reference excludes index construction (main.cu:94-98).

Forms timed on the same index and the same queries (each must return the
plain form's ID sets):
  * plain              - the reference algorithm: level-by-level ring pruning
                         with the k-th pivot distance bound, FP32
                         verification, exact top-k;
  * learned (headline) - the radius network predicts each query's k-th NN
                         distance, one range pass with that radius collects
                         candidates (BF16 hi+lo MFMA verification with an
                         FP32/FP64 re-check band), exact top-k, radius
                         doubling for any query with fewer than k sure
                         results.  With the automatic plan the pass runs as
                         the cell scan: pivot bounds over a Voronoi partition
                         of the rows choose the (64-query block, 32-row
                         chunk) pairs the MFMA kernel verifies;
  * learned_tree       - the network's radius over the GTS tree (ring pruning
                         with that radius, BF16X3 verification of the leaves);
  * learned_exhaustive - the network's radius, every row verified (dense
                         BF16X3 kernel, no pruning);
  * cells_no_network   - the cell scan driven by the k-th distance to one tree
                         level's pivots (the plain bound) instead of the
                         network;
  * host               - the headline form through the host-buffer entry
                         point: query IDs from host memory, IDs, distances and
                         k-th distances back into host memory.
The headline form is also run at k = 8 and k = 32 (a network trained per k).
Each form's timed steps run as min(5, steps) groups; ms_per_step is the median
of the group means.  The JSON line reports the headline form's throughput as
``value``.
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks (dense)
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(headline, kernel, args):
    """HBM bytes per launch of the headline verification kernel, from the
    calibrated FETCH_SIZE / WRITE_SIZE counter runs summarised in
    profiles/<kernel>_traffic.json (null when absent or for another workload)."""
    path = os.path.join(ROOT, "profiles", "%s_traffic.json" % kernel)
    if headline != "learned" or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    w = t.get("workload", {})
    if (w.get("n"), w.get("dim"), w.get("nq"), w.get("k")) != (args.n, args.dim, args.nq, args.k):
        return None
    return t.get("traffic_bytes_per_launch")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # three untimed calls: the first tunes and plans, the second captures the
    # batch's first pass into a graph, the third replays it
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=300)
    ap.add_argument("--clusters", type=int, default=1000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--extra-k", default="8,32", help="further k for the headline form (comma list; '' for none)")
    ap.add_argument("--mode", choices=["plain", "learned"], default=None,
                    help="headline form (default: learned when available)")
    ap.add_argument("--train", type=int, default=4000, help="training queries for the radius network")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--cpu-sample", type=int, default=400,
                    help="queries in the CPU baseline sample (about 10 s on 16 host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", type=int, default=64, help="queries checked against the oracle")
    ap.add_argument("--forms", default="plain,learned,learned_tree,learned_exhaustive,cells_no_network,host",
                    help="forms to time (comma list)")
    ap.add_argument("--plan", choices=["auto", "cells", "exhaustive", "pruned"], default="auto",
                    help="plan of the learned form (results are identical under every plan)")
    ap.add_argument("--side-ms", type=float, default=1500.0,
                    help="timed budget of each non-headline form (its step count is sized to it)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="run one device-to-device copy of this many MiB (calibrates HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import gts_amd
    from gts_amd import DeviceArray, GTSIndex, VERIFY_BF16X3, VERIFY_FP32, datagen

    plans = {"auto": gts_amd.PLAN_AUTO, "cells": gts_amd.PLAN_CELLS, "exhaustive": gts_amd.PLAN_EXHAUSTIVE,
             "pruned": gts_amd.PLAN_PRUNED}
    t = time.time()
    ds, qids, _ = datagen.gaussian_mixture(n=args.n, dim=args.dim, clusters=args.clusters,
                                           nq=args.nq, seed=0)
    log("data generated in %.1f s" % (time.time() - t))
    # the objects handed to the library (the reference's load, file.cuh:28),
    # then the index built from them (indexConstru, tree.cuh:332)
    t = time.perf_counter()
    nds = gts_amd.NativeDataset.from_python(ds)
    setup = {"dataset_load_s": round(time.perf_counter() - t, 3)}
    t = time.perf_counter()
    idx = GTSIndex(nds, 10, 20, 0)
    setup["index_build_s"] = round(time.perf_counter() - t, 3)
    log("index built in %.2f s (objects loaded in %.2f s): %s" % (setup["index_build_s"], setup["dataset_load_s"],
                                                                  idx.info()))
    nq = args.nq
    qh = np.ascontiguousarray(qids, np.int32)
    # query IDs and result buffers resident in device memory (HBM)
    d_q = DeviceArray.from_host(idx, qh)
    bufs = {}

    def outputs(k):
        if k not in bufs:
            bufs[k] = (idx.device_array((nq, k), np.int32), idx.device_array((nq, k), np.float32),
                       idx.device_array(nq, np.float32))
        return bufs[k]

    def train(k):
        rng = np.random.default_rng(1)
        tr = rng.integers(0, args.n, size=args.train).astype(np.int32)
        t0 = time.perf_counter()
        loss = idx.learned_train(tr, kmax=k, epochs=args.epochs)
        dt = time.perf_counter() - t0
        log("radius network (kmax %d) trained in %.2f s, loss %.4f" % (k, dt, loss))
        return dt

    have_learned = args.mode != "plain"
    if have_learned:
        try:
            setup["train_s"] = round(train(args.k), 3)
            setup["train_queries"] = args.train
            setup["train_epochs"] = args.epochs
        except gts_amd.GTSError as e:
            if args.mode == "learned":
                raise
            log("learned form unavailable:", e)
            have_learned = False
    headline = args.mode or ("learned" if have_learned else "plain")

    # form -> (learned, verify, plan, radius source, host buffers)
    forms_def = {
        "plain": (False, VERIFY_FP32, None, None, False),
        "learned": (True, VERIFY_BF16X3, plans[args.plan], gts_amd.RADIUS_NETWORK, False),
        "learned_tree": (True, VERIFY_BF16X3, gts_amd.PLAN_PRUNED, gts_amd.RADIUS_NETWORK, False),
        "learned_exhaustive": (True, VERIFY_BF16X3, gts_amd.PLAN_EXHAUSTIVE, gts_amd.RADIUS_NETWORK, False),
        "cells_no_network": (True, VERIFY_BF16X3, gts_amd.PLAN_CELLS, gts_amd.RADIUS_PIVOTS, False),
        "host": (True, VERIFY_BF16X3, plans[args.plan], gts_amd.RADIUS_NETWORK, True),
    }
    L = gts_amd.lib()
    host_out = {}

    def ptr(a):
        return a.ctypes.data_as(ctypes.c_void_p)

    def runner(name, k):
        learned, verify, plan, src, host = forms_def[name]
        if learned:
            idx.set_knn_plan(plan)
            idx.set_radius_source(src)
        if host:
            if k not in host_out:
                host_out[k] = (np.zeros((nq, k), np.int32), np.zeros((nq, k), np.float32),
                               np.zeros(nq, np.float32))
            hi, hd, hk = host_out[k]

            def run():
                gts_amd._check(L.gts_knn_query_ex(idx.h, ptr(qh), nq, None, k, k, ptr(hi), ptr(hd), ptr(hk),
                                                  int(learned), int(verify)))
            return run, hi.copy
        ids, dists, kth = outputs(k)

        def run():
            idx.knn_query_dev(d_q, k, ids, dists, kth, learned=learned, verify=verify)
        return run, ids.to_host

    stat_keys = ("ms_verify", "verify_launches", "verified_pairs", "candidates", "retries", "band_pairs",
                 "graph_replays", "live_chunks", "ms_verify_batch", "verify_launches_batch", "live_chunks_batch")

    def timed(name, k, steps=None):
        run, fetch = runner(name, k)
        # the timed calls record no HIP events (instrumentation only: the
        # work and results are the same); the statistics calls do
        idx.set_timing(False)
        first = None
        last_ms = 0.0
        for i in range(max(1, args.warmup)):
            t0 = time.perf_counter()
            run()
            last_ms = (time.perf_counter() - t0) * 1000.0
            if i == 0:
                first = idx.stats()
        if steps is None:
            steps = max(2, min(args.steps, int(args.side_ms / max(last_ms, 1e-3))))
        ngroups = min(5, steps)
        sizes = [steps // ngroups + (1 if g < steps % ngroups else 0) for g in range(ngroups)]
        groups = []
        for sz in sizes:
            # every call returns after its results are complete (device or host)
            t0 = time.perf_counter()
            for _ in range(sz):
                run()
            groups.append((time.perf_counter() - t0) * 1000.0 / sz)
        ms = statistics.median(groups)
        ids = fetch()
        # per-step statistics (pairs, launches, verification time on HIP
        # events) from further calls outside the timed region, after two calls
        # that capture the timed setting's graph
        idx.set_timing(True)
        for _ in range(2):
            run()
        st = {key: 0.0 for key in stat_keys}
        for _ in range(steps):
            run()
            s = idx.stats()
            for key in st:
                st[key] += s[key]
        st = {key: v / steps for key, v in st.items()}
        rec = {"qps": round(nq / (ms / 1000.0), 1), "ms_per_step": round(ms, 4), "steps": steps,
               "group_ms": [round(g, 4) for g in groups], "candidates_per_step": round(st["candidates"], 1),
               "verified_pairs_per_step": round(st["verified_pairs"], 1), "retries_per_step": st["retries"]}
        log("%s k=%d: %.3f ms/step (median of %d groups), %.0f QPS, stats/step %s" %
            (name, k, ms, ngroups, rec["qps"], st))
        return rec, st, ids, first

    if args.calib_copy_mb > 0:
        nb = args.calib_copy_mb << 20
        a = idx.device_array(nb // 4, np.float32)
        b = idx.device_array(nb // 4, np.float32)
        b.copy_from(a)
        log("calibration copy: %d bytes device to device" % nb)
        a.free()
        b.free()

    want = [f.strip() for f in args.forms.split(",") if f.strip()]
    forms = [f for f in forms_def if f in want and (f == "plain" or have_learned)]
    if headline not in forms:
        forms.insert(0, headline)
    if "plain" not in forms:
        forms.insert(0, "plain")   # the exactness reference
    res, ids_of = {}, {}
    head_st, s_last = None, None
    for f in forms:
        rec, st, ids, first = timed(f, args.k, args.steps if f == headline else None)
        res[f] = rec
        ids_of[f] = ids
        if f == headline:
            head_st = st
            s_last = idx.stats()
        if first is not None and first["ms_cells_build"] > 0:
            setup["cell_partition_s"] = round(first["ms_cells_build"] / 1000.0, 3)

    # exactness: every form's ID sets equal the plain form's
    exact = True
    for f in forms:
        if f != "plain":
            same = bool((ids_of[f] == ids_of["plain"]).all())
            res[f]["ids_equal_plain"] = same
            log("%s == plain ID sets: %s" % (f, same))
            exact &= same

    oracle_ok = None
    od = None
    threads = min(len(os.sched_getaffinity(0)), 16)
    sel = np.arange(0, nq, max(1, nq // max(args.check, 1)))[:args.check]
    if args.check > 0 or not args.no_cpu:
        import oracle
        od = oracle.OracleData(ds)
    if args.check > 0:
        bids, _ = oracle.bf_knn(od, qh[sel], args.k, threads=threads)
        oracle_ok = bool((ids_of["plain"][sel] == bids).all())
        log("oracle check on %d queries: %s" % (len(sel), oracle_ok))
        exact &= oracle_ok

    # the headline form at further k: a network trained per k, checked
    # against the plain form on every query and the oracle on the sample
    k_rows = {}
    if headline == "learned":
        for kx in [int(x) for x in args.extra_k.split(",") if x.strip()]:
            setup["train_s_k%d" % kx] = round(train(kx), 3)
            rec, _, ids, _ = timed("learned", kx)
            prec, _, pids, _ = timed("plain", kx, 2)
            same = bool((ids == pids).all())
            if args.check > 0:
                bids, _ = oracle.bf_knn(od, qh[sel], kx, threads=threads)
                same &= bool((pids[sel] == bids).all())
            rec["ids_equal_plain_and_oracle"] = same
            rec["plain_qps"] = prec["qps"]
            rec["speedup_vs_plain"] = round(prec["ms_per_step"] / rec["ms_per_step"], 2)
            k_rows["k%d" % kx] = rec
            exact &= same

    cpu = None
    if not args.no_cpu:
        t = time.time()
        otree = oracle.OracleTree(od, 10, 20, 0, threads=threads)
        log("oracle tree built in %.1f s" % (time.time() - t))
        sample = qh[:args.cpu_sample]
        t0 = time.perf_counter()
        otree.knn(sample, args.k, threads=threads)
        dt = time.perf_counter() - t0
        cpu = {"value": round(len(sample) / dt, 2), "unit": "queries/s", "cores": threads,
               "kind": "port",
               "sample": "first %d of the %d query IDs, k=%d, full %d x %d dataset, "
                         "OpenMP over queries, FP64 distances" % (len(sample), nq, args.k, args.n, args.dim),
               "nproc": os.cpu_count(), "cpus_available": len(os.sched_getaffinity(0)),
               "cpu_model": cpu_model(),
               "threads_note": "one thread per available CPU, at most 16 (the host share of one GPU)"}
        log("cpu baseline:", cpu)

    ms = res[headline]["ms_per_step"]
    st = head_st
    sched = ({"rows_per_range": int(s_last["dense_rows"]),
              "blocks": "persistent" if s_last["dense_persistent"] else "grid"}
             if headline == "learned" and s_last["dense_rows"] else None)
    value = nq / (ms / 1000.0)
    # dominant kernel: verification.  Plain and exhaustive forms: algorithmic
    # flops per verified pair (MFMA / FP32 bound).  Cell scan: the rows of
    # every verified (64-query block, 32-row chunk) pair, BF16 hi and FP8
    # residual planes, 32 x dim x 3 B (HBM bound: each chunk is read once per
    # query block that needs it).  The cell scan's dominant kernel is the
    # batch's launch (k_verify_cells_bf16<false>); the reruns of a few padded
    # query blocks run as their own instance (<true>, latency-bound), in
    # ms_per_step but not in the roofline.
    cells = headline == "learned" and s_last["cell_scan"] > 0
    if cells and st["verify_launches_batch"] > 0:
        launches = st["verify_launches_batch"]
        avg_ms = st["ms_verify_batch"] / launches
    else:
        launches = max(1, st["verify_launches"])
        avg_ms = st["ms_verify"] / launches
    pairs_per_launch = st["verified_pairs"] / max(1, st["verify_launches"])
    kernel = "cell_verify" if cells else "dense_verify"
    if cells:
        chunks_per_launch = st["live_chunks_batch"] / launches
        bound, unit, peak = "hbm", "GB/s", PEAK_HBM_GBS
        achieved = chunks_per_launch * 32 * args.dim * 3 / (avg_ms / 1000.0) / 1e9 if avg_ms > 0 else 0.0
    else:
        bound, unit = "mfma", "TFLOP/s"
        flops = (6.0 if headline == "learned" else 3.0) * args.dim * pairs_per_launch
        peak = PEAK_BF16_TFLOPS if headline == "learned" else PEAK_FP32_TFLOPS
        achieved = flops / (avg_ms / 1000.0) / 1e12 if avg_ms > 0 else 0.0
    traffic = measured_traffic(headline, kernel, args)
    qps = {f: res[f]["qps"] for f in res}
    line = {
        "metric": "exact queries/sec (1M×300-d L2, 10k kNN k=16): learned-radius vs plain, 1 GPU",
        "value": round(value, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3+fp32" if headline == "learned" else "fp32",
        "data": "synthetic (1,000-cluster Gaussian mixture, seed 0)",
        "config": {"workload": "kNN k=%d, %d queries, %d x %d FP32 L2, pivot tree T=10 leaf=20"
                               % (args.k, nq, args.n, args.dim),
                   "form": headline,
                   "plan": (("cell scan, %d cells" % s_last["cell_scan"]) if cells else args.plan)
                   if headline == "learned" else "tree (reference)",
                   "dense_schedule": sched,
                   "timing": "median of min(5, steps) group means; query IDs and results in device memory",
                   "graph_replays_per_step": st["graph_replays"] if headline == "learned" else None},
        "plain_qps": qps.get("plain"),
        "learned_qps": qps.get("learned"),
        "learned_tree_qps": qps.get("learned_tree"),
        "learned_exhaustive_qps": qps.get("learned_exhaustive"),
        "cells_no_network_qps": qps.get("cells_no_network"),
        "host_qps": qps.get("host"),
        "speedup_vs_plain": round(res["plain"]["ms_per_step"] / ms, 3),
        "forms": res,
        "k_rows": k_rows,
        "setup": setup,
        "exact": exact,
        "roofline": {"bound": bound, "achieved": round(achieved, 2), "peak": peak,
                     "unit": unit, "frac": round(achieved / peak, 4), "traffic": traffic,
                     "kernel": ("k_verify_cells_bf16<false> (cell scan batch launch, bf16x3 MFMA, FP8 residual rows)"
                                if cells else
                                "k_verify_dense_bf16 (bf16x3 MFMA)" if headline == "learned" else
                                "k_verify (fp32 tiles)"),
                     "avg_launch_ms": round(avg_ms, 4),
                     "launches_per_step": launches,
                     "pairs_per_launch": round(pairs_per_launch, 1),
                     "chunk_blocks_per_launch": round(chunks_per_launch, 1) if cells else None,
                     "rerun_launches_per_step": (st["verify_launches"] - st["verify_launches_batch"]) if cells else None,
                     "rerun_ms_per_step": round(st["ms_verify"] - st["ms_verify_batch"], 4) if cells else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
