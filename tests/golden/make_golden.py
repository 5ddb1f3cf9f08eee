"""Writes the known-answer fixtures for the reference's shipped sample.

The reference ships three input files (``Datasets Sample/``) and no expected
outputs.  This script computes, with numpy in FP64 (independent of the C
oracle), the answers to the README's own invocations on them:

* ``GTS sbw_example.txt sbw_example_qid.txt 0 8``  (README.md:52): kNN, k = 8,
  angular metric 5 -> per distinct query ID, the 8 nearest (ID, degrees);
* ``GTS sbw_example.txt sbw_example_u.txt 2 59``   (README.md:79): the update
  stream with r = 59 degrees -> one count per query op.

It also stores the sample's rows (``sbw_sample.npy``, FP32 as parsed) and a
40-row slice of them as small, self-contained inputs, so the fixtures travel
without the reference.  Run from the repo root:
``python tests/golden/make_golden.py``.
"""
import json
import os

import numpy as np

SAMPLE = "/root/reference/Datasets Sample"
OUT = os.path.dirname(os.path.abspath(__file__))


def load_rows(path):
    with open(path) as f:
        dim, n, metric = (int(v) for v in f.readline().split())
        rows = np.array([[float(v) for v in line.split()] for line in f if line.strip()],
                        dtype=np.float32)
    assert rows.shape == (n, dim)
    return rows, metric


def angular(q, X):
    """Degrees, with the reference's constant 3.1415926 (tree.cuh:126)."""
    q = q.astype(np.float64)
    X = X.astype(np.float64)
    c = (X @ q) / (np.linalg.norm(X, axis=1) * np.linalg.norm(q))
    c = np.clip(c, -1.0, 1.0)
    return np.abs(np.arccos(c) * 180.0 / 3.1415926)


def knn(X, qid, k):
    d = angular(X[qid], X)
    d[qid] = 0.0
    order = np.lexsort((np.arange(len(d)), d))[:k]
    return [int(i) for i in order], [float(d[i]) for i in order]


def update_counts(X, ops, r, cache_cap=10):
    """Plain-Python replay of the op stream over physical rows (update.cuh:449-638)."""
    rows = list(range(len(X)))
    dead = [False] * len(rows)
    cache = []
    out = []
    for flag, oid in ops:
        if flag == 0:
            cache.append(rows[oid])
            if len(cache) == cache_cap:
                rows = [rows[i] for i in range(len(rows)) if not dead[i]] + cache
                dead = [False] * len(rows)
                cache = []
        elif flag == 1:
            live = [i for i in range(len(rows)) if not dead[i]]
            if oid < len(live):
                dead[live[oid]] = True
            else:
                del cache[oid - len(live)]
        else:
            q = X[rows[oid]]
            cnt = 0
            for i, h in enumerate(rows):
                if dead[i]:
                    continue
                d = 0.0 if i == oid else angular(q, X[h:h + 1])[0]
                cnt += d <= r
            for h in cache:
                cnt += angular(q, X[h:h + 1])[0] <= r
            out.append(int(cnt))
    return out


def main():
    X, metric = load_rows(os.path.join(SAMPLE, "sbw_example.txt"))
    with open(os.path.join(SAMPLE, "sbw_example_qid.txt")) as f:
        nq = int(f.readline())
        qids = [int(f.readline()) for _ in range(nq)]
    with open(os.path.join(SAMPLE, "sbw_example_u.txt")) as f:
        nu = int(f.readline())
        ops = [tuple(int(v) for v in f.readline().split()) for _ in range(nu)]
    distinct = sorted(set(qids))
    res = {"source": "Datasets Sample/sbw_example*.txt", "metric": metric,
           "n": int(X.shape[0]), "dim": int(X.shape[1]), "qids": qids,
           "knn8": {str(q): knn(X, q, 8) for q in distinct},
           "update_ops": ops, "update_r": 59.0,
           "update_counts": update_counts(X, ops, np.float32(59.0))}
    with open(os.path.join(OUT, "sbw_sample_kat.json"), "w") as f:
        json.dump(res, f, indent=1)
    # the whole sample (FP32 as parsed from the text file) so the GPU tests
    # and the CLI test can run the README's invocations without the reference
    np.save(os.path.join(OUT, "sbw_sample.npy"), X)
    # a 40-row slice (rows 0..39) of the sample as a self-contained fixture
    sl = X[:40]
    np.save(os.path.join(OUT, "sbw_slice40.npy"), sl)
    sl_res = {"knn5": {str(q): knn(sl, q, 5) for q in range(0, 40, 7)}}
    with open(os.path.join(OUT, "sbw_slice40_kat.json"), "w") as f:
        json.dump(sl_res, f, indent=1)
    print("wrote fixtures for", len(distinct), "distinct queries;",
          "update counts", res["update_counts"])


if __name__ == "__main__":
    main()
