"""Deterministic synthetic workloads for the GTS pivot-tree index.

Every generator draws from ``numpy.random.default_rng(seed)`` (PCG64) in a fixed
order, so the same seed gives the same bytes on any machine.  The objects are
written in the reference's text data format (``dim N metric`` header, then one
row of floats per line, or one raw string per line for metric 6; the loader is
``load`` in GTS/include/file.cuh:28-105) and queries are object IDs
(``loadQuery``, GTS/include/file.cuh:112-151).

Workloads (sizes are the defaults; every generator takes smaller sizes for tests):

* ``gaussian_mixture``  -- clustered FP32 vectors, rows L2-normalised
  (20,000 x 300 with 50 clusters; 1,000,000 x 300 with 1,000 clusters).
* ``strings_edit``       -- lowercase strings made by 1-3 random edits of Zipf
  root words (edit-distance metric 6).
* ``points2d_pareto``    -- 2-d points from a Pareto-sized Gaussian mixture on the
  unit square.
* ``uniform_l1``         -- uniform [0,1)^dim vectors under L1 plus an update stream.
"""
from __future__ import annotations

import hashlib
import numpy as np

METRIC_LINF, METRIC_L1, METRIC_L2, METRIC_ANGULAR, METRIC_EDIT = 0, 1, 2, 5, 6


class Dataset:
    """Objects of one metric space: FP32 rows or strings."""

    def __init__(self, metric, vectors=None, strings=None):
        self.metric = int(metric)
        self.vectors = vectors          # np.ndarray float32 [n, dim] or None
        self.strings = strings          # list[bytes] or None

    @property
    def n(self):
        return len(self.strings) if self.strings is not None else self.vectors.shape[0]

    @property
    def dim(self):
        if self.strings is not None:
            return max(len(s) for s in self.strings) if self.strings else 0
        return self.vectors.shape[1]

    def string_matrix(self, stride=None):
        """Strings packed at a fixed stride (uint8 [n, stride]) plus lengths."""
        assert self.strings is not None
        if stride is None:
            stride = max(1, max(len(s) for s in self.strings))
        mat = np.zeros((len(self.strings), stride), dtype=np.uint8)
        lens = np.zeros(len(self.strings), dtype=np.int32)
        for i, s in enumerate(self.strings):
            mat[i, :len(s)] = np.frombuffer(s, dtype=np.uint8)
            lens[i] = len(s)
        return mat, lens

    def sha256(self):
        h = hashlib.sha256()
        if self.strings is not None:
            for s in self.strings:
                h.update(s + b"\n")
        else:
            h.update(np.ascontiguousarray(self.vectors).tobytes())
        return h.hexdigest()


def _pairwise_quantile(ds, rng, npairs, q):
    """Quantile ``q`` of FP64 distances over ``npairs`` random pairs i != j."""
    n = ds.n
    i = rng.integers(0, n, size=npairs)
    j = rng.integers(0, n - 1, size=npairs)
    j = j + (j >= i)                      # uniform over j != i
    if ds.metric == METRIC_EDIT:
        raise ValueError("edit-distance workloads use integer radii 1..3")
    else:
        a = ds.vectors[i].astype(np.float64)
        b = ds.vectors[j].astype(np.float64)
        diff = a - b
        if ds.metric == METRIC_L2:
            d = np.sqrt((diff * diff).sum(1))
        elif ds.metric == METRIC_L1:
            d = np.abs(diff).sum(1)
        elif ds.metric == METRIC_LINF:
            d = np.abs(diff).max(1)
        else:
            raise ValueError("quantile radius not defined for metric %d" % ds.metric)
    return float(np.float32(np.quantile(d, q)))


def gaussian_mixture(n=20000, dim=300, clusters=50, sigma=0.3, nq=256, seed=0,
                     npairs=10000, selectivity=0.01, chunk=65536):
    """Clustered, L2-normalised FP32 rows (metric L2).

    Draw order: centres C[clusters, dim] ~ N(0,1); labels ~ U{0..clusters-1};
    noise N(0, I) row-major; query IDs ~ U[0, n); then ``npairs`` pairs for the
    range radius (the ``selectivity`` quantile of their FP64 L2 distances).
    Returns (Dataset, qids int32[nq], radius float32).
    """
    rng = np.random.default_rng(seed)
    centres = rng.standard_normal((clusters, dim))
    labels = rng.integers(0, clusters, size=n)
    x = np.empty((n, dim), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        rows = centres[labels[s:e]] + sigma * rng.standard_normal((e - s, dim))
        rows /= np.linalg.norm(rows, axis=1, keepdims=True)
        x[s:e] = rows.astype(np.float32)
    qids = rng.integers(0, n, size=nq).astype(np.int32)
    ds = Dataset(METRIC_L2, vectors=x)
    radius = _pairwise_quantile(ds, rng, npairs, selectivity)
    return ds, qids, np.float32(radius)


def _zipf_letters(rng, size, s=1.1):
    ranks = np.arange(1, 27, dtype=np.float64)
    p = ranks ** (-s)
    p /= p.sum()
    return rng.choice(26, size=size, p=p)


def strings_edit(n=1000000, roots=20000, nq=10000, seed=0, minlen=4, maxlen=20):
    """Strings from random edits of Zipf(1.1) root words (metric 6, edit distance).

    Draw order: root lengths ~ U{minlen..maxlen}; root letters (Zipf over a-z);
    then per string: root index, number of edits U{1,2,3}, and per edit an op
    (insert/delete/substitute), a position and a letter; query IDs ~ U[0, n).
    Lengths stay within [minlen, maxlen]: an edit that would leave the range is
    replaced by a substitution.
    """
    rng = np.random.default_rng(seed)
    rlen = rng.integers(minlen, maxlen + 1, size=roots)
    letters = _zipf_letters(rng, int(rlen.sum()))
    alphabet = np.frombuffer(b"abcdefghijklmnopqrstuvwxyz", dtype=np.uint8)
    root_words = []
    off = 0
    for L in rlen:
        root_words.append(alphabet[letters[off:off + L]].tobytes())
        off += L
    which = rng.integers(0, roots, size=n)
    nedit = rng.integers(1, 4, size=n)
    total = int(nedit.sum())
    ops = rng.integers(0, 3, size=total)
    posr = rng.random(total)
    lets = alphabet[_zipf_letters(rng, total)]
    out = []
    t = 0
    for i in range(n):
        w = bytearray(root_words[which[i]])
        for _ in range(nedit[i]):
            op = ops[t]
            if op == 0 and len(w) >= maxlen:
                op = 2
            if op == 1 and len(w) <= minlen:
                op = 2
            if op == 0:     # insert
                p = int(posr[t] * (len(w) + 1))
                w.insert(p, lets[t])
            elif op == 1:   # delete
                p = int(posr[t] * len(w))
                del w[p]
            else:           # substitute
                p = int(posr[t] * len(w))
                w[p] = lets[t]
            t += 1
        out.append(bytes(w))
    qids = rng.integers(0, n, size=nq).astype(np.int32)
    return Dataset(METRIC_EDIT, strings=out), qids


def points2d_pareto(n=10000000, clusters=500, alpha=1.5, nq=100000, ntrain=50000,
                    kmax=64, seed=0):
    """2-d points: Pareto(alpha) cluster sizes, centres U[0,1]^2, sigma U[0.001,0.05].

    Points outside the unit square are redrawn.  Returns (Dataset, qids, ks,
    train_qids) with per-query k ~ U{1..kmax}.
    """
    rng = np.random.default_rng(seed)
    w = rng.pareto(alpha, size=clusters) + 1.0
    raw = w / w.sum() * n
    sizes = np.floor(raw).astype(np.int64)
    rem = n - int(sizes.sum())
    order = np.argsort(-(raw - sizes), kind="stable")
    sizes[order[:rem]] += 1
    centres = rng.random((clusters, 2))
    sig = rng.uniform(0.001, 0.05, size=clusters)
    pts = np.empty((n, 2), dtype=np.float32)
    off = 0
    for c in range(clusters):
        m = int(sizes[c])
        got = np.empty((0, 2))
        while got.shape[0] < m:
            cand = centres[c] + sig[c] * rng.standard_normal((m - got.shape[0], 2))
            ok = (cand >= 0).all(1) & (cand <= 1).all(1)
            got = np.concatenate([got, cand[ok]], 0)
        pts[off:off + m] = got[:m].astype(np.float32)
        off += m
    qids = rng.integers(0, n, size=nq).astype(np.int32)
    ks = rng.integers(1, kmax + 1, size=nq).astype(np.int32)
    train = rng.integers(0, n, size=ntrain).astype(np.int32)
    return Dataset(METRIC_L2, vectors=pts), qids, ks, train


def uniform_l1(n0=1000000, dim=64, nops=1000000, p=(0.4, 0.2, 0.4), seed=0,
               npairs=10000, selectivity=0.001):
    """Uniform [0,1)^dim rows under L1 plus an op stream (insert/delete/query).

    Fresh rows for inserts are appended after row ``n0``: ``vectors`` holds
    n0 + (#inserts) rows, and insert op ``i`` carries the row index ``n0 + i``.
    Deletes name a uniform live logical ID and queries a uniform live object;
    both are drawn by the consumer (``update_stream``) from ``op_u`` so that the
    stream depends only on the seed.
    Returns (Dataset, ops int8[nops], op_u float64[nops], radius).
    """
    rng = np.random.default_rng(seed)
    ops = rng.choice(3, size=nops, p=list(p)).astype(np.int8)
    nins = int((ops == 0).sum())
    x = rng.random((n0 + nins, dim), dtype=np.float64).astype(np.float32)
    op_u = rng.random(nops)
    ds0 = Dataset(METRIC_L1, vectors=x[:n0])
    radius = _pairwise_quantile(ds0, rng, npairs, selectivity)
    return Dataset(METRIC_L1, vectors=x), ops, op_u, np.float32(radius)


# ---------------------------------------------------------------- text formats
def write_data(path, ds):
    """Reference data format: ``dim N metric`` then rows (file.cuh:28-105)."""
    with open(path, "w") as f:
        if ds.metric == METRIC_EDIT:
            f.write("%d %d %d\n" % (ds.dim, ds.n, ds.metric))
            for s in ds.strings:
                f.write(s.decode("ascii") + "\n")
        else:
            n, dim = ds.vectors.shape
            f.write("%d %d %d\n" % (dim, n, ds.metric))
            for row in ds.vectors:
                f.write(" ".join("%.6f" % v for v in row) + "\n")


def write_queries(path, qids):
    """Reference query format: count, then one ID per line (file.cuh:112-151)."""
    with open(path, "w") as f:
        f.write("%d\n" % len(qids))
        for q in qids:
            f.write("%d\n" % int(q))


def write_updates(path, ops):
    """Reference update format: count, then ``flag id`` per line (update.cuh:43-95)."""
    with open(path, "w") as f:
        f.write("%d\n" % len(ops))
        for flag, oid in ops:
            f.write("%d %d\n" % (int(flag), int(oid)))
