"""GPU parity of the angular metric (code 5, tree.cuh:106-127): the
reference's own sample through the library and the CLI, and seeded data
against the FP64 oracle.

Angles are computed in FP32 on the GPU with a rigorous error bound (the cosine
errs by at most (4 dim + 32) 2^-24 and acos is 1/2-Hoelder with constant
pi / sqrt 2); pairs inside the band are re-decided in FP64, so result sets are
exact.  Distances: within 1e-5 relative (1e-5 absolute below 1).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from gts_amd import VERIFY_FP32, GTSIndex, datagen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RTOL = 1e-5


def _close(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    return np.all(np.abs(got - want) <= RTOL * np.maximum(np.abs(want), 1.0))


def _sample(golden_dir):
    X = np.load(os.path.join(golden_dir, "sbw_sample.npy"))
    kat = json.load(open(os.path.join(golden_dir, "sbw_sample_kat.json")))
    return X, kat


def test_reference_sample_knn8_and_update_stream(golden_dir):
    """README.md:52 (``0 8``) and README.md:79 (``2 59``) on the shipped sample."""
    X, kat = _sample(golden_dir)
    idx = GTSIndex(datagen.Dataset(5, vectors=X))
    qs = np.array(sorted(int(q) for q in kat["knn8"]), np.int32)
    ids, dists, kth = idx.knn_query(qs, 8)
    for row, d, q in zip(ids, dists, qs):
        want_ids, want_d = kat["knn8"][str(q)]
        assert list(row) == want_ids
        assert _close(d, want_d)
    counts = idx.update(np.array(kat["update_ops"], np.int32), kat["update_r"])
    assert list(counts) == kat["update_counts"]


def _write_text(path, X, metric):
    with open(path, "w") as f:
        f.write("%d %d %d\n" % (X.shape[1], X.shape[0], metric))
        for row in X:
            f.write(" ".join("%.9g" % v for v in row) + "\n")


def _trailer(lines, names):
    """The result file's trailing lines: each "<name><sep><float>" in order,
    with the reference's exact labels (main.cu:89-164)."""
    assert len(lines) == len(names), lines
    for line, name in zip(lines, names):
        assert line.startswith(name), (line, name)
        float(line[len(name):])


def test_cli_reference_sample(golden_dir, tmp_path):
    """The CLI with the README's argv on the sample: every line of the result
    files (main.cu:89-164), and the saveK / saveR outputs (file.cuh:154-199)
    against the known answers and brute force."""
    X, kat = _sample(golden_dir)
    data = tmp_path / "sbw_example.txt"
    _write_text(data, X, 5)
    qf = tmp_path / "qid.txt"
    qf.write_text("%d\n" % len(kat["qids"]) + "".join("%d\n" % q for q in kat["qids"]))
    uf = tmp_path / "u.txt"
    uf.write_text("%d\n" % len(kat["update_ops"]) + "".join("%d %d\n" % tuple(o) for o in kat["update_ops"]))
    exe = os.path.join(ROOT, "gts_amd", "bin", "gts")
    out, fid, fd = tmp_path / "knn.txt", tmp_path / "knn_ids.txt", tmp_path / "knn_dis.txt"
    subprocess.run([exe, str(data), str(qf), "0", "8", str(out), "--ids", str(fid), "--dists", str(fd)],
                   check=True, capture_output=True, timeout=120)
    lines = out.read_text().splitlines()
    assert lines[0] == "Knn search num: 8"
    assert lines[1] == "Result radius: "
    kth = [float(v) for v in lines[2].split()]
    want = [kat["knn8"][str(q)][1][-1] for q in kat["qids"]]
    assert len(kth) == len(want)
    assert np.all(np.abs(np.array(kth) - np.array(want)) <= 1e-5 * np.array(want) + 1e-6)  # "%f" keeps 6 decimals
    _trailer(lines[3:], ["Time of index construction: ", "Search time: "])
    # saveK: one line per query, k IDs / k "%f" distances, each followed by a space
    id_lines, d_lines = fid.read_text().split("\n"), fd.read_text().split("\n")
    assert id_lines[-1] == "" and d_lines[-1] == "" and len(id_lines) == len(kat["qids"]) + 1
    for q, li, ld in zip(kat["qids"], id_lines, d_lines):
        ids_k, d_k = kat["knn8"][str(q)]
        assert li == "".join("%d " % v for v in ids_k)
        got = np.array([float(v) for v in ld.split()])
        assert ld.endswith(" ") and np.all(np.abs(got - np.array(d_k)) <= 1e-5 * np.array(d_k) + 1e-6)
    # range with saveR on the same sample, against brute force
    od = oracle.OracleData(datagen.Dataset(5, vectors=X))
    r = float(np.median(kth))
    out, fid, fd = tmp_path / "rng.txt", tmp_path / "rng_ids.txt", tmp_path / "rng_dis.txt"
    fid.write_text("")   # saveR appends
    subprocess.run([exe, str(data), str(qf), "1", "%.6f" % r, str(out), "--ids", str(fid), "--dists", str(fd)],
                   check=True, capture_output=True, timeout=120)
    lines = out.read_text().splitlines()
    assert lines[0] == "Range search radius: %f" % np.float32("%.6f" % r)
    assert lines[1] == "Result num: "
    q = np.array(kat["qids"], np.int32)
    bc, boffs, bids, bd = oracle.bf_range(od, q, np.float32("%.6f" % r), with_ids=True)
    assert [int(v) for v in lines[2].split()] == list(bc)
    _trailer(lines[3:], ["Time of index construction: ", "Search time: "])
    id_lines = fid.read_text().split("\n")[:-1]
    for i, li in enumerate(id_lines):
        assert li == "".join("%d " % v for v in bids[boffs[i]:boffs[i + 1]])
    # the update stream: counts and the reference's four trailer lines
    out = tmp_path / "upd.txt"
    subprocess.run([exe, str(data), str(uf), "2", "59", str(out)], check=True, capture_output=True, timeout=120)
    lines = out.read_text().splitlines()
    assert lines[0] == "Range search radius (check for updates): 59.000000"
    assert lines[1] == "Result num: "
    assert [int(v) for v in lines[2].split()] == kat["update_counts"]
    _trailer(lines[3:], ["Time of index construction: ", "Total update time: ", "Search time in update : ",
                         "Update time in update: "])


def _angular_ds(n, dim, seed, dup=0):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((40, dim))
    x = c[rng.integers(0, 40, n)] + 0.4 * rng.standard_normal((n, dim))
    x *= rng.uniform(0.1, 10.0, (n, 1))            # angles ignore the scale
    if dup:
        # near-duplicates (angles ~1e-3 degrees) and exact scaled copies exercise the band near 0
        x[1:dup] = x[0] * rng.uniform(0.5, 2.0, (dup - 1, 1)) + 1e-5 * rng.standard_normal((dup - 1, dim))
        x[dup] = 3.0 * x[0]
    return datagen.Dataset(5, vectors=x.astype(np.float32))


@pytest.mark.parametrize("n,dim", [(20000, 64), (5000, 300), (3000, 7)])
def test_angular_range_and_knn_match_oracle(n, dim):
    ds = _angular_ds(n, dim, seed=dim, dup=30)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(1)
    qids = np.concatenate([[0, 1, 30], rng.integers(0, n, 197)]).astype(np.int32)
    # radius at ~1% selectivity from sampled pairs (FP64 angles)
    i, j = rng.integers(0, n, 4000), rng.integers(0, n, 4000)
    xa, xb = ds.vectors[i].astype(np.float64), ds.vectors[j].astype(np.float64)
    cos = (xa * xb).sum(1) / np.linalg.norm(xa, axis=1) / np.linalg.norm(xb, axis=1)
    r = float(np.float32(np.quantile(np.degrees(np.arccos(np.clip(cos, -1, 1))) * np.pi / 3.1415926, 0.01)))
    counts, offs, ids, dists = idx.range_query(qids, r, with_ids=True)
    bc, boffs, bids, bd = oracle.bf_range(od, qids, r, with_ids=True)
    assert (counts == bc).all()
    assert (ids == bids).all()
    assert _close(dists, bd)
    for k in (1, 10, 40):
        ids, dists, kth = idx.knn_query(qids, k)
        bids, bd = oracle.bf_knn(od, qids, k)
        assert (ids == bids).all()
        assert _close(dists, bd)
    # tiny radius: only the query itself and its near-copies
    counts = idx.range_query(qids[:3], 0.01)
    assert (counts == oracle.bf_range(od, qids[:3], 0.01)).all()


def test_angular_rejects_zero_vectors():
    x = np.ones((100, 8), np.float32)
    x[17] = 0.0
    from gts_amd import GTSError
    with pytest.raises(GTSError):
        GTSIndex(datagen.Dataset(5, vectors=x))


def test_angular_learned_radius_is_exact():
    ds = _angular_ds(30000, 32, seed=9)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds)
    rng = np.random.default_rng(3)
    idx.learned_train(rng.integers(0, 30000, 2000).astype(np.int32), kmax=16, epochs=40)
    qids = rng.integers(0, 30000, 300).astype(np.int32)
    ids, dists, _ = idx.knn_query(qids, 16, learned=True, verify=VERIFY_FP32)
    bids, bd = oracle.bf_knn(od, qids, 16)
    assert (ids == bids).all()
    assert _close(dists, bd)
