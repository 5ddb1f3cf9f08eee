"""GPU construction at sizes that reach every sort path of a level.

The distance keys of a level are FP32 (16-byte loads when dim % 4 == 0, else
4-byte loads); nodes of at most 16,384 objects are sorted by one workgroup in
LDS, larger ones by the global radix sort (keyed by the node's rank at the
level when every object is in a splitting node of it, else by the range
start).  Each tree is compared with the oracle's indexConstru restatement
(tree.cuh:332-393): node existence, leaf flags, sizes and range starts
exactly; inside every split node the children tile the range in order of
FP64 distance to the node's pivot (up to the FP32 error of the keys: 1e-5
relative, or the angle's bound), each
child's min_dis is its smallest distance, a leaf child keeps the node's order
(ascending by (distance, ID)), and the pivot of a split child is the object
of rank (size - 1) // 2 of the child's range by (FP64 distance to its
parent's pivot, ID) wherever that rank's key is distinct.
"""
import numpy as np
import pytest

import oracle
from gts_amd import GTSIndex, datagen

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _tol(metric, dim, d):
    """The FP32 key's error bound: 1e-5 relative (1e-5 absolute below 1); for
    the angle, the bound of its FP32 cosine's error through acos,
    (180 / pi)(pi / sqrt 2) sqrt((4 dim + 32) 2^-24) + 1e-4 degrees."""
    if metric == 5:
        return 57.29578 * 2.2215 * np.sqrt((4 * dim + 32) * 2.0 ** -24) + 1e-4 + 0.0 * np.asarray(d)
    return TOL * np.maximum(1.0, d)


def _fp64_dist(x, rows, piv, metric):
    a = x[rows].astype(np.float64)
    b = x[piv].astype(np.float64)
    if metric == 2:
        return np.sqrt(((a - b) ** 2).sum(axis=1))
    if metric == 1:
        return np.abs(a - b).sum(axis=1)
    if metric == 0:
        return np.abs(a - b).max(axis=1)
    c = (a @ b) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b))
    return np.abs(np.arccos(np.clip(c, -1.0, 1.0)) * 180.0 / 3.1415926)


def _check_tree(x, metric, T, ms, mh=0):
    ds = datagen.Dataset(metric, vectors=x)
    od = oracle.OracleData(ds)
    idx = GTSIndex(ds, T, ms, mh)
    tr = idx.export_tree()
    ot = oracle.OracleTree(od, T, ms, mh)
    n = x.shape[0]
    assert tr["height"] == ot.height
    perm = tr["perm"]
    assert (np.sort(perm) == np.arange(n)).all()
    ex = (tr["flags"] & 1) != 0
    assert (ex == (ot.exists != 0)).all()
    leaf = (tr["flags"] & 2) != 0
    assert (leaf[ex] == (ot.is_leaf[ex] != 0)).all()
    assert (tr["size"][ex] == ot.size[ex]).all() and (tr["lid"][ex] == ot.lid[ex]).all()
    decided = total = 0
    dim = x.shape[1]
    for nd in np.nonzero(ex & ~leaf)[0]:
        lid, size, piv = int(tr["lid"][nd]), int(tr["size"][nd]), int(tr["pivot"][nd])
        members = perm[lid:lid + size]
        d = _fp64_dist(x, members, piv, metric)
        prev_hi = -np.inf
        for c in range(T):
            ch = nd * T + 1 + c
            if ch >= len(ex) or not ex[ch]:
                continue
            cl, cs = int(tr["lid"][ch]) - lid, int(tr["size"][ch])
            if cs == 0:
                continue
            cm, cd = members[cl:cl + cs], d[cl:cl + cs]
            # the children tile the range in distance order (up to the keys'
            # FP32 error); min_dis is the child's smallest distance
            assert cd.min() >= prev_hi - _tol(metric, dim, max(prev_hi, 0.0)), (nd, ch)
            assert abs(tr["min_dis"][ch] - cd.min()) <= _tol(metric, dim, cd.min()), (nd, ch)
            prev_hi = cd.max()
            if leaf[ch]:
                # a leaf keeps this node's order: ascending by (distance, ID)
                # (FP32 keys: identical rows tie exactly and fall in ID order;
                # other FP64 ties may fall either way)
                assert (cd[1:] >= cd[:-1] - _tol(metric, dim, cd[1:])).all(), (nd, ch)
                same = np.all(x[cm[1:]] == x[cm[:-1]], axis=1)
                assert (cm[1:][same] > cm[:-1][same]).all(), (nd, ch)
                continue
            # the pivot rule for a split child, by this node's order
            order = np.lexsort((cm, cd))
            mid = (cs - 1) // 2
            ks = cd[order]
            total += 1
            if (np.abs(ks - ks[mid]) <= 2 * _tol(metric, dim, ks[mid])).sum() == 1:
                assert int(tr["pivot"][ch]) == int(cm[order[mid]]), (nd, ch)
                decided += 1
    # most pivots are decided beyond the keys' error (not for the angle, whose
    # FP32 error of ~0.25 degree exceeds the spacing of these rows' angles)
    assert metric == 5 or decided >= total // 2
    # the search over the tree still equals brute force
    qids = np.arange(0, n, max(1, n // 24), dtype=np.int32)
    ids, dists, _ = idx.knn_query(qids, 5)
    bids, bd = oracle.bf_knn(od, qids, 5)
    assert (ids == bids).all()
    return tr


def test_lds_and_rank_keyed_global_sorts_float4_rows():
    """60,000 x 8 L2: the root by the global sort (rank keys, 32 bits), the
    levels below by the LDS sort; 16-byte key loads."""
    rng = np.random.default_rng(21)
    x = rng.random((60_000, 8)).astype(np.float32)
    x[5000:5300] = x[17]                       # exact duplicates: ties by ID
    _check_tree(x, 2, 10, 20)


def test_global_sort_with_leaves_at_the_level():
    """544,031 x 4 L1, T = 32, leaf capacity 17,010: level 1 has 31 leaves of
    17,000 and one node of 17,031 > 16,384, sorted by the global sort keyed by
    range starts (the 31 leaves keep the root's order); its 32 children of
    532 are leaves."""
    rng = np.random.default_rng(22)
    n = 32 * 17_000 + 31
    x = rng.random((n, 4)).astype(np.float32)
    tr = _check_tree(x, 1, 32, 17_010)
    sizes = tr["size"][1:33]
    assert (sizes == 17_000).sum() == 31 and sizes.max() == 17_031


@pytest.mark.parametrize("metric,dim", [(2, 5), (0, 12), (5, 8)])
def test_scalar_and_float4_keys_other_metrics(metric, dim):
    """20,000 rows: dim 5 takes the 4-byte key loads; L-inf and angular the
    16-byte ones; T = 4 puts the root above the LDS limit."""
    rng = np.random.default_rng(23 + dim)
    x = (rng.random((20_000, dim)) + 0.05).astype(np.float32)
    _check_tree(x, metric, 4, 20)
