"""C ABI checks that need no GPU: exported symbols and the text loaders."""
import os
import re

import numpy as np
import pytest

import gts_amd
from gts_amd import GTSError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "gts.h")).read()
    names = set(re.findall(r"\b(gts_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 20
    L = gts_amd.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_loaders_roundtrip(tmp_path):
    from gts_amd import datagen
    rng = np.random.default_rng(0)
    x = rng.random((7, 3)).astype(np.float32)
    p = str(tmp_path / "d.txt")
    datagen.write_data(p, datagen.Dataset(1, vectors=x))
    ds = gts_amd.load_dataset(p)
    assert ds.info() == (7, 3, 1)
    q = str(tmp_path / "q.txt")
    datagen.write_queries(q, [3, 1, 4])
    assert list(gts_amd.load_queries(q)) == [3, 1, 4]
    u = str(tmp_path / "u.txt")
    datagen.write_updates(u, [(0, 2), (1, 5), (2, 3)])
    assert gts_amd.load_updates(u).tolist() == [[0, 2], [1, 5], [2, 3]]
    s = str(tmp_path / "s.txt")
    datagen.write_data(s, datagen.Dataset(6, strings=[b"abc", b"", b"hello"]))
    assert gts_amd.load_dataset(s).info()[::2] == (3, 6)


@pytest.mark.parametrize("text,code", [("", 2), ("3 2\n1 2 3\n", 2), ("3 2 1\n1 2 3\n", 2),
                                       ("2 2 9\n1 2\n3 4\n", 2), ("3 1 1\n1 2\n", 2)])
def test_malformed_data_files(tmp_path, text, code):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(GTSError) as e:
        gts_amd.load_dataset(str(p))
    assert e.value.code == code


def test_missing_file_is_io_error():
    with pytest.raises(GTSError) as e:
        gts_amd.load_dataset("/nonexistent/file.txt")
    assert e.value.code == 1


def test_reference_sample_loads():
    path = "/root/reference/Datasets Sample/sbw_example.txt"
    if not os.path.exists(path):
        pytest.skip("reference sample not mounted")
    ds = gts_amd.load_dataset(path)
    assert ds.info() == (1000, 300, 5)
    q = gts_amd.load_queries("/root/reference/Datasets Sample/sbw_example_qid.txt")
    assert len(q) == 1000 and q[0] == 2
    u = gts_amd.load_updates("/root/reference/Datasets Sample/sbw_example_u.txt")
    assert u.tolist()[:2] == [[0, 2], [1, 11]]
