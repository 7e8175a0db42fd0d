"""Shared test helpers: the README fixture, CSR comparison, oracle calls."""
import hashlib
import lzma
import os

import numpy as np

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
TEST_DAT_SHA256 = "dc6ad96a6a7a56460563a48096fe35c0debe291c12f6416e0d695ae048557b64"
README_COUNT = 27_982_028  # README.md:82-84, test_dat.txt, eps=1.0, FP32, no optimizations


def load_test_dat_text():
    raw = lzma.open(os.path.join(GOLDEN, "test_dat.txt.xz")).read()
    assert hashlib.sha256(raw).hexdigest() == TEST_DAT_SHA256
    return raw


def load_test_dat(dtype=np.float32):
    v = O.parse_text(load_test_dat_text(), dtype)
    return v.reshape(-1, 2)


def sorted_rows(counts, ids):
    """CSR with every row sorted ascending (the comparison form)."""
    counts = np.asarray(counts, dtype=np.int64)
    ids = np.asarray(ids, dtype=np.uint32)
    offs = np.zeros(len(counts) + 1, np.int64)
    np.cumsum(counts, out=offs[1:])
    row = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
    order = np.lexsort((ids.astype(np.int64), row))
    return ids[order]


def assert_csr_equal(c_gpu, ids_gpu, c_ref, ids_ref):
    c_gpu = np.asarray(c_gpu, dtype=np.uint32)
    c_ref = np.asarray(c_ref, dtype=np.uint32)
    bad = np.nonzero(c_gpu != c_ref)[0]
    assert len(bad) == 0, "counts differ at %d points, first %s: gpu %s ref %s" % (
        len(bad), bad[:5], c_gpu[bad[:5]], c_ref[bad[:5]])
    a = sorted_rows(c_gpu, ids_gpu)
    b = sorted_rows(c_ref, ids_ref)
    assert a.shape == b.shape
    diff = np.nonzero(a != b)[0]
    assert len(diff) == 0, "neighbor ids differ at %d positions" % len(diff)


def oracle_params(p):
    return dict(k=p["k"], ilp=p["ilp"], shortcircuit=p["shortcircuit"], reorder=p["reorder"])
