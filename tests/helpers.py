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


def dense_sample(x, eps, p, frac=0.01, seed=77, dense_cells=16):
    """Seeded query sample of about frac * N points that contains every point
    of the dense_cells most occupied grid cells (the heaviest neighborhoods)
    and fills up with uniformly drawn points.  Cells are taken over the k
    highest-variance dims of an approximate reordering; the selection only
    decides which queries are checked."""
    x = np.asarray(x)
    n, d = x.shape
    k = p["k"]
    if p.get("reorder"):
        var = x[::100].astype(np.float64).var(axis=0)
        dims = np.argsort(-var, kind="stable")[:k]
    else:
        dims = np.arange(k)
    xi = x[:, dims].astype(np.float64)
    lo = xi.min(axis=0) - eps
    c = np.floor((xi - lo) / eps).astype(np.int64)
    key = np.zeros(n, np.int64)
    for j in range(k):
        key = key * (int(c[:, j].max()) + 1) + c[:, j]
    uk, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    want = max(1, int(round(frac * n)))
    top = np.argsort(-cnt, kind="stable")[:dense_cells]
    dense = np.nonzero(np.isin(inv, top))[0]
    rng = np.random.default_rng(seed)
    if len(dense) > want // 2:
        dense = np.sort(rng.choice(dense, want // 2, replace=False))
    rest = np.setdiff1d(np.arange(n), dense)
    extra = rng.choice(rest, max(0, want - len(dense)), replace=False)
    return np.sort(np.concatenate([dense, extra])).astype(np.uint32)


def fold_straddle_data(dt=np.float32, d=6, eps=1.5, want=16, seed=11):
    """A point q plus points c at distance about eps from it whose squared
    distance, summed in the two readings of the ILP fold (ILP = 8,
    SHORTCIRCUIT = 1: fused or rounded products added), falls on opposite
    sides of the threshold, plus uniform background points.  Returns
    (data, eps, number of straddling points)."""
    dt = np.dtype(dt).type
    e = dt(eps)
    thr = float(np.float32(e) * np.float32(e)) if dt is np.float32 else float(e) * float(e)
    import math
    t = dt(thr)
    # the largest DTYPE value whose correctly rounded square root is <= eps
    while math.sqrt(float(t)) > float(e):
        t = np.nextafter(t, dt(0))
    while math.sqrt(float(np.nextafter(t, dt(np.inf)))) <= float(e):
        t = np.nextafter(t, dt(np.inf))
    rng = np.random.default_rng(seed)
    q = np.full(d, 50.25, dtype=dt)
    pts = [q]
    found = 0
    for _ in range(200000):
        v = rng.standard_normal(d)
        v /= np.linalg.norm(v)
        c = (q.astype(np.float64) + v * float(e) * (1.0 + rng.uniform(-4.0, 4.0) * np.finfo(dt).eps)).astype(dt)
        a0 = O.acc(q, c, ilp=8, shortcircuit=1, fold_fused=0)
        a1 = O.acc(q, c, ilp=8, shortcircuit=1, fold_fused=1)
        if (a0 <= t) != (a1 <= t):
            pts.append(c)
            found += 1
            if found == want:
                break
    bg = (rng.random((400, d)) * 8.0 + 46.0).astype(dt)
    data = np.concatenate([np.stack(pts), bg]).astype(dt)
    return np.ascontiguousarray(data), float(e), found
