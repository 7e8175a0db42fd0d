"""Unicomp (STAMP, stamploops.h:2-201, and LINEARSTAMP, kernel.cu:272-289):
each pair of adjacent cells is evaluated from one side (the odd-coordinate
owner, or the lower linear id) and both directions are emitted.  The result set is
the one without unicomp (the pair's arithmetic is symmetric:
fl(a - b) = -fl(b - a)); only the order inside a row may differ, so rows are
compared sorted, and with sort_neighbors the tables are byte-identical."""
import numpy as np
import pytest

from gpu_self_join_amd import _lib
from oracle import oracle as O
from tests.helpers import assert_csr_equal, sorted_rows

pytestmark = pytest.mark.gpu


def _join(sj, x, eps, **kw):
    import torch
    r = sj.join(torch.from_numpy(np.ascontiguousarray(x)).cuda(), eps, **kw)
    return r.counts.cpu().numpy().view(np.uint32), r.ids.cpu().numpy().view(np.uint32), r


CASES = [  # n, dims, k, dtype, ilp, sc, reorder, spread
    (20000, 2, 2, np.float32, 0, 0, 0, "uniform"),
    (15000, 3, 3, np.float32, 8, 1, 1, "gmm"),
    (12000, 4, 2, np.float64, 0, 0, 0, "gmm"),
    (10000, 5, 5, np.float32, 8, 0, 1, "gmm"),
    (9000, 6, 6, np.float32, 8, 1, 1, "gmm"),
    (6000, 8, 4, np.float64, 8, 1, 0, "uniform"),
    (8000, 7, 3, np.float32, 4, 1, 0, "gmm"),  # generic-dims kernel
]


def _data(n, d, dt, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return (rng.random((n, d)) * 20).astype(dt)
    c = rng.random((8, d)) * 20
    return (c[rng.integers(0, 8, n)] + rng.standard_normal((n, d))).astype(dt)


@pytest.mark.parametrize("mode", [1, 2], ids=["stamp", "linearstamp"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_d%d_k%d_%s_ilp%d_sc%d_r%d_%s" % (
    c[0], c[1], c[2], np.dtype(c[3]).name, c[4], c[5], c[6], c[7]))
def test_stamp_equals_oracle_and_full_neighborhood(sj, case, mode):
    n, d, k, dt, ilp, sc, reorder, kind = case
    x = _data(n, d, dt, kind, n + d)
    eps = float(np.asarray(0.35 if kind == "uniform" and d <= 3 else 0.9 + 0.15 * d, dtype=dt))
    kw = dict(k=k, ilp=ilp, shortcircuit=sc, reorder=reorder, filter=_lib.FILTER_PLAIN)
    cs, is_, rs = _join(sj, x, eps, stamp=mode, **kw)
    cf, if_, rf = _join(sj, x, eps, stamp=0, **kw)
    print(case, "pairs", rs.total, "candidates stamp %d full %d" % (rs.stats["candidates"], rf.stats["candidates"]),
          flush=True)
    assert np.array_equal(cs, cf)
    assert np.array_equal(sorted_rows(cs, is_), sorted_rows(cf, if_))
    # each cross-cell pair is evaluated once: fewer candidate pairs
    assert rs.stats["candidates"] < rf.stats["candidates"]
    ref = O.join(x, eps, k=k, ilp=ilp, shortcircuit=sc, reorder=reorder)
    assert_csr_equal(cs, is_, ref["counts"], ref["ids"])
    # sorted rows: byte-identical tables
    c2, i2, _ = _join(sj, x, eps, stamp=mode, sort_neighbors=1, **kw)
    c3, i3, _ = _join(sj, x, eps, stamp=0, sort_neighbors=1, **kw)
    assert np.array_equal(c2, c3) and np.array_equal(i2, i3)


@pytest.mark.parametrize("mode", [1, 2], ids=["stamp", "linearstamp"])
def test_stamp_dense_cells_span_several_tiles(sj, mode):
    # cells holding hundreds of points: a cell's queries span several tiles,
    # own-cell pairs are found from both of their tiles, cross-cell pairs once
    rng = np.random.default_rng(4)
    x = (rng.random((30000, 2)) * 4).astype(np.float32)
    eps = float(np.float32(0.5))
    cs, is_, rs = _join(sj, x, eps, k=2, ilp=0, shortcircuit=0, reorder=0, stamp=mode)
    ref = O.join(x, eps, k=2)
    assert_csr_equal(cs, is_, ref["counts"], ref["ids"])


def test_stamp_ignored_with_estimator_and_gram(sj):
    # unicomp applies to the exact passes of the plain kernels; with the Gram
    # filter or an estimator the full neighborhood is evaluated (same result)
    rng = np.random.default_rng(6)
    x = (rng.random((5000, 16)) * 3).astype(np.float32)
    eps = float(np.float32(1.5))
    cg, ig, rg = _join(sj, x, eps, k=4, stamp=1, filter=_lib.FILTER_GRAM)
    cp, ip, rp = _join(sj, x, eps, k=4, stamp=0, filter=_lib.FILTER_PLAIN)
    assert np.array_equal(cg, cp) and np.array_equal(ig, ip)


def test_both_ownership_rules_evaluate_each_cell_pair_once(sj):
    # without pruning (k < 4) both rules evaluate every own-cell pair and
    # every adjacent cell pair exactly once: the same number of candidates
    x = _data(20000, 3, np.float32, "uniform", 9)
    eps = float(np.float32(0.6))
    kw = dict(k=3, ilp=0, shortcircuit=0, reorder=0, filter=_lib.FILTER_PLAIN)
    _, _, r1 = _join(sj, x, eps, stamp=1, **kw)
    _, _, r2 = _join(sj, x, eps, stamp=2, **kw)
    assert r1.stats["candidates"] == r2.stats["candidates"]
    with pytest.raises(_lib.GdsjError):
        _join(sj, x, eps, stamp=3, **kw)
