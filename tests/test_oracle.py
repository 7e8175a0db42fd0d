"""Pins the CPU oracle before it is trusted as the checker.

- README known answer (README.md:82-84): 27,982,028 on test_dat.txt, eps=1.0,
  FP32, no optimizations -- pins ILP=0 arithmetic, the predicate, self
  inclusion and the candidate set.
- exact-rational restatement of each accumulation mode (Fraction arithmetic
  with correct rounding done here, not by the C library), so a dropped FMA,
  a wrong fold order or a transposed difference fails.
- the grid's candidate restriction against an all-pairs scan, and the closed
  forms "eps below the minimum distance => every count is 1" and
  "eps at least the diameter => every count is n".
- REORDER against a float64 variance ranking where variances are well
  separated.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import README_COUNT, load_test_dat


def test_readme_known_answer():
    data = load_test_dat(np.float32)
    assert data.shape == (100_000, 2)
    r = O.join(data, 1.0, k=2, with_ids=False)
    assert int(r["counts"].sum(dtype=np.uint64)) == README_COUNT
    assert int(r["stats"][0]) == README_COUNT


def test_readme_other_orders_differ():
    # the known answer separates the fused order from the unfused one (ILP=8 with
    # SHORTCIRCUIT=1 on 2 dims is fma(t0,t0,0) + round(t1*t1))
    data = load_test_dat(np.float32)
    r = O.join(data, 1.0, k=2, ilp=8, shortcircuit=1, with_ids=False)
    assert int(r["counts"].sum(dtype=np.uint64)) != README_COUNT


# ---- exact arithmetic restatement -----------------------------------------
def _round(fr, dt):
    """Correctly rounded (ties to even) value of the rational fr in dtype dt."""
    if dt == np.float64:
        return Fraction(float(fr))
    c = np.float32(float(fr))
    cands = [np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))]
    best = None
    for v in cands:
        if not np.isfinite(v):
            continue
        err = abs(Fraction(float(v)) - fr)
        key = (err, int(np.array(v, np.float32).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, v)
    return Fraction(float(best[1]))


def _acc_exact(p, q, ilp, sc, dt):
    P = [Fraction(float(x)) for x in p]
    Q = [Fraction(float(x)) for x in q]
    d = len(P)
    sub = lambda a, b: _round(a - b, dt)
    fma = lambda a, b, c: _round(a * b + c, dt)
    add = lambda a, b: _round(a + b, dt)
    mul = lambda a, b: _round(a * b, dt)
    if ilp == 0:
        acc = Fraction(0)
        for l in range(d):
            t = sub(Q[l], P[l])
            acc = fma(t, t, acc)
        return acc
    if sc:
        r0 = Fraction(0)
        for l in range(0, d, ilp):
            t = sub(Q[l], P[l])
            r0 = fma(t, t, r0)
            for j in range(1, ilp):
                if l + j < d:
                    t = sub(Q[l + j], P[l + j])
                    r0 = add(r0, mul(t, t))
        return r0
    r = [Fraction(0)] * ilp
    for l in range(0, d, ilp):
        for j in range(ilp):
            if l + j < d:
                t = sub(Q[l + j], P[l + j])
                r[j] = fma(t, t, r[j])
    acc = r[0]
    for j in range(1, ilp):
        acc = add(acc, r[j])
    return acc


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("ilp,sc", [(0, 0), (0, 1), (8, 1), (8, 0), (3, 1), (3, 0)])
def test_accumulation_matches_exact_rational(dt, ilp, sc):
    rng = np.random.default_rng(7 + ilp * 3 + sc)
    for trial in range(40):
        d = int(rng.integers(1, 20))
        scale = 10.0 ** rng.integers(-3, 4)
        p = (rng.standard_normal(d) * scale).astype(dt)
        q = (p + rng.standard_normal(d) * scale * 1e-3).astype(dt)
        got = O.acc(p, q, ilp=ilp, shortcircuit=sc)
        want = _acc_exact(p, q, ilp, sc, dt)
        assert Fraction(float(got)) == want, (trial, d, ilp, sc)


def test_fused_fold_reading_is_sequential_fma():
    # fold_fused=1 turns every dim into fma(t,t,r0): the ILP=0 order
    rng = np.random.default_rng(3)
    for _ in range(50):
        d = int(rng.integers(1, 30))
        p = rng.standard_normal(d).astype(np.float32)
        q = (p + 0.01 * rng.standard_normal(d)).astype(np.float32)
        assert O.acc(p, q, ilp=8, shortcircuit=1, fold_fused=1) == O.acc(p, q, ilp=0)


# ---- candidate set -----------------------------------------------------------
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("d,k", [(1, 1), (2, 2), (3, 2), (5, 3), (8, 4), (17, 6)])
def test_grid_equals_bruteforce(dt, d, k):
    rng = np.random.default_rng(100 + d * 7 + k)
    n = 600
    x = (rng.random((n, d)) * 10).astype(dt)
    # pick eps so a point has a few dozen neighbors
    eps = float(np.asarray(np.sort(np.linalg.norm(x[:50, None, :].astype(np.float64) - x[None, :, :], axis=2),
                                   axis=1)[:, 20].mean(), dtype=dt))
    for ilp, sc, reorder in [(0, 0, 0), (8, 1, 1), (8, 0, 0)]:
        g = O.join(x, eps, k=k, ilp=ilp, shortcircuit=sc, reorder=reorder, with_ids=True)
        b = O.bruteforce(x, eps, ilp=ilp, shortcircuit=sc, reorder=reorder)
        assert np.array_equal(g["counts"], b["counts"])
        assert np.array_equal(g["hashes"], b["hashes"])
        assert np.array_equal(O.csr_hashes(g["counts"], g["ids"]), b["hashes"])


def test_closed_form_isolated_points():
    # points on a lattice of spacing 1, eps = 0.4: every point only finds itself
    g = np.stack(np.meshgrid(np.arange(30), np.arange(30), np.arange(3), indexing="ij"), -1).reshape(-1, 3)
    x = g.astype(np.float32)
    r = O.join(x, 0.4, k=3)
    assert np.all(r["counts"] == 1)
    assert np.array_equal(r["ids"], np.arange(len(x), dtype=np.uint32))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_closed_form_all_pairs(dt):
    rng = np.random.default_rng(5)
    x = rng.random((300, 4)).astype(dt)
    r = O.join(x, 2.5, k=4)  # diameter of the unit 4-cube is 2
    assert np.all(r["counts"] == 300)


def test_self_inclusion_and_symmetry():
    rng = np.random.default_rng(9)
    x = (rng.random((2000, 3)) * 20).astype(np.float32)
    r = O.join(x, 1.3, k=3, ilp=8, shortcircuit=1)
    counts, ids = r["counts"], r["ids"]
    row = np.repeat(np.arange(len(x)), counts)
    pairs = set(zip(row.tolist(), ids.tolist()))
    assert all((i, i) in pairs for i in range(len(x)))
    assert all((j, i) in pairs for i, j in pairs)


def test_reorder_matches_float64_ranking():
    rng = np.random.default_rng(11)
    scales = np.array([1.0, 9.0, 3.0, 27.0, 0.3, 5.0])
    x = (rng.standard_normal((50_000, 6)) * scales).astype(np.float32)
    perm, var = O.reorder_perm(x)
    ref = np.argsort(-np.var(x[::100].astype(np.float64), axis=0), kind="stable")
    assert list(perm) == list(ref)


def test_parse_text_direct_decimal_to_float():
    v = O.parse_text("0.1,0.2\n3.4028235e38, 1e-45\n", np.float32)
    assert v.dtype == np.float32
    assert list(v) == [np.float32("0.1"), np.float32("0.2"), np.float32(3.4028235e38), np.float32(1e-45)]


def test_cpu_baselines_rtree_equals_bruteforce():
    """The restated reference CPU paths (timing baselines) find the same
    counts: R-tree search + filter == all-pairs scan with the same filter."""
    rng = np.random.default_rng(21)
    for d, dt, eps in ((2, np.float32, 0.3), (5, np.float64, 1.1), (12, np.float32, 2.5)):
        x = (rng.random((3000, d)) * 6).astype(dt)
        q = np.arange(0, 3000, 3, dtype=np.uint32)
        e = float(np.asarray(eps, dtype=dt))
        cb, _ = O.cpu_bruteforce(x, e, q, 2)
        cr, _, _ = O.cpu_rtree(x, e, q)
        assert np.array_equal(cb, cr)
        assert np.all(cb >= 1)
        # the grid contract fuses its arithmetic: counts may differ only at
        # the epsilon boundary, which random data almost never hits
        g = O.join(x, e, k=min(d, 6), with_ids=False)["counts"][q]
        assert int(np.abs(g.astype(np.int64) - cb).sum()) <= 2


# ---- the two readings of the ILP fold ----------------------------------------
def _pairs(counts, ids):
    rows = np.repeat(np.arange(len(counts), dtype=np.uint64), counts.astype(np.int64))
    return set((rows << np.uint64(32) | ids.astype(np.uint64)).tolist())


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_fold_ambiguity_count_equals_table_difference(dt):
    """stats[3] (pairs on which the fused and the unfused fold disagree) is
    the size of the symmetric difference of the two readings' tables, on
    data built to contain such pairs."""
    from tests.helpers import fold_straddle_data
    x, eps, found = fold_straddle_data(dt)
    assert found >= 8
    r0 = O.join(x, eps, k=6, ilp=8, shortcircuit=1, reorder=0, fold_fused=0)
    r1 = O.join(x, eps, k=6, ilp=8, shortcircuit=1, reorder=0, fold_fused=1)
    diff = _pairs(r0["counts"], r0["ids"]) ^ _pairs(r1["counts"], r1["ids"])
    assert len(diff) >= 2 * found
    assert int(r0["stats"][3]) == len(diff) == int(r1["stats"][3])
    # the fused reading is the ILP = 0 table
    s = O.join(x, eps, k=6, ilp=0, reorder=0)
    assert np.array_equal(s["counts"], r1["counts"]) and np.array_equal(s["ids"], r1["ids"])


def test_fold_ambiguity_on_golden_fixtures():
    """On the ILP > 0 / SHORTCIRCUIT fixtures the counter agrees with the
    tables of both readings."""
    import glob
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    seen = 0
    for path in sorted(glob.glob(os.path.join(here, "grid_*.npz"))):
        z = np.load(path)
        k, ilp, sc, reorder = (int(v) for v in z["params"])
        if ilp == 0 or sc == 0:
            continue
        x, eps = z["data"], float(z["eps"])
        r0 = O.join(x, eps, k=k, ilp=ilp, shortcircuit=sc, reorder=reorder, fold_fused=0)
        r1 = O.join(x, eps, k=k, ilp=ilp, shortcircuit=sc, reorder=reorder, fold_fused=1)
        diff = _pairs(r0["counts"], r0["ids"]) ^ _pairs(r1["counts"], r1["ids"])
        assert int(r0["stats"][3]) == len(diff), path
        print(os.path.basename(path), "ambiguous pairs", len(diff))
        seen += 1
    assert seen >= 3


# ---- SORT (kernel.cu:400-467) --------------------------------------------
def _lattice(d, side=7, base=10.0):
    g = np.stack(np.meshgrid(*[np.arange(side, dtype=np.float32)] * d, indexing="ij"), -1).reshape(-1, d)
    return np.ascontiguousarray(g + np.float32(base))


def test_sort_equals_plain_join_away_from_the_boundary():
    """On random data SORT only skips candidates that fail the predicate: the
    tables equal the join without SORT, with fewer candidates evaluated."""
    rng = np.random.default_rng(5)
    for d, k in ((3, 3), (6, 6), (8, 4), (5, 2)):
        c = rng.random((6, d)) * 10
        x = (c[rng.integers(0, 6, 3000)] + rng.standard_normal((3000, d))).astype(np.float32)
        a = O.join(x, 0.8, k=k, ilp=8, shortcircuit=1, reorder=1)
        b = O.join(x, 0.8, k=k, ilp=8, shortcircuit=1, reorder=1, sort=1)
        assert np.array_equal(a["counts"], b["counts"]) and np.array_equal(a["ids"], b["ids"])
        assert b["stats"][1] < a["stats"][1]


def test_sortidu_window_ends_strictly_below_eps():
    """With a sort dim beyond the indexed ones, the walk runs while value -
    query < eps (kernel.cu:455): on an integer lattice with eps = 1 exactly
    the pairs (q, q + e_sd) drop out, and (q + e_sd, q) stays (the search
    steps back while query - value <= eps)."""
    x = _lattice(3)
    a = O.join(x, 1.0, k=2, reorder=0)
    b = O.join(x, 1.0, k=2, reorder=0, sort=1)
    n = len(x)
    rows = np.repeat(np.arange(n), a["counts"].astype(np.int64))
    keep = ~((x[a["ids"], 2] - x[rows, 2] == 1.0) & np.all(x[a["ids"], :2] == x[rows, :2], axis=1))
    want = np.bincount(rows[keep], minlength=n)
    assert np.array_equal(b["counts"], want)
    assert int(a["counts"].sum() - b["counts"].sum()) == 294  # 7 x 7 columns x 6 steps
    # every dim indexed (sort dim 0): the walk stops only past eps, so the
    # lattice keeps every pair
    c = O.join(x, 1.0, k=3, reorder=0, sort=1)
    assert np.array_equal(c["counts"], O.join(x, 1.0, k=3, reorder=0)["counts"])
