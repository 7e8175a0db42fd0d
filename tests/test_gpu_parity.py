"""Parity of the HIP path (libgdsjoin.so through its C ABI) with the oracle.

Bar: bit-exact counts and identical per-point neighbor sets (rows compared
after sorting, since the reference's row order follows its atomics).
"""
import ctypes as C

import numpy as np
import pytest

from gpu_self_join_amd import _lib, gdsjoin
from oracle import oracle as O
from tests.helpers import README_COUNT, assert_csr_equal, load_test_dat

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _gpu_join(sj, x, eps, **kw):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    r = sj.join(t, eps, **kw)
    counts = r.counts.cpu().numpy().view(np.uint32)
    ids = r.ids.cpu().numpy().view(np.uint32)
    offs = r.offsets.cpu().numpy()
    assert offs[0] == 0 and int(offs[-1]) == r.total
    assert np.array_equal(np.diff(offs), counts.astype(np.int64))
    return counts, ids, r


def _check(sj, x, eps, k, ilp=0, sc=0, reorder=0, **kw):
    c, ids, r = _gpu_join(sj, x, eps, k=k, ilp=ilp, shortcircuit=sc, reorder=reorder, **kw)
    ref = O.join(x, eps, k=k, ilp=ilp, shortcircuit=sc, reorder=reorder)
    print("case n=%d d=%d k=%d ilp=%d sc=%d reorder=%d total=%d" % (x.shape[0], x.shape[1], k, ilp, sc, reorder,
                                                                    r.total), flush=True)
    assert_csr_equal(c, ids, ref["counts"], ref["ids"])
    return r


def test_readme_known_answer_gpu(sj):
    x = load_test_dat(np.float32)
    r = _check(sj, x, 1.0, k=2, ilp=0, sc=0, reorder=0, queryreorder=0)
    assert r.total == README_COUNT


def test_readme_through_gdsjoin_dropin(gpu):
    x = load_test_dat(np.float32)
    p = _lib.default_params(k=2, ilp=0, shortcircuit=0, reorder=0, queryreorder=0)
    counts, table = gdsjoin(x.ravel(), 1.0, 2, "float", verbose=False, params=p)
    assert counts.dtype == np.uint32 and table.dtype == np.uint32
    assert int(counts.sum()) == README_COUNT == len(table)
    ref = O.join(x, 1.0, k=2)
    assert_csr_equal(counts, table, ref["counts"], ref["ids"])


def test_host_path_staging_ring_equals_device_join(sj):
    """The drop-in's host path moves the points in and the table out through
    a ring of pinned buffers (one stream, host copy threads): 72 MB of points
    and a table of about 200 MB go round it more than once.  The counts and
    rows must equal the device-resident join of the same points."""
    import torch
    rng = np.random.default_rng(21)
    x = (rng.random((4_500_000, 4)) * np.array([60.0, 60.0, 60.0, 60.0])).astype(np.float32)
    eps = 1.6
    kw = dict(k=4, ilp=0, shortcircuit=0, reorder=0)
    counts, table = gdsjoin(x.ravel(), eps, 4, "float", verbose=False, params=_lib.default_params(**kw))
    print("host path pairs", len(table), flush=True)
    assert x.nbytes > 16 * (4 << 20) and table.nbytes > 16 * (4 << 20)
    res = sj.join(torch.from_numpy(x).cuda(), eps, **kw)
    dc = res.counts.cpu().numpy().view(np.uint32)
    di = res.ids[: res.total].cpu().numpy().view(np.uint32)
    assert_csr_equal(counts, table, dc, di)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_ilp_fold_readings_on_straddling_pairs(sj, dt):
    """Data with pairs on which the two readings of the ILP fold disagree:
    ilp_fold = 0 gives the oracle's default table, ilp_fold = 1 the fused
    one, and the two differ."""
    from tests.helpers import fold_straddle_data
    x, eps, found = fold_straddle_data(dt)
    assert found >= 8
    for fold in (0, 1):
        c, ids, r = _gpu_join(sj, x, eps, k=6, ilp=8, shortcircuit=1, reorder=0, ilp_fold=fold)
        ref = O.join(x, eps, k=6, ilp=8, shortcircuit=1, reorder=0, fold_fused=fold)
        assert_csr_equal(c, ids, ref["counts"], ref["ids"])
        if fold == 0:
            c0 = c
    assert not np.array_equal(c0, c)


CASES = [
    # (n, d, k, dtype, ilp, sc, reorder, neighbors)
    (3000, 2, 2, np.float32, 0, 0, 0, 30),
    (3000, 2, 2, np.float32, 8, 1, 0, 30),
    (4000, 3, 3, np.float32, 8, 1, 1, 40),
    (5000, 4, 4, np.float32, 0, 1, 0, 20),
    (6000, 6, 6, np.float32, 8, 1, 1, 64),
    (4000, 6, 6, np.float32, 8, 0, 1, 30),
    (3000, 5, 3, np.float32, 0, 0, 0, 25),
    (3000, 8, 4, np.float32, 8, 1, 1, 30),
    (2500, 16, 6, np.float32, 8, 1, 1, 30),
    (2000, 32, 6, np.float32, 8, 1, 1, 32),
    (1500, 64, 8, np.float32, 8, 1, 1, 20),
    (2000, 7, 5, np.float32, 3, 1, 1, 25),
    (2000, 10, 6, np.float32, 0, 0, 1, 25),
    (3000, 2, 2, np.float64, 0, 0, 0, 30),
    (3000, 4, 4, np.float64, 8, 1, 1, 30),
    (2000, 16, 6, np.float64, 8, 1, 1, 30),
    (1500, 64, 8, np.float64, 8, 1, 1, 50),
    (1500, 9, 4, np.float64, 8, 0, 0, 20),
]


def _eps_for(x, nb):
    rng = np.random.default_rng(0)
    q = x[rng.choice(len(x), 64, replace=False)].astype(np.float64)
    d = np.sqrt(((q[:, None, :] - x[None, :, :].astype(np.float64)) ** 2).sum(-1))
    return float(np.asarray(np.sort(d, axis=1)[:, min(nb, len(x) - 1)].mean(), dtype=x.dtype))


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_d%d_k%d_%s_ilp%d_sc%d_r%d" % (
    c[0], c[1], c[2], np.dtype(c[3]).name, c[4], c[5], c[6]))
def test_random_parity(sj, case):
    n, d, k, dt, ilp, sc, reorder, nb = case
    rng = np.random.default_rng(n + 31 * d + 7 * k + ilp)
    ncl = 5
    centers = rng.random((ncl, d)) * 20
    x = (centers[rng.integers(0, ncl, n)] + rng.standard_normal((n, d))).astype(dt)
    eps = _eps_for(x, nb)
    _check(sj, x, eps, k=k, ilp=ilp, sc=sc, reorder=reorder)


def _reorder_data(n, d, dt, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "ties":
        # columns scaled apart by below an ulp: rounding decides the ranking,
        # which then follows the exact order of the sums (the reference's
        # sequential, unfused accumulation)
        base = rng.standard_normal(n) * 3.0 + 5.0
        sh = 25 if np.dtype(dt) == np.float32 else 54
        x = np.stack([base * (1.0 + j * 2.0 ** -sh) for j in range(d)], axis=1)
    elif kind == "equal":
        # identical columns: equal variances, ties kept in dimension order
        base = rng.random(n) * 10.0
        x = np.stack([base] * d, axis=1)
    else:
        c = rng.random((6, d)) * 10.0
        x = c[rng.integers(0, 6, n)] + rng.standard_normal((n, d)) * rng.random(d) * 2.0
    return np.ascontiguousarray(x.astype(dt))


# n / 100 samples per dimension: several staging chunks, ragged last chunk
@pytest.mark.parametrize("n,d,dt,kind", [(450_123, 8, np.float32, "ties"), (1_000_000, 12, np.float32, "gmm"),
                                         (333_333, 6, np.float64, "ties"), (205_000, 5, np.float32, "equal"),
                                         (150, 4, np.float64, "gmm")])
def test_reorder_permutation_equals_oracle(sj, n, d, dt, kind):
    x = _reorder_data(n, d, dt, kind, n + d)
    eps = float(np.asarray(1e-3, dtype=dt))
    _, _, r = _gpu_join(sj, x, eps, k=2, reorder=1)
    perm, _ = O.reorder_perm(x)
    print("reorder n=%d d=%d %s perm %s" % (n, d, kind, list(perm)), flush=True)
    assert list(r.stats["perm"][:d]) == [int(v) for v in perm]


def test_dense_cell_many_tiles(sj):
    # one cell holds far more than 64 queries: tiles split the cell
    rng = np.random.default_rng(1)
    x = (rng.random((5000, 2)) * 3).astype(np.float32)
    _check(sj, x, 1.0, k=2)


def test_many_runs_window(sj):
    # k=6 with uniform data: a cell has far more than 64 neighbor runs
    rng = np.random.default_rng(2)
    x = (rng.random((20000, 6)) * 4).astype(np.float32)
    _check(sj, x, 0.9, k=6, ilp=8, sc=1, reorder=1)


HITMASK_CASES = [
    # (name, n, d, k, dtype, ilp, sc, reorder, data scale, eps)
    ("uniform2d", 20000, 2, 2, np.float32, 0, 0, 0, 100.0, 1.0),
    ("dense_cell_tiles", 5000, 2, 2, np.float32, 0, 0, 0, 3.0, 1.0),
    ("k6_many_runs", 20000, 6, 6, np.float32, 8, 1, 1, 4.0, 0.9),
    ("k3_of_5_runtime_dims", 8000, 5, 3, np.float32, 0, 0, 0, 10.0, 1.2),
    ("d10_generic", 6000, 10, 6, np.float32, 0, 0, 1, 3.0, 1.1),
    ("f64_d4_nosc", 8000, 4, 4, np.float64, 8, 0, 0, 8.0, 1.0),
]


@pytest.mark.parametrize("case", HITMASK_CASES, ids=lambda c: c[0])
def test_hitmask_fill_equals_recomputing_fill(sj, case):
    """The fill pass that reads the counting pass's hit masks writes the
    same table, byte for byte (same row order), as the fill pass that
    evaluates every distance again, and both equal the oracle."""
    name, n, d, k, dt, ilp, sc, reorder, scale, eps = case
    rng = np.random.default_rng(n + d)
    x = (rng.random((n, d)) * scale).astype(dt)
    eps = float(np.asarray(eps, dtype=dt))
    kw = dict(k=k, ilp=ilp, shortcircuit=sc, reorder=reorder, stamp=0, estimator=0)
    c1, ids1, r1 = _gpu_join(sj, x, eps, hitmask=1, **kw)
    c2, ids2, r2 = _gpu_join(sj, x, eps, hitmask=2, **kw)
    print("hitmask %s total=%d" % (name, r1.total), flush=True)
    assert np.array_equal(c1, c2)
    assert np.array_equal(ids1, ids2)
    ref = O.join(x, eps, k=k, ilp=ilp, shortcircuit=sc, reorder=reorder)
    assert_csr_equal(c1, ids1, ref["counts"], ref["ids"])


def test_hitmask_rejected_where_unsupported(sj):
    x = (np.random.default_rng(5).random((2000, 2)) * 10).astype(np.float32)
    with pytest.raises(_lib.GdsjError):
        _gpu_join(sj, x, 1.0, k=2, stamp=1, hitmask=1)
    with pytest.raises(_lib.GdsjError):
        _gpu_join(sj, x, 1.0, k=2, estimator=1, hitmask=1)


@pytest.mark.parametrize("d,k,est", [(2, 2, 0), (6, 6, 0), (6, 6, 1), (20, 6, 0)])
def test_rejoin_reuses_the_grid(gpu, d, k, est):
    """gdsj_rejoin_device reruns the passes over the last join's grid: the
    same table as the join (byte for byte on the exact path), no grid build."""
    import torch
    from gpu_self_join_amd import SelfJoin
    rng = np.random.default_rng(40 + d)
    x = (rng.random((15000, d)) * (4.0 if d > 2 else 60.0)).astype(np.float32)
    eps = 0.9 if d > 2 else 1.0
    sj = SelfJoin(0)
    try:
        t = torch.from_numpy(x).cuda()
        kw = dict(k=k, ilp=8, shortcircuit=1, reorder=1, estimator=est)
        r1 = sj.join(t, eps, **kw)
        c1, i1 = r1.counts.cpu().numpy(), r1.ids.cpu().numpy()
        for _ in range(2):
            r2 = sj.rejoin(n=len(x))
            c2, i2 = r2.counts.cpu().numpy(), r2.ids.cpu().numpy()
            assert np.array_equal(c1, c2)
            if est == 0:
                assert np.array_equal(i1, i2)
            assert_csr_equal(c2.view(np.uint32), i2.view(np.uint32), c1.view(np.uint32), i1.view(np.uint32))
            assert r2.stats["ms_grid"] < max(0.05, 0.5 * r1.stats["ms_grid"])
        ref = O.join(x, eps, k=k, ilp=8, shortcircuit=1, reorder=1)
        assert_csr_equal(c2.view(np.uint32), i2.view(np.uint32), ref["counts"], ref["ids"])
    finally:
        sj.close()


def test_rejoin_needs_a_join(gpu):
    from gpu_self_join_amd import SelfJoin
    sj = SelfJoin(0)
    try:
        with pytest.raises(_lib.GdsjError):
            sj.rejoin(n=10)
    finally:
        sj.close()


def test_edge_single_point(sj):
    x = np.array([[1.0, 2.0, 3.0]], np.float32)
    c, ids, r = _gpu_join(sj, x, 0.5, k=3)
    assert list(c) == [1] and list(ids) == [0]


def test_edge_duplicates(sj):
    x = np.ones((777, 3), np.float64)
    c, ids, r = _gpu_join(sj, x, 1e-9, k=2)
    assert np.all(c == 777)


def test_edge_isolated(sj):
    g = np.stack(np.meshgrid(np.arange(40), np.arange(40), indexing="ij"), -1).reshape(-1, 2)
    x = g.astype(np.float32)
    c, ids, r = _gpu_join(sj, x, 0.4, k=2)
    assert np.all(c == 1) and np.array_equal(ids, np.arange(len(x), dtype=np.uint32))


def test_edge_all_pairs(sj):
    rng = np.random.default_rng(4)
    x = rng.random((700, 3)).astype(np.float32)
    c, ids, r = _gpu_join(sj, x, 2.0, k=3)
    assert np.all(c == 700)


def test_edge_empty_input(sj):
    torch = _torch()
    t = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    r = sj.join(t, 1.0, k=3)
    assert r.total == 0


def test_invalid_arguments(sj):
    torch = _torch()
    t = torch.zeros((10, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.GdsjError):
        sj.join(t, 0.0, k=3)  # the reference divides by zero here (main.cu:936)
    with pytest.raises(_lib.GdsjError):
        sj.join(t, 1.0, k=4)  # k > dims
    bad = torch.zeros((10, 3), dtype=torch.float32, device="cuda")
    bad[3, 1] = float("nan")
    with pytest.raises(_lib.GdsjError):
        sj.join(bad, 1.0, k=3)


def test_sorted_rows_option(sj):
    rng = np.random.default_rng(6)
    x = (rng.random((4000, 2)) * 10).astype(np.float32)
    c, ids, r = _gpu_join(sj, x, 0.8, k=2, sort_neighbors=1)
    offs = np.concatenate([[0], np.cumsum(c.astype(np.int64))])
    for i in range(0, len(x), 97):
        row = ids[offs[i]:offs[i + 1]]
        assert np.all(row[1:] > row[:-1])
    ref = O.join(x, 0.8, k=2)
    assert np.array_equal(ids, ref["ids"])  # sorted rows are exactly the oracle's CSR


@pytest.mark.parametrize("kind", ["float64_as_float", "int_list", "float32_2d"])
def test_gdsjoin_input_conversion(gpu, kind):
    # DATASET values go through float64 into DTYPE, as the reference's
    # np.asfarray + ctypes conversion does; floating arrays are cast directly
    rng = np.random.default_rng(12)
    x64 = rng.random((2500, 3)) * 10
    if kind == "float64_as_float":
        data, ref_x = x64, x64.astype(np.float32)
    elif kind == "int_list":
        xi = rng.integers(0, 40, (2500, 3))
        data, ref_x = xi.ravel().tolist(), xi.astype(np.float64).astype(np.float32)
    else:
        data, ref_x = x64.astype(np.float32), x64.astype(np.float32)
    p = _lib.default_params(k=3, ilp=0, shortcircuit=0, reorder=0)
    counts, table = gdsjoin(data, 1.5, 3, "float", params=p)
    ref = O.join(np.ascontiguousarray(ref_x), float(np.float32(1.5)), k=3)
    assert_csr_equal(counts, table, ref["counts"], ref["ids"])


def test_legacy_symbols_two_call_protocol(gpu):
    L = _lib.load()
    rng = np.random.default_rng(8)
    x = (rng.random((3000, 3)) * 10).astype(np.float32)
    L.gdsj_set_legacy_params(C.byref(_lib.default_params(k=3, ilp=0, shortcircuit=0, reorder=0)))
    counts = np.zeros(3000, np.uint32)
    L.GDSJoinPy(x.ctypes.data, 3000, C.c_float(0.7), 3, 0, counts.ctypes.data)
    assert L.gdsj_legacy_status() == 0
    table = np.zeros(int(counts.sum()), np.uint32)
    L.copyResultIntoPythonArray(table.ctypes.data, len(table))
    L.copyResultIntoPythonArray(table.ctypes.data, len(table))  # second call is a no-op
    ref = O.join(x, 0.7, k=3)
    assert_csr_equal(counts, table, ref["counts"], ref["ids"])
    # the pending table is held until copied: a shorter buffer takes its prefix
    L.gdsj_set_legacy_params(C.byref(_lib.default_params(k=3, ilp=0, shortcircuit=0, reorder=0, stamp=0)))
    L.GDSJoinPy(x.ctypes.data, 3000, C.c_float(0.7), 3, 0, counts.ctypes.data)
    full = np.zeros(int(counts.sum()), np.uint32)
    L.copyResultIntoPythonArray(full.ctypes.data, len(full))
    L.GDSJoinPy(x.ctypes.data, 3000, C.c_float(0.7), 3, 0, counts.ctypes.data)
    part = np.zeros(10, np.uint32)
    L.copyResultIntoPythonArray(part.ctypes.data, 10)
    assert L.gdsj_legacy_status() == 0
    assert np.array_equal(part, full[:10])
    assert_csr_equal(counts, full, ref["counts"], ref["ids"])
    L.gdsj_set_legacy_params(None)


def test_repeat_join_reuses_handle(sj):
    rng = np.random.default_rng(10)
    for n in (5000, 1200, 8000):
        x = (rng.random((n, 4)) * 10).astype(np.float32)
        _check(sj, x, 0.9, k=4, ilp=8, sc=1, reorder=1)


@pytest.mark.parametrize("d,k,dt,jitter", [(2, 2, np.float32, 0.0), (3, 3, np.float64, 0.0), (6, 6, np.float32, 0.0),
                                           (6, 6, np.float32, 1e-6), (4, 2, np.float64, 1e-12),
                                           (8, 8, np.float32, 0.0)])
def test_lattice_on_cell_faces(sj, d, k, dt, jitter):
    # points on a lattice of spacing eps: many pairs at distance exactly eps and
    # every point on a cell face, where classification rounding decides the cell
    rng = np.random.default_rng(d * 10 + k)
    side = {2: 60, 3: 16, 4: 7, 6: 5, 8: 3}[d]
    g = np.stack(np.meshgrid(*[np.arange(side)] * d, indexing="ij"), -1).reshape(-1, d)
    eps = 0.37
    x = (g * eps + 3.1).astype(np.float64)
    x += jitter * rng.standard_normal(x.shape)
    x = x.astype(dt)
    for ilp, sc, reorder in [(0, 0, 0), (8, 1, 1)]:
        _check(sj, x, float(np.asarray(eps, dtype=dt)), k=k, ilp=ilp, sc=sc, reorder=reorder)


@pytest.mark.parametrize("d", [3, 4, 6, 8])
def test_vfilter_on_off_same_table(sj, d):
    """The Gram-form prefilter of the plain FP32 kernels (every dim indexed)
    gives the oracle's table, on clustered data and on an integer lattice
    whose pairs at distance exactly eps all fall inside the filter's band
    (the reference's arithmetic decides them)."""
    rng = np.random.default_rng(90 + d)
    centers = rng.random((6, d)) * 12
    x = (centers[rng.integers(0, 6, 5000)] + rng.standard_normal((5000, d))).astype(np.float32)
    eps = _eps_for(x, 40)
    side = {3: 14, 4: 8, 6: 4, 8: 3}[d]
    grid = np.stack(np.meshgrid(*[np.arange(side, dtype=np.float32)] * d, indexing="ij"), -1).reshape(-1, d)
    for data, e in ((x, eps), (grid + np.float32(20.0), 1.0)):
        for vf in (1, 2):
            for kw in (dict(ilp=8, shortcircuit=1), dict(ilp=0, shortcircuit=0)):
                c, ids, r = _gpu_join(sj, data, e, k=d, reorder=0, vfilter=vf, **kw)
                ref = O.join(data, e, k=d, reorder=0, **kw)
                assert_csr_equal(c, ids, ref["counts"], ref["ids"])
            for est in (_lib.EST_LEARNED, _lib.EST_LEARNED):  # train, then the slotted pass
                c, ids, r = _gpu_join(sj, data, e, k=d, reorder=0, vfilter=vf, ilp=8, shortcircuit=1, estimator=est)
            ref = O.join(data, e, k=d, reorder=0, ilp=8, shortcircuit=1)
            assert_csr_equal(c, ids, ref["counts"], ref["ids"])


@pytest.mark.parametrize("d,k,dt", [(3, 3, np.float32), (6, 6, np.float32), (3, 2, np.float32),
                                    (8, 4, np.float32), (5, 3, np.float64), (4, 4, np.float64)])
def test_sort_matches_oracle(sj, d, k, dt):
    """SORT (sort dim 0 when every dim is indexed, else dim k): the GPU's
    table equals the oracle's on clustered data and on an integer lattice
    whose epsilon-boundary pairs SORT treats differently from the plain join,
    with the exact passes and with learned slots."""
    rng = np.random.default_rng(300 + d + k)
    centers = rng.random((6, d)) * 10
    x = (centers[rng.integers(0, 6, 4000)] + rng.standard_normal((4000, d))).astype(dt)
    eps = _eps_for(x, 30)
    side = {3: 9, 4: 6, 5: 4, 6: 3, 8: 3}[d]
    lat = (np.stack(np.meshgrid(*[np.arange(side, dtype=dt)] * d, indexing="ij"), -1).reshape(-1, d) + dt(10))
    for data, e, reorder in ((x, eps, 1), (np.ascontiguousarray(lat), 1.0, 0)):
        ref = O.join(data, e, k=k, ilp=8, shortcircuit=1, reorder=reorder, sort=1)
        c, ids, r = _gpu_join(sj, data, e, k=k, ilp=8, shortcircuit=1, reorder=reorder, sort=1, stamp=0)
        print("sort d=%d k=%d n=%d total=%d" % (d, k, len(data), r.total), flush=True)
        assert_csr_equal(c, ids, ref["counts"], ref["ids"])
        for est in (_lib.EST_LEARNED, _lib.EST_LEARNED):  # train, then the slotted pass
            c, ids, r = _gpu_join(sj, data, e, k=k, ilp=8, shortcircuit=1, reorder=reorder, sort=1, stamp=0,
                                  estimator=est)
        assert_csr_equal(c, ids, ref["counts"], ref["ids"])
