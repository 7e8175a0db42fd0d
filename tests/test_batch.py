"""Batched joins (gdsj_batch_*): many independent self-joins of uneven n,
dims and epsilon in one persistent launch per pass, replayed as a hipGraph.
Every instance's table must equal the oracle's for that instance alone (and
the single-join API's), whatever the autotuner picks and whether the learned
slots or the exact passes produced it."""
import numpy as np
import pytest

from gpu_self_join_amd import _lib
from oracle import oracle as O
from tests.helpers import assert_csr_equal

pytestmark = pytest.mark.gpu


def _gmm(n, d, seed, ncl=6, sigma=1.0):
    rng = np.random.default_rng(seed)
    c = rng.random((ncl, d)) * 20
    return (c[rng.integers(0, ncl, n)] + sigma * rng.standard_normal((n, d))).astype(np.float32)


# (n, dims, eps): specialised dims 2/4/8/16, generic dims 3 and 5, an empty
# and a one-point instance
SMALL = [(4000, 2, 0.35), (0, 4, 1.0), (9000, 4, 0.9), (1, 8, 1.0), (3000, 8, 1.8), (2500, 16, 3.0),
         (6000, 3, 0.6), (1500, 5, 1.2), (12000, 2, 0.2), (700, 16, 2.5), (5000, 8, 1.5), (20000, 4, 0.6)]


def _instances(cases, dtype=np.float32):
    return [_gmm(n, d, 100 + i).astype(dtype) for i, (n, d, _) in enumerate(cases)]


def _snap(r):
    """Copies of a result's tensors, then the batch's buffers are cleared so
    that a replay has to write everything again."""
    out = (r.counts.clone(), r.offsets.clone(), r.ids.clone())
    r.counts.zero_()
    r.offsets.zero_()
    r.ids.zero_()
    return out


def _same(r, snap):
    return (torch_equal(r.counts, snap[0]) and torch_equal(r.offsets, snap[1]) and torch_equal(r.ids, snap[2]))


def torch_equal(a, b):
    import torch
    return a.shape == b.shape and bool(torch.equal(a, b))


def _check_against_oracle(res, xs, cases, params):
    for i, (x, (n, d, e)) in enumerate(zip(xs, cases)):
        c, off, ids = res.instance(i)
        c = c.cpu().numpy().view(np.uint32)
        ids = ids.cpu().numpy().view(np.uint32)
        assert len(c) == n
        if n == 0:
            continue
        eps = float(np.asarray(e, dtype=x.dtype))
        ref = O.join(x, eps, k=min(d, 6), ilp=params["ilp"], shortcircuit=params["shortcircuit"],
                     reorder=params["reorder"])
        print("instance", i, "n", n, "d", d, "pairs", int(c.sum()), flush=True)
        assert_csr_equal(c, ids, ref["counts"], ref["ids"])
        o = off.cpu().numpy()
        assert o[0] == 0 and np.array_equal(np.diff(o), c.astype(np.int64))


@pytest.mark.parametrize("ilp,sc,reorder", [(8, 1, 1), (0, 0, 0), (4, 1, 0), (4, 0, 1)])
def test_batch_matches_oracle_per_instance(gpu, ilp, sc, reorder):
    import torch

    from gpu_self_join_amd import BatchJoin
    xs = _instances(SMALL)
    pts = [torch.from_numpy(x).cuda() for x in xs]
    bj = BatchJoin(pts, [e for _, _, e in SMALL], ilp=ilp, shortcircuit=sc, reorder=reorder)
    try:
        r1 = bj.run()  # grows the id buffer: captures for the small and the grown buffer
        _check_against_oracle(r1, xs, SMALL, dict(ilp=ilp, shortcircuit=sc, reorder=reorder))
        caps = bj.stats()["graph_captures"]
        snap = _snap(r1)
        r2 = bj.run()  # graph replay into the cleared buffers
        assert r2.total == r1.total and _same(r2, snap)
        assert bj.stats()["graph_captures"] == caps
    finally:
        bj.close()


def test_batch_autotuned_ilp_and_replays(gpu):
    import torch

    from gpu_self_join_amd import BatchJoin
    xs = _instances(SMALL)
    pts = [torch.from_numpy(x).cuda() for x in xs]
    bj = BatchJoin(pts, [e for _, _, e in SMALL], ilp=_lib.ILP_AUTO)
    try:
        t = bj.autotune()
        print("tuning", t, flush=True)
        assert t["ilp"] in (0, 8) and t["tile_queries"] in (16, 32, 64) and t["waves_per_block"] in (1, 2, 4)
        r = bj.run()
        # the chosen ILP is the result contract
        _check_against_oracle(r, xs, SMALL, dict(ilp=t["ilp"], shortcircuit=1, reorder=1))
        snap = _snap(r)
        for _ in range(3):
            r = bj.run()
            assert _same(r, snap)
            _snap(r)
        st = bj.stats()
        assert st["graph_captures"] <= 2  # the first run may grow the id buffer once
    finally:
        bj.close()


def test_batch_learned_slots_match_exact(gpu):
    import torch

    from gpu_self_join_amd import BatchJoin
    cases = [(30000, 2, 0.25), (20000, 4, 0.8), (15000, 8, 1.6), (8000, 16, 3.0), (25000, 3, 0.5)] * 2
    xs = _instances(cases)
    pts = [torch.from_numpy(x).cuda() for x in xs]
    eps = [e for _, _, e in cases]
    ex = BatchJoin(pts, eps)
    bl = BatchJoin(pts, eps, estimator=_lib.EST_LEARNED)
    try:
        ref = _snap(ex.run())
        r1 = bl.run()  # exact, then training
        assert r1.total == ref[2].numel() and _same(r1, ref)
        _snap(r1)
        r2 = bl.run()  # slotted single pass
        st = bl.stats()
        print("learned batch stats", st, flush=True)
        assert st["slotted"] == 1 and st["est_capacity"] > 0
        _check_against_oracle(r2, xs[:5], cases[:5], dict(ilp=8, shortcircuit=1, reorder=1))
        assert _same(r2, ref)
        _snap(r2)
        r3 = bl.run()  # replay of the slotted graph
        assert _same(r3, ref)
        st3 = bl.stats()
        assert st3["graph_captures"] == st["graph_captures"]  # replayed, not captured again
    finally:
        ex.close()
        bl.close()


def test_batch_small_output_buffer_reports_overflow(gpu):
    import torch

    from gpu_self_join_amd import BatchJoin
    xs = _instances(SMALL[:4])
    pts = [torch.from_numpy(x).cuda() for x in xs]
    bj = BatchJoin(pts, [e for _, _, e in SMALL[:4]])
    try:
        ids = torch.empty(10, dtype=torch.int32, device="cuda")
        r = bj.run(out_ids=ids)  # too small: reported, then grown and rerun
        snap = _snap(r)
        full = bj.run()
        assert full.total == r.total and _same(full, snap)
    finally:
        bj.close()


def test_batch_rejects_bad_arguments(gpu):
    import torch

    from gpu_self_join_amd import BatchJoin
    x = torch.zeros((10, 4), device="cuda")
    with pytest.raises(_lib.GdsjError):
        BatchJoin([x], [0.5], k=7)  # k > dims
    with pytest.raises(_lib.GdsjError):
        BatchJoin([x], [0.5], filter=_lib.FILTER_GRAM)
    with pytest.raises(_lib.GdsjError):
        BatchJoin([x], [0.5], estimator=_lib.EST_SAMPLE)


@pytest.mark.timeout(900)
def test_batched_workload_equals_single_joins(gpu):
    """The 256-instance workload: every instance's rows equal a single join's,
    and every instance's counts and per-point neighbor-id hashes equal the
    oracle's (the oracle's ILP fold ambiguity count is printed)."""
    import torch

    from gpu_self_join_amd import BatchJoin, SelfJoin, datagen
    m = datagen.BATCHED_COUNT
    xs, eps, host = [], [], []
    for i in range(m):
        x, _ = datagen.batched_instance(i)
        host.append(x)
        xs.append(torch.from_numpy(x).cuda())
        eps.append(datagen.batched_eps(i))
    bj = BatchJoin(xs, eps, estimator=_lib.EST_LEARNED)
    sj = SelfJoin(0)
    amb = 0
    try:
        bj.run()
        r = bj.run()
        print("batched stats", bj.stats(), flush=True)
        for i in range(m):
            c, off, ids = r.instance(i)
            s = sj.join(xs[i], eps[i], filter=_lib.FILTER_PLAIN, index=1)  # the batch's index
            assert torch.equal(c, s.counts), i
            assert torch.equal(ids, s.ids), i
            x = host[i]
            ref = O.join(x, eps[i], k=min(x.shape[1], 6), ilp=8, shortcircuit=1, reorder=1, with_ids=False,
                         nthreads=16)
            cn = c.cpu().numpy().view(np.uint32)
            assert np.array_equal(cn, ref["counts"]), i
            assert np.array_equal(O.csr_hashes(cn, ids.cpu().numpy().view(np.uint32)), ref["hashes"]), i
            amb += int(ref["stats"][3])
            if i % 16 == 15:
                print("batched: instances 0..%d equal the oracle" % i, flush=True)
        print("batched: all %d instances equal the oracle; ILP fold ambiguous pairs %d" % (m, amb), flush=True)
    finally:
        bj.close()
        sj.close()
