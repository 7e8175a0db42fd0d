"""The five workloads at full size (the measured configurations).

Full tables are too large to store, so each is compared point by point
through size-independent summaries:
- per-point counts and order-independent id hashes (sum of splitmix64 of
  the neighbor ids), against the oracle's, for the 2-D and 6-D workloads
  (the oracle runs the whole join);
- for the 32-D and 64-D workloads, whose full oracle join would take hours
  on the host, the oracle's counts and hashes of a seeded 1 % query sample
  that holds the points of the most occupied cells (the heaviest
  neighborhoods), plus byte equality of the MFMA-filtered table with the
  plain table at full size;
- the oracle's count of pairs on which the two readings of the ILP fold
  (kernel.cu:320-334: fused or rounded products added) disagree, printed per
  workload; for gmm6d the GPU's fused reading (ilp_fold = 1) is also checked
  against the oracle's table for that reading;
- CSR invariants: offsets are the prefix sum of counts, every point is its
  own neighbor, the relation is symmetric (sum of counts equals the number
  of (i, j) with i in row j, checked through hashes of transposed pairs on a
  sample).
The learned (slotted) path must give the same table as the exact path."""
import numpy as np
import pytest

from gpu_self_join_amd import _lib, datagen
from oracle import oracle as O
from tests.helpers import dense_sample, sorted_rows

# full-size oracle runs take minutes on the host
pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def _run(sj, pts, eps, **kw):
    r = sj.join(pts, eps, **kw)
    counts = r.counts.cpu().numpy().view(np.uint32)
    offsets = r.offsets.cpu().numpy()
    ids = r.ids.cpu().numpy().view(np.uint32)
    return counts, offsets, ids, r


def _check_csr(counts, offsets, ids):
    assert offsets[0] == 0 and np.array_equal(np.diff(offsets), counts.astype(np.int64))
    assert offsets[-1] == len(ids)
    assert np.all(counts >= 1)
    # every point is in its own row
    n = len(counts)
    rows = np.repeat(np.arange(n, dtype=np.uint32), counts)
    assert np.count_nonzero(rows == ids) == n


def _symmetric_sample(counts, offsets, ids, m=2000, seed=5):
    rng = np.random.default_rng(seed)
    n = len(counts)
    for i in rng.choice(n, m, replace=False):
        for j in ids[offsets[i]:offsets[i + 1]][:8]:
            row = ids[offsets[j]:offsets[j + 1]]
            assert i in row, (i, j)


def _workload(name):
    import torch
    x = datagen.generate(name)
    eps = datagen.workload_eps(name)
    p = datagen.workload_params(name)
    return x, torch.from_numpy(x).cuda(), eps, p


@pytest.mark.parametrize("name", ["uniform2d", "gmm6d"])
def test_full_size_table_equals_oracle(gpu, name):
    from gpu_self_join_amd import SelfJoin
    x, pts, eps, p = _workload(name)
    kw = dict(k=p["k"], ilp=p["ilp"], shortcircuit=p["shortcircuit"], reorder=p["reorder"], stamp=p["stamp"],
              queryreorder=p["queryreorder"])
    sj = SelfJoin(0)
    try:
        counts, offsets, ids, r = _run(sj, pts, eps, **kw)
        print(name, "pairs", r.total, flush=True)
        _check_csr(counts, offsets, ids)
        print(name, "oracle join (default fold reading)", flush=True)
        ref = O.join(x, eps, k=p["k"], ilp=p["ilp"], shortcircuit=p["shortcircuit"], reorder=p["reorder"],
                     with_ids=False, nthreads=16)
        print(name, "ILP fold ambiguous pairs", int(ref["stats"][3]), flush=True)
        assert np.array_equal(counts, ref["counts"])
        assert np.array_equal(O.csr_hashes(counts, ids), ref["hashes"])
        _symmetric_sample(counts, offsets, ids)
        if p["ilp"] > 0 and p["shortcircuit"]:
            # the fused reading of the fold, on the GPU and in the oracle
            cf, of, idf, rf = _run(sj, pts, eps, ilp_fold=1, **kw)
            print(name, "oracle join (fused fold reading)", flush=True)
            reff = O.join(x, eps, k=p["k"], ilp=p["ilp"], shortcircuit=p["shortcircuit"], reorder=p["reorder"],
                          fold_fused=1, with_ids=False, nthreads=16)
            assert np.array_equal(cf, reff["counts"])
            assert np.array_equal(O.csr_hashes(cf, idf), reff["hashes"])
            ndiff = int(np.abs(cf.astype(np.int64) - counts.astype(np.int64)).sum())
            assert ndiff <= int(ref["stats"][3])
        # the learned single pass writes the same table
        sl = SelfJoin(0)
        try:
            _run(sl, pts, eps, estimator=_lib.EST_LEARNED, **kw)  # exact + training
            c2, o2, i2, r2 = _run(sl, pts, eps, estimator=_lib.EST_LEARNED, **kw)
            assert r2.stats["est_tiles"] > 0
            assert np.array_equal(c2, counts)
            if p["stamp"]:
                # unicomp (exact passes) orders the reverse entries of a row by
                # tile completion; the slotted pass evaluates both directions
                assert np.array_equal(sorted_rows(c2, i2), sorted_rows(counts, ids))
            else:
                assert np.array_equal(i2, ids)
        finally:
            sl.close()
    finally:
        sj.close()


@pytest.mark.parametrize("name", ["aniso32d", "lowid64d"])
def test_full_size_filtered_table_equals_plain_and_oracle_sample(gpu, name):
    from gpu_self_join_amd import SelfJoin
    x, pts, eps, p = _workload(name)
    kw = dict(k=p["k"], ilp=p["ilp"], shortcircuit=p["shortcircuit"], reorder=p["reorder"], stamp=p["stamp"],
              queryreorder=p["queryreorder"])
    sj = SelfJoin(0)
    try:
        cg, og, ig, rg = _run(sj, pts, eps, filter=_lib.FILTER_GRAM, **kw)
        st = rg.stats
        print(name, "pairs", rg.total, "rechecks", st["filter_rechecks"], "hits", st["filter_hits"],
              "max filter err", st["filter_max_err"], flush=True)
        assert st["filter_hits"] == rg.total
        _check_csr(cg, og, ig)
        cp, op, ip, rp = _run(sj, pts, eps, filter=_lib.FILTER_PLAIN, **kw)
        assert np.array_equal(cg, cp) and np.array_equal(ig, ip)
        q = dense_sample(x, eps, p, frac=0.01)
        print(name, "oracle sample join", flush=True)
        ref = O.join_subset(x, eps, p["k"], q, ilp=p["ilp"], shortcircuit=p["shortcircuit"], reorder=p["reorder"],
                            nthreads=16)
        print(name, "oracle sample", len(q), "queries,", int(ref["stats"][1]), "candidates; ILP fold ambiguous pairs",
              int(ref["stats"][3]), flush=True)
        assert np.array_equal(cg[q], ref["counts"])
        h = O.csr_hashes(cg, ig)
        assert np.array_equal(h[q], ref["hashes"])
        _symmetric_sample(cg, og, ig, m=500)
    finally:
        sj.close()
