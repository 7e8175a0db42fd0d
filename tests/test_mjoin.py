"""The matrix-core join of low-dimensional data (gdsj_params.mfma): every dim
indexed, 3 to 7 dims, FP32 MFMA on the Gram form in tile-local coordinates,
pairs within the proven error band decided by the reference's arithmetic.

Bar: the oracle's table (counts and sorted rows), and rows identical in
order to the plain kernels' (mfma = 2), for every pass (exact counting and
fill, learned slots, the sampling estimator's slots), on clustered data, on
integer lattices whose pairs at distance exactly eps sit in the band, and on
a cluster where every pair is a neighbor (every block queues 256 hits)."""
import numpy as np
import pytest

from gpu_self_join_amd import _lib
from oracle import oracle as O
from tests.helpers import assert_csr_equal

pytestmark = pytest.mark.gpu


def _run(sj, x, eps, **kw):
    import torch
    r = sj.join(torch.from_numpy(np.ascontiguousarray(x)).cuda(), eps, **kw)
    counts = r.counts.cpu().numpy().view(np.uint32)
    ids = r.ids.cpu().numpy().view(np.uint32)
    return counts, ids, r


def _eps_for(x, nb):
    rng = np.random.default_rng(0)
    q = x[rng.choice(len(x), 64, replace=False)].astype(np.float64)
    d = np.sqrt(((q[:, None, :] - x[None, :, :].astype(np.float64)) ** 2).sum(-1))
    return float(np.asarray(np.sort(d, axis=1)[:, min(nb, len(x) - 1)].mean(), dtype=x.dtype))


def _clustered(n, d, dt, seed):
    rng = np.random.default_rng(seed)
    centers = rng.random((6, d)) * 14
    return (centers[rng.integers(0, 6, n)] + rng.standard_normal((n, d))).astype(dt)


def _lattice(d, dt):
    side = {3: 12, 4: 7, 5: 5, 6: 4, 7: 3}[d]
    g = np.stack(np.meshgrid(*[np.arange(side, dtype=np.float64)] * d, indexing="ij"), -1).reshape(-1, d)
    return np.ascontiguousarray((g + 17.0).astype(dt))


MODES = [dict(ilp=8, shortcircuit=1, reorder=1), dict(ilp=0, shortcircuit=0, reorder=0),
         dict(ilp=8, shortcircuit=0, reorder=1)]


@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("d", [3, 4, 5, 6, 7])
def test_mjoin_equals_plain_and_oracle(sj, d, dt):
    x = _clustered(6000, d, dt, 500 + d)
    eps = _eps_for(x, 48)
    lat = _lattice(d, dt)
    for data, e in ((x, eps), (lat, 1.0)):
        for kw in MODES:
            ref = O.join(data, e, k=d, **kw)
            cp, ip, rp = _run(sj, data, e, k=d, mfma=2, **kw)
            assert rp.stats["mfma_tiles"] == 0
            _, _, ra = _run(sj, data, e, k=d, mfma=0, **kw)
            assert ra.stats["mfma_tiles"] == 0  # auto: off
            for mf in (1, 8):
                c, ids, r = _run(sj, data, e, k=d, mfma=mf, **kw)
                print("d=%d %s n=%d mfma=%d %s total=%d mfma_tiles=%d" % (
                    d, np.dtype(dt).name, len(data), mf, kw, r.total, r.stats["mfma_tiles"]), flush=True)
                if mf == 1:
                    assert r.stats["mfma_tiles"] == r.stats["n_tiles"]
                assert_csr_equal(c, ids, ref["counts"], ref["ids"])
                assert np.array_equal(c, cp) and np.array_equal(ids, ip)


@pytest.mark.parametrize("d", [3, 6, 7])
def test_mjoin_slotted_passes(sj, d):
    """Learned slots (train on the exact passes, then the slotted pass) and
    the sampling estimator's slots, every tile on the matrix cores."""
    from gpu_self_join_amd import SelfJoin
    x = _clustered(8000, d, np.float32, 700 + d)
    eps = _eps_for(x, 60)
    ref = O.join(x, eps, k=d, ilp=8, shortcircuit=1, reorder=1)
    kw = dict(k=d, ilp=8, shortcircuit=1, reorder=1, mfma=1)
    cp, ip, _ = _run(sj, x, eps, k=d, ilp=8, shortcircuit=1, reorder=1, mfma=2)
    for est in (_lib.EST_LEARNED, _lib.EST_SAMPLE):
        s2 = SelfJoin(0)
        try:
            _run(s2, x, eps, estimator=est, **kw)
            c, ids, r = _run(s2, x, eps, estimator=est, **kw)
            print("d=%d est=%d total=%d overflow tiles=%d" % (d, est, r.total, r.stats["est_overflow_tiles"]),
                  flush=True)
            assert r.stats["est_tiles"] > 0 and r.stats["mfma_tiles"] > 0
            assert_csr_equal(c, ids, ref["counts"], ref["ids"])
            assert np.array_equal(ids, ip)
        finally:
            s2.close()


@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_mjoin_every_pair_a_neighbor(sj, dt):
    """3000 points within a ball of radius eps / 4: every pair is a
    neighbor (9 M pairs), every block of 16 x 16 pairs queues 256 hits."""
    rng = np.random.default_rng(11)
    d = 5
    x = (rng.random((3000, d)) * 0.1 + 40.0).astype(dt)
    eps = 1.0
    for kw in (dict(ilp=8, shortcircuit=1), dict(ilp=0, shortcircuit=0)):
        c, ids, r = _run(sj, x, eps, k=d, reorder=0, mfma=1, **kw)
        assert r.total == 3000 * 3000
        assert np.all(c == 3000)
        cp, ip, _ = _run(sj, x, eps, k=d, reorder=0, mfma=2, **kw)
        assert np.array_equal(ids, ip)


def test_mjoin_rejoin_and_small_tiles(sj):
    """Tiles of every size up to 64 queries (mfma = 1), occupancy
    thresholds of 3 and 16 points per cell, and a rejoin over the same
    grid."""
    x = _clustered(20000, 6, np.float32, 91)
    eps = _eps_for(x, 30)
    ref = O.join(x, eps, k=6, ilp=8, shortcircuit=1, reorder=1)
    for mf in (1, 3, 16):
        c, ids, r = _run(sj, x, eps, k=6, ilp=8, shortcircuit=1, reorder=1, mfma=mf)
        assert_csr_equal(c, ids, ref["counts"], ref["ids"])
        assert r.stats["mfma_tiles"] > 0
    r2 = sj.rejoin(n=len(x))
    assert r2.stats["mfma_tiles"] == r.stats["mfma_tiles"]
    assert np.array_equal(r2.counts.cpu().numpy().view(np.uint32), c)
    assert np.array_equal(r2.ids.cpu().numpy().view(np.uint32), ids)
