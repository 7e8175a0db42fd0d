"""The Gram filter on principal axes (gdsj_params.gram_axes): the fp16
operands hold each point's coordinates on 16 or 32 principal axes of the
REORDER sample instead of all its dims, with eps' widened by the basis'
spectral norm and the bound by the projection's rounding.  The filter only
drops pairs farther apart than eps'; every kept pair is rechecked by the
reference's arithmetic (kernel.cu:307-365), so the table cannot change.

Bar: the oracle's table, and rows identical in order to the Gram filter on
the points' own dims (gram_axes = 1), on data near a rotated subspace (where
the axes hold nearly all of a pair's distance), on full-rank data with the
axes forced (a weak filter that must still keep every pair), on pairs placed
at distance eps (one part in 1e7 either side) along the subspace, in FP64
and FP32, for the exact and the learned passes."""
import numpy as np
import pytest

from gpu_self_join_amd import _lib
from oracle import oracle as O
from tests.helpers import assert_csr_equal

pytestmark = pytest.mark.gpu


def _run(sj, x, eps, **kw):
    import torch
    r = sj.join(torch.from_numpy(np.ascontiguousarray(x)).cuda(), eps, **kw)
    return r.counts.cpu().numpy().view(np.uint32), r.ids.cpu().numpy().view(np.uint32), r


def _subspace(n, d, r, seed, noise=0.05, offset=3.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((d, r)))
    x = rng.standard_normal((n, r)) * 10.0 @ q.T + rng.standard_normal((n, d)) * noise + offset
    return x, q


def _eps_for(x, nb):
    rng = np.random.default_rng(0)
    q = x[rng.choice(len(x), 64, replace=False)].astype(np.float64)
    d = np.sqrt(((q[:, None, :] - x[None, :, :].astype(np.float64)) ** 2).sum(-1))
    return float(np.asarray(np.sort(d, axis=1)[:, min(nb, len(x) - 1)].mean(), dtype=x.dtype))


BASE = dict(ilp=8, shortcircuit=1, reorder=1, filter=_lib.FILTER_GRAM)


def _check_axes(sj, x, eps, k, want_auto, **kw):
    ref = O.join(x, eps, k=k, ilp=8, shortcircuit=1, reorder=1)
    c1, i1, r1 = _run(sj, x, eps, k=k, gram_axes=1, **BASE, **kw)
    assert r1.stats["gram_axes"] == 0
    assert_csr_equal(c1, i1, ref["counts"], ref["ids"])
    d = x.shape[1]
    out = {}
    for ax in (16, 32, 0):
        c, ids, r = _run(sj, x, eps, k=k, gram_axes=ax, **BASE, **kw)
        used = r.stats["gram_axes"]
        print("n=%d d=%d %s gram_axes=%d used=%d rechecks %d (own dims %d) total %d" % (
            len(x), d, x.dtype.name, ax, used, r.stats["filter_rechecks"], r1.stats["filter_rechecks"], r.total),
            flush=True)
        if ax == 0:
            assert used == want_auto
        else:
            assert used == (ax if ax < (d + 15) // 16 * 16 else 0)
        assert_csr_equal(c, ids, ref["counts"], ref["ids"])
        assert np.array_equal(c, c1) and np.array_equal(ids, i1)
        out[ax] = r
    return r1, out


@pytest.mark.parametrize("dt", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("d,r,k", [(64, 8, 8), (32, 6, 6)])
def test_axes_near_subspace(sj, d, r, k, dt):
    x, _ = _subspace(12000, d, r, 300 + d)
    x = np.ascontiguousarray(x.astype(dt))
    eps = _eps_for(x, 40)
    r1, out = _check_axes(sj, x, eps, k, 16)
    # the 16 axes hold nearly all of a pair's distance: about as few rechecks
    assert out[16].stats["filter_rechecks"] <= 1.2 * r1.stats["filter_rechecks"] + 1000


@pytest.mark.parametrize("dt", [np.float64, np.float32], ids=["f64", "f32"])
def test_axes_forced_on_full_rank_data(sj, dt):
    """Anisotropic full-rank 48-D data: 16 axes leave much of the distance
    out (auto keeps the own dims), so the forced filter keeps many more
    pairs; the table must not change."""
    rng = np.random.default_rng(21)
    scale = np.linspace(3.0, 0.5, 48)
    x = np.ascontiguousarray((rng.standard_normal((8000, 48)) * scale + 100.0).astype(dt))
    eps = _eps_for(x, 25)
    r1, out = _check_axes(sj, x, eps, 8, 0)
    assert out[16].stats["filter_rechecks"] >= r1.stats["filter_rechecks"]


@pytest.mark.parametrize("dt", [np.float64, np.float32], ids=["f64", "f32"])
def test_axes_pairs_at_eps(sj, dt):
    """Partners at distance eps (1 -+ 1e-7) along directions inside the
    subspace, where the axes hold the whole distance and the pair sits at
    the threshold: the band must keep every pair the reference accepts."""
    rng = np.random.default_rng(5)
    d, r = 64, 8
    x, q = _subspace(6000, d, r, 77, offset=40.0)
    eps = 4.0
    base = x[:3000]
    u = rng.standard_normal((3000, r))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    f = np.where(rng.random(3000) < 0.5, 1.0 - 1e-7, 1.0 + 1e-7)
    part = base + (u @ q.T) * (eps * f)[:, None]
    x = np.ascontiguousarray(np.concatenate([x, part]).astype(dt))
    ref = O.join(x, eps, k=8, ilp=8, shortcircuit=1, reorder=1)
    c1, i1, _ = _run(sj, x, eps, k=8, gram_axes=1, **BASE)
    assert_csr_equal(c1, i1, ref["counts"], ref["ids"])
    for ax in (16, 32):
        c, ids, r2 = _run(sj, x, eps, k=8, gram_axes=ax, **BASE)
        print("%s gram_axes=%d used=%d total=%d" % (np.dtype(dt).name, ax, r2.stats["gram_axes"], r2.total),
              flush=True)
        assert r2.stats["gram_axes"] == ax
        assert_csr_equal(c, ids, ref["counts"], ref["ids"])
        assert np.array_equal(ids, i1)


def test_axes_learned_slots_index_and_rejoin(sj):
    """Learned slots (exact joins that train, then slotted ones) with the
    operands on 16 axes, on the principal-axes index and on the reference's
    dims, then a rejoin over the same grid."""
    from gpu_self_join_amd import SelfJoin
    x, _ = _subspace(16000, 64, 8, 9)
    eps = _eps_for(x, 40)
    ref = O.join(x, eps, k=8, ilp=8, shortcircuit=1, reorder=1)
    for idx in (0, 1):
        s2 = SelfJoin(0)
        try:
            for est in (_lib.EST_EXACT, _lib.EST_LEARNED, _lib.EST_LEARNED):
                c, ids, r = _run(s2, x, eps, k=8, index=idx, estimator=est, gram_axes=16, **BASE)
                assert r.stats["gram_axes"] == 16
                assert r.stats["index_used"] == (2 if idx == 0 else 1)
                assert_csr_equal(c, ids, ref["counts"], ref["ids"])
            r2 = s2.rejoin(n=len(x))
            assert_csr_equal(r2.counts.cpu().numpy().view(np.uint32), r2.ids.cpu().numpy().view(np.uint32),
                             ref["counts"], ref["ids"])
        finally:
            s2.close()


def test_axes_param_validation(sj):
    import torch
    x = torch.zeros((100, 32), dtype=torch.float64).cuda()
    with pytest.raises(Exception, match="gram_axes"):
        sj.join(x, 1.0, k=4, gram_axes=8)
