"""The principal-axes index (gdsj_params.index): the grid is built on the
top k principal axes of the REORDER sample instead of the reference's k dims
of largest variance (main.cu:981-1075), with cells slightly wider than eps,
and used only when no point lies within rounding of a face of the
reference's own grid.  The candidate set changes, the pair test does not.

Bar: the oracle's table (the reference's grid, dims and arithmetic) for the
plain kernels, the Gram filter and learned slots, on data near a rotated
subspace (where the principal axes prune) in FP64 and FP32; the fallback to
the reference's dims on lattices whose coordinates sit on cell faces."""
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


def _subspace(n, d, r, dt, seed, noise=0.05):
    """n points on a random r-dim linear subspace of R^d plus isotropic noise."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((d, r)))
    x = rng.standard_normal((n, r)) * 10.0 @ q.T + rng.standard_normal((n, d)) * noise + 3.0
    return np.ascontiguousarray(x.astype(dt))


def _eps_for(x, nb):
    rng = np.random.default_rng(0)
    q = x[rng.choice(len(x), 64, replace=False)].astype(np.float64)
    d = np.sqrt(((q[:, None, :] - x[None, :, :].astype(np.float64)) ** 2).sum(-1))
    return float(np.asarray(np.sort(d, axis=1)[:, min(nb, len(x) - 1)].mean(), dtype=x.dtype))


@pytest.mark.parametrize("dt", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("d,k", [(32, 6), (20, 4)])
def test_principal_axes_equal_oracle(sj, d, k, dt):
    x = _subspace(12000, d, k, dt, 100 + d)
    eps = _eps_for(x, 30)
    ref = O.join(x, eps, k=k, ilp=8, shortcircuit=1, reorder=1)
    base = dict(k=k, ilp=8, shortcircuit=1, reorder=1)
    c1, i1, r1 = _run(sj, x, eps, index=1, **base)
    assert r1.stats["index_used"] == 1
    assert_csr_equal(c1, i1, ref["counts"], ref["ids"])
    for filt in (_lib.FILTER_PLAIN, _lib.FILTER_GRAM):
        for idx in (2, 0):
            c, ids, r = _run(sj, x, eps, index=idx, filter=filt, **base)
            print("d=%d k=%d %s filter=%d index=%d used=%d risky=%d candidates %d (reference dims %d)" % (
                d, k, np.dtype(dt).name, filt, idx, r.stats["index_used"], r.stats["index_risky"],
                r.stats["candidates"], r1.stats["candidates"]), flush=True)
            assert_csr_equal(c, ids, ref["counts"], ref["ids"])
            if dt == np.float64:
                # no point within ~1e-14 of a face: the principal axes are used
                assert r.stats["index_used"] == 2 and r.stats["index_risky"] == 0
                assert r.stats["candidates"] < r1.stats["candidates"] * 0.75
            else:
                assert r.stats["index_used"] == (2 if r.stats["index_risky"] == 0 else 1)


def test_principal_axes_learned_slots_and_rejoin(sj):
    from gpu_self_join_amd import SelfJoin
    x = _subspace(15000, 24, 5, np.float64, 7)
    eps = _eps_for(x, 40)
    ref = O.join(x, eps, k=5, ilp=8, shortcircuit=1, reorder=1)
    s2 = SelfJoin(0)
    try:
        for est in (_lib.EST_EXACT, _lib.EST_LEARNED, _lib.EST_LEARNED):
            c, ids, r = _run(s2, x, eps, k=5, ilp=8, shortcircuit=1, reorder=1, index=2, estimator=est)
            assert r.stats["index_used"] == 2
            assert_csr_equal(c, ids, ref["counts"], ref["ids"])
        r2 = s2.rejoin(n=len(x))
        assert_csr_equal(r2.counts.cpu().numpy().view(np.uint32), r2.ids.cpu().numpy().view(np.uint32),
                         ref["counts"], ref["ids"])
    finally:
        s2.close()


def test_faces_fall_back_to_reference_dims(sj):
    """Points within rounding below a face of the reference's grid
    (quotients (x - min) / eps just under an integer m + 1) with points on
    the face two cells up (quotients m + 2) in the same dim: the reference's
    dims are used.  Points on faces alone (an integer lattice) are not at
    risk."""
    rng = np.random.default_rng(3)
    x = rng.integers(0, 4, size=(6000, 10)).astype(np.float64)
    x[:, :3] *= 3.0  # variance concentrated in a few dims
    x[0] = 0.0       # the minimum: min - eps = -1 in every dim
    ref = O.join(x, 1.0, k=3, ilp=8, shortcircuit=1, reorder=1)
    c, ids, r = _run(sj, x, 1.0, k=3, ilp=8, shortcircuit=1, reorder=1, index=2)
    assert r.stats["index_used"] == 2 and r.stats["index_risky"] == 0  # on faces only
    assert_csr_equal(c, ids, ref["counts"], ref["ids"])
    y = x.copy()
    y[1:40, 0] = 6.0 - 2.0 ** -50  # quotient 7 - 2^-50 in the dim of largest variance
    y[40:80, 0] = 7.0             # quotient 8: on the face two cells up
    ref = O.join(y, 1.0, k=3, ilp=8, shortcircuit=1, reorder=1)
    c, ids, r = _run(sj, y, 1.0, k=3, ilp=8, shortcircuit=1, reorder=1, index=2)
    assert r.stats["index_used"] == 1 and r.stats["index_risky"] > 0
    assert_csr_equal(c, ids, ref["counts"], ref["ids"])


def test_auto_index(sj):
    """Auto takes the principal axes on data near a rotated subspace and the
    reference's dims on axis-aligned anisotropic data (whose principal axes
    are the coordinate axes); both tables equal the oracle's."""
    x = _subspace(8000, 16, 3, np.float64, 11)
    eps = _eps_for(x, 20)
    c, ids, r = _run(sj, x, eps, k=3, ilp=8, shortcircuit=1, reorder=1)
    assert r.stats["index_used"] == 2
    ref = O.join(x, eps, k=3, ilp=8, shortcircuit=1, reorder=1)
    assert_csr_equal(c, ids, ref["counts"], ref["ids"])
    rng = np.random.default_rng(12)
    y = (rng.standard_normal((8000, 16)) * (4.0 * 0.8 ** np.arange(16))).astype(np.float64)
    e2 = _eps_for(y, 20)
    c, ids, r = _run(sj, y, e2, k=3, ilp=8, shortcircuit=1, reorder=1)
    assert r.stats["index_used"] == 1
    ref = O.join(y, e2, k=3, ilp=8, shortcircuit=1, reorder=1)
    assert_csr_equal(c, ids, ref["counts"], ref["ids"])
