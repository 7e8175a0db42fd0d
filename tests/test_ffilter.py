"""The FMA-chain filter of the plain kernels' counting passes
(gdsj_params.ffilter): with ILP > 0 each pair is first decided from its
squared differences summed by one FMA per dim, and a 64-pair chunk holding a
pair within the proven band around the threshold is evaluated in the
reference's order (kernel.cu:307-357).

Bar: the oracle's table, and rows identical in order to those of
ffilter = 2 (every pair by the reference's arithmetic), on data built to
put many pairs within a few ulps of eps (pairs placed at distance eps along
random directions, integer lattices with eps = 1 where the two sums agree
exactly at the threshold), for both ILP orders (SHORTCIRCUIT 1 with its
early exit between ILP groups, SHORTCIRCUIT 0 with ILP partials)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import assert_csr_equal

pytestmark = pytest.mark.gpu


def _run(sj, x, eps, **kw):
    import torch
    r = sj.join(torch.from_numpy(np.ascontiguousarray(x)).cuda(), eps, **kw)
    return r.counts.cpu().numpy().view(np.uint32), r.ids.cpu().numpy().view(np.uint32), r


def _shell_pairs(n, d, dt, eps, seed):
    """n base points, each with 3 partners at distance eps (rounded to dt)
    along random directions, spread so that clusters do not overlap much."""
    rng = np.random.default_rng(seed)
    base = rng.random((n, d)) * (6.0 * eps * n ** (1.0 / d)) + 20.0
    pts = [base]
    for _ in range(3):
        v = rng.standard_normal((n, d))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        pts.append(base + eps * v)
    return np.concatenate(pts).astype(dt)


def _lattice(d, dt):
    side = {2: 40, 3: 12, 6: 4, 9: 3, 16: 2}[d]
    g = np.stack(np.meshgrid(*[np.arange(side, dtype=np.float64)] * d, indexing="ij"), -1).reshape(-1, d)
    return np.ascontiguousarray((g * 0.5 + 3.0).astype(dt))


MODES = [dict(ilp=8, shortcircuit=1), dict(ilp=4, shortcircuit=0), dict(ilp=2, shortcircuit=1)]


@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("d", [3, 6, 16])
def test_ffilter_equals_reference_order_and_oracle(sj, d, dt):
    eps = 0.75
    shell = _shell_pairs(4000 if d < 16 else 2000, d, dt, eps, 40 + d)
    lat = _lattice(d, dt)
    for data, e in ((shell, eps), (lat, 0.5)):
        k = min(d, 6)
        for kw in MODES:
            ref = O.join(data, e, k=k, reorder=1, **kw)
            c2, i2, _ = _run(sj, data, e, k=k, reorder=1, ffilter=2, **kw)
            for ff in (0, 1):
                c, ids, r = _run(sj, data, e, k=k, reorder=1, ffilter=ff, **kw)
                print("d=%d %s n=%d ffilter=%d %s total=%d" % (d, np.dtype(dt).name, len(data), ff, kw, r.total),
                      flush=True)
                assert_csr_equal(c, ids, ref["counts"], ref["ids"])
                assert np.array_equal(c, c2) and np.array_equal(ids, i2)


def test_ffilter_slotted_and_hit_mask_passes(sj):
    """The exact passes with hit masks (the counting pass stores the
    filter's ballots, the fill reads them) and the learned slots sized from
    them, on boundary-heavy 6-D data."""
    from gpu_self_join_amd import SelfJoin, _lib
    x = _shell_pairs(20000, 6, np.float32, 0.6, 7)
    ref = O.join(x, 0.6, k=6, ilp=8, shortcircuit=1, reorder=1)
    c2, i2, _ = _run(sj, x, 0.6, k=6, ilp=8, shortcircuit=1, reorder=1, ffilter=2)
    s2 = SelfJoin(0)
    try:
        for _ in range(2):
            c, ids, r = _run(s2, x, 0.6, k=6, ilp=8, shortcircuit=1, reorder=1, estimator=_lib.EST_LEARNED)
        assert r.stats["est_tiles"] > 0
        assert_csr_equal(c, ids, ref["counts"], ref["ids"])
        assert np.array_equal(ids, i2)
    finally:
        s2.close()


def test_ffilter_param_validation(sj):
    import torch
    x = torch.zeros((10, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(Exception):
        sj.join(x, 0.5, ffilter=3)
