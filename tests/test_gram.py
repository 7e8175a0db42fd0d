"""The MFMA Gram-form filter (dims >= 16): identical tables to the plain join
and to the oracle, including pairs at exactly eps, and filter statistics.
Both operand formats: fp16 (the default) and, for FP32 data, bf16 hi + lo."""
import numpy as np
import pytest

from gpu_self_join_amd import _lib
from oracle import oracle as O
from tests.helpers import assert_csr_equal

pytestmark = pytest.mark.gpu


def _join(sj, x, eps, **kw):
    import torch
    r = sj.join(torch.from_numpy(np.ascontiguousarray(x)).cuda(), eps, **kw)
    return r.counts.cpu().numpy().view(np.uint32), r.ids.cpu().numpy().view(np.uint32), r


def _clustered(n, d, dt, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    c = rng.random((6, d)) * 30
    return (c[rng.integers(0, 6, n)] + spread * rng.standard_normal((n, d))).astype(dt)


def _eps(x, nb):
    rng = np.random.default_rng(1)
    q = x[rng.choice(len(x), 64, replace=False)].astype(np.float64)
    dd = np.sqrt(((q[:, None, :] - x[None, :, :].astype(np.float64)) ** 2).sum(-1))
    return float(np.asarray(np.sort(dd, axis=1)[:, nb].mean(), dtype=x.dtype))


CASES = [(3000, 16, 6, np.float32, 8, 1, 1), (3000, 32, 6, np.float32, 8, 1, 1), (2000, 64, 8, np.float32, 0, 0, 1),
         (2500, 20, 4, np.float32, 8, 0, 0), (2000, 100, 6, np.float32, 8, 1, 1),
         (2000, 16, 6, np.float64, 8, 1, 1), (2000, 64, 8, np.float64, 8, 1, 1), (1500, 40, 5, np.float64, 0, 0, 0)]


MODES = [_lib.FILTER_GRAM, _lib.FILTER_GRAM_BF16]


@pytest.mark.parametrize("mode", MODES, ids=["fp16", "bf16"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_d%d_k%d_%s_ilp%d_sc%d_r%d" % (
    c[0], c[1], c[2], np.dtype(c[3]).name, c[4], c[5], c[6]))
def test_gram_equals_plain_and_oracle(sj, case, mode):
    n, d, k, dt, ilp, sc, reorder = case
    if mode == _lib.FILTER_GRAM_BF16 and dt == np.float64:
        pytest.skip("bf16 operands apply to FP32 data only")
    x = _clustered(n, d, dt, n + d)
    eps = _eps(x, 25)
    kw = dict(k=k, ilp=ilp, shortcircuit=sc, reorder=reorder)
    cg, ig, rg = _join(sj, x, eps, filter=mode, **kw)
    cp, ip, rp = _join(sj, x, eps, filter=_lib.FILTER_PLAIN, **kw)
    print("case", case, "total", rg.total, "rechecks", rg.stats["filter_rechecks"], "hits",
          rg.stats["filter_hits"], flush=True)
    assert np.array_equal(cg, cp) and np.array_equal(ig, ip)  # same rows, same order
    assert rg.stats["filter_hits"] == rg.total
    assert rg.stats["filter_rechecks"] >= rg.total
    ref = O.join(x, eps, k=k, ilp=ilp, shortcircuit=sc, reorder=reorder)
    assert_csr_equal(cg, ig, ref["counts"], ref["ids"])


@pytest.mark.parametrize("mode", MODES, ids=["fp16", "bf16"])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_gram_pairs_at_exactly_eps(sj, dt, mode):
    # points on a 16-D lattice of spacing eps: neighbors at distance exactly
    # eps, sqrt(2) eps, ... -- the filter must send every boundary pair to the
    # exact recheck
    side = 3
    g = np.stack(np.meshgrid(*[np.arange(side)] * 4, indexing="ij"), -1).reshape(-1, 4)
    rng = np.random.default_rng(3)
    basis = np.zeros((4, 16))
    basis[:, :4] = np.eye(4)
    eps = 0.75
    x = (g @ basis * eps + rng.random(16) * 50).astype(dt)
    x = np.concatenate([x, x + dt(0.01 * eps)], axis=0)
    e = float(np.asarray(eps, dtype=dt))
    for ilp in (0, 8):
        cg, ig, rg = _join(sj, x, e, k=4, ilp=ilp, shortcircuit=1, reorder=0, filter=mode)
        ref = O.join(x, e, k=4, ilp=ilp, shortcircuit=1, reorder=0)
        assert_csr_equal(cg, ig, ref["counts"], ref["ids"])


@pytest.mark.parametrize("mode", MODES, ids=["fp16", "bf16"])
def test_gram_large_coordinates(sj, mode):
    # far from the origin: the Gram form cancels catastrophically unless the
    # bound accounts for it; the result must still be exact
    x = _clustered(2000, 24, np.float32, 5) + np.float32(3.0e4)
    eps = _eps(x, 20)
    cg, ig, rg = _join(sj, x, eps, k=6, ilp=8, shortcircuit=1, reorder=1, filter=mode)
    ref = O.join(x, eps, k=6, ilp=8, shortcircuit=1, reorder=1)
    assert_csr_equal(cg, ig, ref["counts"], ref["ids"])


def test_gram_with_learned_slots(sj):
    from gpu_self_join_amd import SelfJoin
    x = _clustered(20000, 32, np.float32, 9, spread=2.0)
    eps = _eps(x, 30)
    kw = dict(k=6, ilp=8, shortcircuit=1, reorder=1, filter=_lib.FILTER_GRAM, estimator=_lib.EST_LEARNED)
    h = SelfJoin(0)
    try:
        _join(h, x, eps, **kw)
        c, ids, r = _join(h, x, eps, **kw)
        assert r.stats["est_tiles"] > 0
    finally:
        h.close()
    ref = O.join(x, eps, k=6, ilp=8, shortcircuit=1, reorder=1)
    assert_csr_equal(c, ids, ref["counts"], ref["ids"])


def test_gram_formats_agree_and_unknown_filter_rejected(sj):
    x = _clustered(4000, 32, np.float32, 21, spread=1.5)
    eps = _eps(x, 30)
    kw = dict(k=6, ilp=8, shortcircuit=1, reorder=1)
    c1, i1, r1 = _join(sj, x, eps, filter=_lib.FILTER_GRAM, **kw)
    c2, i2, r2 = _join(sj, x, eps, filter=_lib.FILTER_GRAM_BF16, **kw)
    assert np.array_equal(c1, c2) and np.array_equal(i1, i2)
    with pytest.raises(_lib.GdsjError):
        _join(sj, x, eps, filter=7, **kw)


@pytest.mark.parametrize("mode", MODES, ids=["fp16", "bf16"])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_gram_dense_cluster_every_pair_survives(sj, dt, mode):
    # 700 points within about 0.2 of each other (eps = 1) among a sparse
    # background: the cluster's tiles are 128 queries whose every candidate
    # is a neighbor, so each 32 x 32 row block queues 1024 survivors (the
    # recheck queue's largest load, flushed within the block) and the fill
    # emits many hits of one row per recheck round
    from gpu_self_join_amd import SelfJoin
    if mode == _lib.FILTER_GRAM_BF16 and dt == np.float64:
        pytest.skip("bf16 operands apply to FP32 data only")
    rng = np.random.default_rng(11)
    d = 32
    core = 50.0 + 0.02 * rng.standard_normal((700, d))
    bg = rng.random((1500, d)) * 100.0
    x = np.concatenate([core, bg])[rng.permutation(2200)].astype(dt)
    eps = 1.0
    kw = dict(k=6, ilp=8, shortcircuit=1, reorder=1)
    ref = O.join(x, eps, k=6, ilp=8, shortcircuit=1, reorder=1)
    assert int(np.asarray(ref["counts"]).sum()) >= 700 * 700
    h = SelfJoin(0)
    try:
        cp, ip, _ = _join(h, x, eps, filter=_lib.FILTER_PLAIN, **kw)
        for est in (_lib.EST_EXACT, _lib.EST_LEARNED, _lib.EST_LEARNED):
            cg, ig, rg = _join(h, x, eps, filter=mode, estimator=est, **kw)
            print("dense cluster", np.dtype(dt).name, "estimator", est, "total", rg.total, "rechecks",
                  rg.stats["filter_rechecks"], flush=True)
            assert np.array_equal(cg, cp) and np.array_equal(ig, ip)  # same rows, same order
            assert_csr_equal(cg, ig, ref["counts"], ref["ids"])
    finally:
        h.close()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_gram_forced_overflow_fill_split_by_candidates(sj, dt):
    # a model that predicts no hits: nearly every tile overflows its slot and
    # is filled exactly on sub-tiles whose candidates are split 32 ways
    # (equal ranges of the cell's candidate order, cut inside windows and
    # blocks), counted per split and placed by a prefix over the splits; the
    # rows must come out as the plain join's, in the same order
    import ctypes as C
    from gpu_self_join_amd import SelfJoin
    from oracle import mlp as M
    L = _lib.load()
    x = _clustered(6000, 32, dt, 77, spread=2.0)
    eps = _eps(x, 40)
    kw = dict(k=6, ilp=8, shortcircuit=1, reorder=1)
    h = SelfJoin(0)
    try:
        cp, ip, _ = _join(h, x, eps, filter=_lib.FILTER_PLAIN, **kw)
        p = np.zeros(M.P, np.float32)
        assert L.gdsj_model_set(h.h, p.ctypes.data, M.P, C.c_double(0.0)) == 0
        cg, ig, rg = _join(h, x, eps, filter=_lib.FILTER_GRAM, estimator=_lib.EST_LEARNED, train=2, **kw)
        print("forced overflow", np.dtype(dt).name, "tiles", rg.stats["est_tiles"], "overflowed",
              rg.stats["est_overflow_tiles"], "pairs", rg.total, flush=True)
        assert rg.stats["est_overflow_tiles"] > 0
        assert np.array_equal(cg, cp) and np.array_equal(ig, ip)  # same rows, same order
    finally:
        h.close()
    ref = O.join(x, eps, k=6, ilp=8, shortcircuit=1, reorder=1)
    assert_csr_equal(cg, ig, ref["counts"], ref["ids"])
