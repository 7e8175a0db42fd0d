"""Result-size estimators of the single-pass join.

CPU: the float64 MLP restatement's analytic gradient against central finite
differences (pins oracle/mlp.py).  GPU: the HIP forward/backward kernels
against that restatement; slotted joins (learned and sampling estimators,
forced slot overflow) against the oracle table; what the trained network
gains over the sampling estimator on clustered data.
"""
import ctypes as C

import numpy as np
import pytest

from gpu_self_join_amd import _lib
from oracle import mlp as M
from oracle import oracle as O
from tests.helpers import assert_csr_equal


def test_mlp_restatement_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init(1)
    x = rng.uniform(0, 1, (40, M.F))
    t = rng.uniform(2, 10, 40)
    _, g = M.loss_and_grad(p, x, t)
    idx = rng.choice(M.P, 200, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, x, t)[0] - M.loss_and_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), i


def test_mlp_layout_constants_match_library():
    L = _lib.load()
    assert L.gdsj_mlp_param_count() == M.P
    assert L.gdsj_mlp_feature_count() == M.F


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


@pytest.mark.gpu
def test_mlp_forward_kernel_matches_restatement(gpu):
    L = _lib.load()
    rng = np.random.default_rng(2)
    p = _f32(M.init(3))
    x = _f32(rng.uniform(0, 1.5, (1000, M.F)))
    y = np.zeros(1000, np.float32)
    assert L.gdsj_mlp_forward(p.ctypes.data, x.ctypes.data, 1000, y.ctypes.data) == 0
    ref, _ = M.forward(p.astype(np.float64), x.astype(np.float64))
    np.testing.assert_allclose(y, ref, rtol=2e-5, atol=2e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 63, 64, 65, 300, 2048])
def test_mlp_gradient_kernel_matches_restatement(gpu, n):
    L = _lib.load()
    rng = np.random.default_rng(n)
    p = _f32(M.init(4))
    x = _f32(rng.uniform(0, 1.5, (n, M.F)))
    t = _f32(rng.uniform(1, 12, n))
    g = np.zeros(M.P, np.float32)
    loss = C.c_float(0)
    assert L.gdsj_mlp_grad(p.ctypes.data, x.ctypes.data, t.ctypes.data, n, g.ctypes.data, C.byref(loss)) == 0
    lref, gref = M.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(loss.value - lref) <= 1e-4 * max(1.0, lref)
    scale = np.max(np.abs(gref))
    np.testing.assert_allclose(g, gref, rtol=1e-3, atol=1e-5 * scale)


def _gmm(n, d, seed, ncl=12, sigma=1.0):
    rng = np.random.default_rng(seed)
    centers = rng.random((ncl, d)) * 40
    w = rng.dirichlet(np.ones(ncl) * 0.7)
    lab = rng.choice(ncl, n, p=w)
    s = sigma * rng.uniform(0.3, 2.0, ncl)
    return (centers[lab] + rng.standard_normal((n, d)) * s[lab, None]).astype(np.float32)


def _join(sj, x, eps, **kw):
    import torch
    r = sj.join(torch.from_numpy(x).cuda(), eps, **kw)
    return r.counts.cpu().numpy().view(np.uint32), r.ids.cpu().numpy().view(np.uint32), r


@pytest.mark.gpu
def test_learned_estimator_slotted_join_matches_oracle(gpu):
    from gpu_self_join_amd import SelfJoin
    sj = SelfJoin(0)
    try:
        x = _gmm(60000, 4, 7)
        eps = 0.35
        kw = dict(k=4, ilp=8, shortcircuit=1, reorder=1, estimator=_lib.EST_LEARNED)
        ref = O.join(x, eps, k=4, ilp=8, shortcircuit=1, reorder=1)
        c1, i1, r1 = _join(sj, x, eps, **kw)  # untrained: exact passes, then training
        assert r1.stats["est_tiles"] == 0 and r1.stats["ms_train"] > 0
        assert_csr_equal(c1, i1, ref["counts"], ref["ids"])
        c2, i2, r2 = _join(sj, x, eps, **kw)  # trained: slotted single pass
        st = r2.stats
        print("learned:", {k: st[k] for k in ("est_predicted", "est_capacity", "est_overflow_tiles", "est_tiles",
                                               "est_mean_rel_err", "est_max_rel_err")}, flush=True)
        assert st["est_tiles"] > 0 and st["est_capacity"] >= r2.total * 0.5
        assert_csr_equal(c2, i2, ref["counts"], ref["ids"])
        assert np.array_equal(i1, i2)  # same row order as the exact fill
    finally:
        sj.close()


@pytest.mark.gpu
def test_forced_slot_overflow_falls_back_exactly(gpu):
    from gpu_self_join_amd import SelfJoin
    L = _lib.load()
    sj = SelfJoin(0)
    try:
        x = _gmm(30000, 3, 8)
        eps = 0.4
        # a model that predicts 0 hits everywhere: every slot holds only its
        # floor, so nearly all tiles overflow and are completed exactly
        p = np.zeros(M.P, np.float32)
        assert L.gdsj_model_set(sj.h, p.ctypes.data, M.P, C.c_double(0.0)) == 0
        c, ids, r = _join(sj, x, eps, k=3, ilp=0, reorder=0, estimator=_lib.EST_LEARNED, train=2)
        assert r.stats["est_overflow_tiles"] > 0
        ref = O.join(x, eps, k=3)
        assert_csr_equal(c, ids, ref["counts"], ref["ids"])
    finally:
        sj.close()


@pytest.mark.gpu
def test_sampling_estimator_slotted_join_matches_oracle(gpu):
    from gpu_self_join_amd import SelfJoin
    sj = SelfJoin(0)
    try:
        x = _gmm(40000, 5, 9)
        eps = 0.6
        c, ids, r = _join(sj, x, eps, k=5, ilp=8, shortcircuit=1, reorder=1, estimator=_lib.EST_SAMPLE)
        assert r.stats["est_tiles"] > 0
        ref = O.join(x, eps, k=5, ilp=8, shortcircuit=1, reorder=1)
        assert_csr_equal(c, ids, ref["counts"], ref["ids"])
    finally:
        sj.close()


@pytest.mark.gpu
def test_learned_estimator_beats_sampling_on_clustered_data(gpu):
    """What the trained network gains over the method without it."""
    from gpu_self_join_amd import SelfJoin
    sj = SelfJoin(0)
    try:
        x = _gmm(200000, 6, 10, ncl=20, sigma=1.5)
        eps = 0.9
        kw = dict(k=6, ilp=8, shortcircuit=1, reorder=1)
        _join(sj, x, eps, estimator=_lib.EST_LEARNED, **kw)  # train
        _, _, rl = _join(sj, x, eps, estimator=_lib.EST_LEARNED, **kw)
        _, _, rs = _join(sj, x, eps, estimator=_lib.EST_SAMPLE, **kw)
        sl, ss = rl.stats, rs.stats
        ovl = sl["est_overflow_tiles"] / sl["est_tiles"]
        ovs = ss["est_overflow_tiles"] / ss["est_tiles"]
        print("learned overflow %.4f mean rel err %.3f | sampling overflow %.4f mean rel err %.3f" % (
            ovl, sl["est_mean_rel_err"], ovs, ss["est_mean_rel_err"]), flush=True)
        assert ovl < ovs
        assert sl["est_mean_rel_err"] < ss["est_mean_rel_err"]
    finally:
        sj.close()


@pytest.mark.gpu
def test_learned_estimator_gain_holds_out_of_sample(gpu):
    """The network trained on one dataset, then frozen (train = 2), sizes
    joins it has not seen: another seed of the same family and the training
    data at 0.9 and 1.1 times epsilon.  On each it beats the sampling
    estimator in mean relative error per tile and in overflowed tiles."""
    from gpu_self_join_amd import SelfJoin
    sj = SelfJoin(0)
    try:
        kw = dict(k=6, ilp=8, shortcircuit=1, reorder=1)
        x = _gmm(200000, 6, 10, ncl=20, sigma=1.5)
        eps = 0.9
        _join(sj, x, eps, estimator=_lib.EST_LEARNED, train=0, **kw)  # exact passes + training
        held = [("seed 12", _gmm(200000, 6, 12, ncl=20, sigma=1.5), eps),
                ("eps x 0.9", x, float(np.float32(eps * 0.9))),
                ("eps x 1.1", x, float(np.float32(eps * 1.1)))]
        for name, xh, eh in held:
            _, _, rl = _join(sj, xh, eh, estimator=_lib.EST_LEARNED, train=2, **kw)
            _, _, rs = _join(sj, xh, eh, estimator=_lib.EST_SAMPLE, **kw)
            sl, ss = rl.stats, rs.stats
            assert sl["est_tiles"] > 0 and ss["est_tiles"] > 0
            ovl = sl["est_overflow_tiles"] / sl["est_tiles"]
            ovs = ss["est_overflow_tiles"] / ss["est_tiles"]
            print("%s: learned overflow %.4f mean rel err %.3f capacity/pairs %.2f | sampling overflow %.4f "
                  "mean rel err %.3f capacity/pairs %.2f" % (
                      name, ovl, sl["est_mean_rel_err"], sl["est_capacity"] / rl.total, ovs,
                      ss["est_mean_rel_err"], ss["est_capacity"] / rs.total), flush=True)
            assert sl["est_mean_rel_err"] < ss["est_mean_rel_err"], name
            assert ovl < ovs, name
    finally:
        sj.close()
