"""Golden CSR fixtures (tests/golden/grid_*.npz, written by
tests/golden/make_fixtures.py with the oracle): the oracle reproduces them
(regression of the CPU restatement) and the GPU path matches them in every
mode (plain, Gram filter for dims >= 16, learned slots)."""
import glob
import os

import numpy as np
import pytest

from gpu_self_join_amd import _lib
from oracle import oracle as O
from tests.helpers import sorted_rows

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = sorted(glob.glob(os.path.join(HERE, "golden", "grid_*.npz")))


def _load(path):
    z = np.load(path)  # plain arrays only (allow_pickle stays False)
    k, ilp, sc, reorder = (int(v) for v in z["params"])
    return z["data"], float(z["eps"]), dict(k=k, ilp=ilp, shortcircuit=sc, reorder=reorder), z["counts"], z["ids"]


def test_fixtures_present():
    assert len(FIX) >= 6


@pytest.mark.parametrize("path", FIX, ids=lambda p: os.path.basename(p)[5:-4])
def test_oracle_reproduces_fixture(path):
    x, eps, p, counts, ids = _load(path)
    r = O.join(x, eps, **p)
    assert np.array_equal(r["counts"], counts)
    assert np.array_equal(sorted_rows(r["counts"], r["ids"]), ids)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIX, ids=lambda p: os.path.basename(p)[5:-4])
def test_gpu_matches_fixture(sj, path):
    import torch
    x, eps, p, counts, ids = _load(path)
    pts = torch.from_numpy(x).cuda()
    modes = [dict(filter=_lib.FILTER_PLAIN), dict(filter=_lib.FILTER_PLAIN, stamp=1)]
    if x.shape[1] >= 16:
        modes.append(dict(filter=_lib.FILTER_GRAM))
    for m in modes:
        r = sj.join(pts, eps, **p, **m)
        c = r.counts.cpu().numpy().view(np.uint32)
        i = r.ids.cpu().numpy().view(np.uint32)
        assert np.array_equal(c, counts), m
        assert np.array_equal(sorted_rows(c, i), ids), m
    if p["ilp"] > 0 and p["shortcircuit"]:
        # the other reading of the ILP fold (every dim fused) against the
        # oracle's table for that reading
        ref = O.join(x, eps, fold_fused=1, **p)
        print(os.path.basename(path), "ambiguous pairs", int(ref["stats"][3]), flush=True)
        for m in modes:
            r = sj.join(pts, eps, ilp_fold=1, **p, **m)
            c = r.counts.cpu().numpy().view(np.uint32)
            i = r.ids.cpu().numpy().view(np.uint32)
            assert np.array_equal(c, ref["counts"]), m
            assert np.array_equal(sorted_rows(c, i), sorted_rows(ref["counts"], ref["ids"])), m
