"""The all-pairs join (gdsj_bruteforce_device; the reference's search mode 1,
GPU.cu:1285-1480, kernel.cu:486-515) against the oracle's all-pairs join,
and its relation to the grid join: every grid row is contained in the
all-pairs row (the grid only leaves out pairs whose cells are not adjacent)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import README_COUNT, load_test_dat

pytestmark = pytest.mark.gpu


def _rows(r):
    counts = r.counts.cpu().numpy().view(np.uint32)
    ids = r.ids.cpu().numpy().view(np.uint32)
    offs = r.offsets.cpu().numpy()
    assert offs[0] == 0 and int(offs[-1]) == r.total
    return counts, ids, offs


CASES = [
    # (n, d, dtype, ilp, sc)
    (2000, 2, np.float32, 0, 0),
    (1500, 3, np.float32, 8, 1),
    (1200, 7, np.float32, 3, 0),
    (1000, 16, np.float32, 8, 1),
    (800, 5, np.float64, 0, 0),
    (700, 33, np.float64, 8, 0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_d%d_%s_ilp%d_sc%d" % (
    c[0], c[1], np.dtype(c[2]).name, c[3], c[4]))
def test_bruteforce_equals_oracle(sj, case):
    import torch
    n, d, dt, ilp, sc = case
    rng = np.random.default_rng(n + d)
    x = (rng.standard_normal((n, d)) * 2.0).astype(dt)
    eps = float(np.asarray(1.5 + 0.1 * d, dtype=dt))
    r = sj.bruteforce(torch.from_numpy(x).cuda(), eps, ilp=ilp, shortcircuit=sc)
    counts, ids, offs = _rows(r)
    ref = O.bruteforce(x, eps, ilp=ilp, shortcircuit=sc, reorder=0)
    print("bruteforce n=%d d=%d total=%d" % (n, d, r.total), flush=True)
    assert np.array_equal(counts, ref["counts"])
    assert np.array_equal(O.csr_hashes(counts, ids), ref["hashes"])
    # each row ascending, the point itself included
    for i in range(0, n, max(1, n // 50)):
        row = ids[offs[i]:offs[i + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0) and i in row


def test_bruteforce_contains_grid_rows_readme(sj):
    """test_dat.txt at eps 1.0: the grid join's rows (27,982,028 pairs,
    README.md:82-84) are contained in the all-pairs rows."""
    import torch
    x = load_test_dat(np.float32)
    t = torch.from_numpy(x).cuda()
    b = sj.bruteforce(t, 1.0, ilp=0, shortcircuit=0)
    g = sj.join(t, 1.0, k=2, ilp=0, shortcircuit=0, reorder=0, queryreorder=0, sort_neighbors=1)
    bc, bi, bo = _rows(b)
    gc, gi, go = _rows(g)
    print("test_dat all-pairs %d grid %d" % (b.total, g.total), flush=True)
    assert g.total == README_COUNT
    assert b.total >= g.total and np.all(bc >= gc)
    # identical rows wherever the counts agree; contained everywhere
    same = np.nonzero(bc == gc)[0]
    for i in same[:: max(1, len(same) // 2000)]:
        assert np.array_equal(bi[bo[i]:bo[i + 1]], gi[go[i]:go[i + 1]])
    for i in np.nonzero(bc != gc)[0][:200]:
        assert np.all(np.isin(gi[go[i]:go[i + 1]], bi[bo[i]:bo[i + 1]]))


def test_bruteforce_edge_cases(sj):
    import torch
    from gpu_self_join_amd import _lib
    x = np.zeros((0, 3), np.float32)
    r = sj.bruteforce(torch.from_numpy(np.zeros((1, 3), np.float32)).cuda(), 0.5)
    assert r.total == 1
    dup = np.ones((130, 4), np.float64)
    r = sj.bruteforce(torch.from_numpy(dup).cuda(), 1e-9)
    assert r.total == 130 * 130
    with pytest.raises(_lib.GdsjError):
        sj.bruteforce(torch.from_numpy(np.ones((4, 2), np.float32)).cuda(), 0.0)
    del x
