"""Writes the small golden CSR fixtures (tests/golden/grid_*.npz).

Each fixture holds a seeded input, its epsilon and parameters, and the
oracle's table (counts and each point's neighbor ids sorted ascending).
The oracle computed them in this repository's container; the tests compare
both the oracle (regression) and the GPU path against them.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

# name, n, dims, k, dtype, ilp, shortcircuit, reorder, eps, generator
FIXTURES = [
    ("uniform2d_ilp0", 1500, 2, 2, np.float32, 0, 0, 0, 0.9, "uniform"),
    ("gmm3d_ilp8_sc", 1200, 3, 3, np.float32, 8, 1, 1, 0.7, "gmm"),
    ("gmm6d_ilp8_sc_k6", 1000, 6, 6, np.float32, 8, 1, 1, 1.6, "gmm"),
    ("gmm5d_ilp4_nosc_f64", 900, 5, 4, np.float64, 4, 0, 1, 1.6, "gmm"),
    ("aniso20d_ilp8_k6", 800, 20, 6, np.float32, 8, 1, 1, 5.5, "aniso"),
    ("lowid24d_f64_k8", 700, 24, 8, np.float64, 8, 1, 1, 1.5, "lowid"),
]


def make_input(n, d, dt, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        x = rng.random((n, d)) * 20.0
    elif kind == "gmm":
        c = rng.random((5, d)) * 10.0
        x = c[rng.integers(0, 5, n)] + rng.standard_normal((n, d))
    elif kind == "aniso":
        std = 2.0 * 0.9 ** np.arange(d)
        c = rng.random((4, d)) * 10.0
        x = c[rng.integers(0, 4, n)] + rng.standard_normal((n, d)) * std
    else:  # low intrinsic dimension: a 3-D subspace plus small noise
        q, _ = np.linalg.qr(rng.standard_normal((d, 3)))
        x = rng.standard_normal((n, 3)) @ q.T * 3.0 + 0.05 * rng.standard_normal((n, d))
    return np.ascontiguousarray(x.astype(dt))


def sorted_ids(counts, ids):
    off = np.zeros(len(counts) + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    out = ids.copy()
    for i in range(len(counts)):
        out[off[i]:off[i + 1]] = np.sort(ids[off[i]:off[i + 1]])
    return out


def main():
    for i, (name, n, d, k, dt, ilp, sc, reorder, eps, kind) in enumerate(FIXTURES):
        x = make_input(n, d, dt, kind, 900 + i)
        e = float(np.asarray(eps, dtype=dt))
        r = O.join(x, e, k=k, ilp=ilp, shortcircuit=sc, reorder=reorder)
        np.savez_compressed(os.path.join(HERE, "grid_%s.npz" % name), data=x, eps=np.asarray(e, dtype=dt),
                            params=np.array([k, ilp, sc, reorder], np.uint32), counts=r["counts"],
                            ids=sorted_ids(r["counts"], r["ids"]))
        print(name, "pairs", int(r["counts"].sum()))


if __name__ == "__main__":
    main()
