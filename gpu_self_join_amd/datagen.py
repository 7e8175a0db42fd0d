"""Seeded synthetic inputs for the self-join workloads.

All generators are deterministic for a given seed (numpy PCG64 streams and
elementwise float64 arithmetic, cast once to the DTYPE), so the oracle and
the device path always see the same array.

Workloads (N, dims, distribution, epsilon target):
  uniform2d : 1,000,000 x 2,  U[0,100]^2 FP32, eps = 0.5 (STAMP=1, ILP=0, no high-D options)
  gmm6d     : 2,000,000 x 6,  32 isotropic clusters (sigma 2) + 5% uniform, FP32, ~64 neighbors
  aniso32d  : 1,000,000 x 32, 16 clusters, per-dim std 4*0.9^d, FP32, ~32 neighbors, k=6
  lowid64d  :   500,000 x 64, 8-D linear subspace + N(0,0.05) noise, FP64, ~50 neighbors, k=8
  batched   : 256 instances, N log-uniform [1k,200k], dims in {2,4,8,16}, FP32 GMMs, 10-100 neighbors
Epsilon values were calibrated once on a 1% seeded query sample (see
calibrate_eps) and are stored as exact bit patterns in eps_table.json.
"""
import json
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def uniform(n, d, lo, hi, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    return (lo + (hi - lo) * rng.random((n, d))).astype(dtype)


def gmm(n, d, nclusters, sigma, background, seed, dtype=np.float32, box=100.0):
    """Isotropic Gaussian mixture (equal weights) plus a uniform background fraction."""
    rng = np.random.default_rng(seed)
    centers = box * rng.random((nclusters, d))
    nb = int(round(n * background))
    nc = n - nb
    lab = rng.integers(0, nclusters, nc)
    pts = centers[lab] + sigma * rng.standard_normal((nc, d))
    bg = box * rng.random((nb, d))
    x = np.concatenate([pts, bg], axis=0)
    perm = rng.permutation(n)
    return x[perm].astype(dtype)


def aniso(n, d, nclusters, std0, decay, seed, dtype=np.float32, box=100.0):
    """Clusters with per-dimension std std0*decay^j (j = 0..d-1)."""
    rng = np.random.default_rng(seed)
    centers = box * rng.random((nclusters, d))
    std = std0 * decay ** np.arange(d, dtype=np.float64)
    lab = rng.integers(0, nclusters, n)
    x = centers[lab] + rng.standard_normal((n, d)) * std[None, :]
    return x.astype(dtype)


def _mgs_qr(a):
    """Modified Gram-Schmidt in plain Python floats (deterministic on any host)."""
    rows, cols = len(a), len(a[0])
    q = [[a[i][j] for j in range(cols)] for i in range(rows)]
    for j in range(cols):
        for p in range(j):
            dot = math.fsum(q[i][p] * q[i][j] for i in range(rows))
            for i in range(rows):
                q[i][j] -= dot * q[i][p]
        nrm = math.sqrt(math.fsum(q[i][j] * q[i][j] for i in range(rows)))
        for i in range(rows):
            q[i][j] /= nrm
    return np.array(q, dtype=np.float64)


def lowid(n, d, latent, coeff_std, noise_std, seed, dtype=np.float64):
    """Points on a random latent-dimensional linear subspace plus isotropic noise."""
    rng = np.random.default_rng(seed)
    basis = _mgs_qr(rng.standard_normal((d, latent)).tolist())  # d x latent, orthonormal columns
    coeff = coeff_std * rng.standard_normal((n, latent))
    x = np.zeros((n, d), dtype=np.float64)
    for j in range(latent):  # elementwise accumulation, no BLAS
        x += coeff[:, j:j + 1] * basis[None, :, j]
    x += noise_std * rng.standard_normal((n, d))
    return x.astype(dtype)


# ---------------------------------------------------------------- workloads --
def _eps_table():
    path = os.path.join(_HERE, "eps_table.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def eps_from_hex(h, dtype):
    if np.dtype(dtype) == np.float32:
        return float(np.frombuffer(bytes.fromhex(h), dtype=">f4")[0])
    return float(np.frombuffer(bytes.fromhex(h), dtype=">f8")[0])


def eps_to_hex(e, dtype):
    if np.dtype(dtype) == np.float32:
        return np.array([e], dtype=">f4").tobytes().hex()
    return np.array([e], dtype=">f8").tobytes().hex()


WORKLOADS = {
    "uniform2d": dict(n=1_000_000, d=2, dtype=np.float32, k=2, ilp=0, shortcircuit=0, reorder=0, stamp=1,
                      queryreorder=0, target=None),
    "gmm6d": dict(n=2_000_000, d=6, dtype=np.float32, k=6, ilp=8, shortcircuit=1, reorder=1, stamp=0,
                  queryreorder=1, target=64),
    "aniso32d": dict(n=1_000_000, d=32, dtype=np.float32, k=6, ilp=8, shortcircuit=1, reorder=1, stamp=0,
                     queryreorder=1, target=32),
    "lowid64d": dict(n=500_000, d=64, dtype=np.float64, k=8, ilp=8, shortcircuit=1, reorder=1, stamp=0,
                     queryreorder=1, target=50),
}


WORKLOAD_SEEDS = {"uniform2d": 1, "gmm6d": 2, "aniso32d": 3, "lowid64d": 4}


def generate(name, n=None, seed=None):
    """Data of a workload (optionally only its first n points, same stream);
    seed replaces the workload's own (another draw of the same family)."""
    w = WORKLOADS[name]
    nn = w["n"]
    sd = WORKLOAD_SEEDS[name] if seed is None else seed
    if name == "uniform2d":
        x = uniform(nn, 2, 0.0, 100.0, seed=sd)
    elif name == "gmm6d":
        x = gmm(nn, 6, 32, 2.0, 0.05, seed=sd)
    elif name == "aniso32d":
        x = aniso(nn, 32, 16, 4.0, 0.9, seed=sd)
    elif name == "lowid64d":
        x = lowid(nn, 64, 8, 10.0, 0.05, seed=sd)
    else:
        raise KeyError(name)
    return x if n is None else x[:n].copy()


def workload_eps(name):
    w = WORKLOADS[name]
    if name == "uniform2d":
        return float(np.float32(0.5))
    tab = _eps_table()
    if name not in tab:
        raise KeyError("epsilon of %s not calibrated (eps_table.json)" % name)
    return eps_from_hex(tab[name], w["dtype"])


def workload_params(name):
    w = WORKLOADS[name]
    return {k: w[k] for k in ("k", "ilp", "shortcircuit", "reorder", "stamp", "queryreorder")}


# ------------------------------------------------------------ batched set --
BATCHED_COUNT = 256


def batched_instance(i):
    """Instance i of the batched workload: (data, meta).  seed 5000+i."""
    rng = np.random.default_rng(5000 + i)
    n = int(round(math.exp(rng.uniform(math.log(1_000), math.log(200_000)))))
    d = int(rng.choice([2, 4, 8, 16]))
    ncl = int(rng.integers(4, 17))
    sigma = float(rng.uniform(0.5, 3.0))
    target = int(rng.integers(10, 101))
    x = gmm(n, d, ncl, sigma, 0.0, seed=100_000 + i)
    meta = dict(n=n, d=d, nclusters=ncl, sigma=sigma, target=target, k=min(d, 6), ilp=8, shortcircuit=1,
                reorder=1, stamp=0, queryreorder=1)
    return x, meta


def batched_eps(i):
    tab = _eps_table()
    key = "batched/%d" % i
    if key not in tab:
        raise KeyError("epsilon of %s not calibrated" % key)
    return eps_from_hex(tab[key], np.float32)


# ------------------------------------------------------------ calibration --
def _sample_counts_fn(x, qidx):
    """Returns f(eps) -> mean neighbor count of the sampled queries (float64 distances)."""
    from scipy.spatial import cKDTree
    d = x.shape[1]
    x64 = x.astype(np.float64)
    q = x64[qidx]
    if d <= 16:
        tree = cKDTree(x64)
        return lambda e: float(np.mean(tree.query_ball_point(q, e, return_length=True)))
    # high dims: keep each query's smallest 4096 squared distances (Gram form, chunks)
    keep = min(4096, x.shape[0])
    sq = np.einsum("ij,ij->i", x64, x64)
    best = np.empty((len(qidx), keep))
    for s in range(0, len(qidx), 32):
        qq = q[s:s + 32]
        d2 = sq[None, :] + np.einsum("ij,ij->i", qq, qq)[:, None] - 2.0 * (qq @ x64.T)
        part = np.partition(d2, keep - 1, axis=1)[:, :keep]
        best[s:s + 32] = np.sort(part, axis=1)

    def f(e):
        return float(np.mean([np.searchsorted(r, e * e, side="right") for r in best]))
    return f


def calibrate_eps(x, target, seed, frac=0.01, min_q=200):
    """Bisects eps so the mean neighbor count of a seeded query sample is ~target."""
    rng = np.random.default_rng(seed)
    nq = max(min_q, int(round(frac * x.shape[0])))
    nq = min(nq, x.shape[0])
    qidx = np.sort(rng.choice(x.shape[0], nq, replace=False))
    f = _sample_counts_fn(x, qidx)
    lo, hi = 0.0, 1.0
    while f(hi) < target:
        hi *= 2.0
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if f(mid) < target:
            lo = mid
        else:
            hi = mid
    return hi
