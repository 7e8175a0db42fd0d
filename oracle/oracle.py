"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() (as the
checker) and bench.py's cpu_baseline leg.  The product package never imports
this module.  See gdsj_oracle.c for the reference lines each step restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("ilp", C.c_uint32), ("shortcircuit", C.c_uint32),
                ("reorder", C.c_uint32), ("fold_fused", C.c_uint32), ("nthreads", C.c_int32),
                ("sort", C.c_uint32)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        for sfx, ct in (("f32", C.c_float), ("f64", C.c_double)):
            P = C.c_void_p
            getattr(L, "orc_join_" + sfx).argtypes = [P, C.c_uint32, C.c_uint32, ct, C.POINTER(OrcParams),
                                                      P, P, P, C.c_uint64, P]
            getattr(L, "orc_join_subset_" + sfx).argtypes = [P, C.c_uint32, C.c_uint32, ct, C.POINTER(OrcParams),
                                                             P, C.c_uint32, P, P, P]
            getattr(L, "orc_bruteforce_" + sfx).argtypes = [P, C.c_uint32, C.c_uint32, ct, C.POINTER(OrcParams), P, P]
            getattr(L, "orc_grid_info_" + sfx).argtypes = [P, C.c_uint32, C.c_uint32, ct, C.POINTER(OrcParams),
                                                           P, P, P, P, P, P, P]
            getattr(L, "orc_reorder_perm_" + sfx).argtypes = [P, C.c_uint32, C.c_uint32, P, P]
            getattr(L, "orc_parse_text_" + sfx).argtypes = [C.c_char_p, C.c_size_t, P, C.c_size_t]
            getattr(L, "orc_parse_text_" + sfx).restype = C.c_size_t
            getattr(L, "cpu_bruteforce_" + sfx).argtypes = [P, C.c_uint32, C.c_uint32, ct, P, C.c_uint32, C.c_int,
                                                            P, C.POINTER(C.c_double)]
            getattr(L, "cpu_rtree_" + sfx).argtypes = [P, C.c_uint32, C.c_uint32, ct, P, C.c_uint32, P,
                                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
        for sfx, ct in (("f32", C.c_float), ("f64", C.c_double)):
            f = getattr(L, "orc_acc_" + sfx)
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
            f.restype = ct
        L.orc_mix64.argtypes = [C.c_uint64]
        L.orc_mix64.restype = C.c_uint64
        _LIB = L
    return _LIB


def _sfx(dt):
    return "f32" if np.dtype(dt) == np.float32 else "f64"


def _params(k, ilp=0, shortcircuit=0, reorder=0, fold_fused=0, nthreads=0, sort=0):
    return OrcParams(k, ilp, shortcircuit, reorder, fold_fused, nthreads, sort)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _prep(data):
    data = np.ascontiguousarray(data)
    assert data.ndim == 2 and data.dtype in (np.float32, np.float64)
    return data


def _scalar(dt, eps):
    return C.c_float(eps) if np.dtype(dt) == np.float32 else C.c_double(eps)


def join(data, eps, k, ilp=0, shortcircuit=0, reorder=0, fold_fused=0, with_ids=True, nthreads=0, sort=0):
    """Full self-join on the CPU.  Returns dict(counts, hashes, ids|None, stats)."""
    data = _prep(data)
    n, d = data.shape
    p = _params(k, ilp, shortcircuit, reorder, fold_fused, nthreads, sort)
    counts = np.zeros(n, np.uint32)
    hashes = np.zeros(n, np.uint64)
    stats = np.zeros(8, np.uint64)
    f = getattr(lib(), "orc_join_" + _sfx(data.dtype))
    rc = f(_ptr(data), n, d, _scalar(data.dtype, eps), C.byref(p), _ptr(counts), _ptr(hashes),
           None, 0, _ptr(stats))
    if rc != 0:
        raise RuntimeError("oracle join failed: %d" % rc)
    ids = None
    if with_ids:
        total = int(counts.sum(dtype=np.uint64))
        ids = np.zeros(max(total, 1), np.uint32)
        c2 = np.zeros(n, np.uint32)
        rc = f(_ptr(data), n, d, _scalar(data.dtype, eps), C.byref(p), _ptr(c2), None,
               _ptr(ids), total, None)
        if rc != 0:
            raise RuntimeError("oracle join (fill) failed: %d" % rc)
        ids = ids[:total]
    return {"counts": counts, "hashes": hashes, "ids": ids, "stats": stats}


def join_subset(data, eps, k, qids, ilp=0, shortcircuit=0, reorder=0, fold_fused=0, nthreads=0):
    data = _prep(data)
    n, d = data.shape
    qids = np.ascontiguousarray(qids, dtype=np.uint32)
    p = _params(k, ilp, shortcircuit, reorder, fold_fused, nthreads)
    counts = np.zeros(len(qids), np.uint32)
    hashes = np.zeros(len(qids), np.uint64)
    stats = np.zeros(8, np.uint64)
    f = getattr(lib(), "orc_join_subset_" + _sfx(data.dtype))
    rc = f(_ptr(data), n, d, _scalar(data.dtype, eps), C.byref(p), _ptr(qids), len(qids),
           _ptr(counts), _ptr(hashes), _ptr(stats))
    if rc != 0:
        raise RuntimeError("oracle subset join failed: %d" % rc)
    return {"counts": counts, "hashes": hashes, "stats": stats}


def bruteforce(data, eps, ilp=0, shortcircuit=0, reorder=0, fold_fused=0):
    data = _prep(data)
    n, d = data.shape
    p = _params(1, ilp, shortcircuit, reorder, fold_fused, 0)
    counts = np.zeros(n, np.uint32)
    hashes = np.zeros(n, np.uint64)
    f = getattr(lib(), "orc_bruteforce_" + _sfx(data.dtype))
    rc = f(_ptr(data), n, d, _scalar(data.dtype, eps), C.byref(p), _ptr(counts), _ptr(hashes))
    if rc != 0:
        raise RuntimeError("oracle brute force failed: %d" % rc)
    return {"counts": counts, "hashes": hashes}


def grid_info(data, eps, k, reorder=0):
    data = _prep(data)
    n, d = data.shape
    p = _params(k, 0, 0, reorder)
    minarr = np.zeros(k, data.dtype)
    ncells = np.zeros(k, np.uint32)
    total = np.zeros(1, np.uint64)
    ncell = np.zeros(1, np.uint32)
    perm = np.zeros(d, np.uint32)
    cellids = np.zeros(max(n, 1), np.uint64)
    order = np.zeros(max(n, 1), np.uint32)
    f = getattr(lib(), "orc_grid_info_" + _sfx(data.dtype))
    rc = f(_ptr(data), n, d, _scalar(data.dtype, eps), C.byref(p), _ptr(minarr), _ptr(ncells),
           _ptr(total), _ptr(ncell), _ptr(perm), _ptr(cellids), _ptr(order))
    if rc != 0:
        raise RuntimeError("oracle grid failed: %d" % rc)
    nc = int(ncell[0])
    return {"min": minarr, "ncells": ncells, "total_cells": int(total[0]), "n_nonempty": nc,
            "perm": perm, "cell_ids": cellids[:nc], "order": order[:n]}


def acc(p, q, ilp=0, shortcircuit=0, fold_fused=0):
    """The oracle's squared-distance accumulation of candidate q against query p."""
    p = np.ascontiguousarray(p)
    q = np.ascontiguousarray(q, dtype=p.dtype)
    f = getattr(lib(), "orc_acc_" + _sfx(p.dtype))
    return f(_ptr(p), _ptr(q), len(p), ilp, shortcircuit, fold_fused)


def reorder_perm(data):
    data = _prep(data)
    n, d = data.shape
    perm = np.zeros(d, np.uint32)
    var = np.zeros(d, data.dtype)
    getattr(lib(), "orc_reorder_perm_" + _sfx(data.dtype))(_ptr(data), n, d, _ptr(perm), _ptr(var))
    return perm, var


def parse_text(text, dtype=np.float32, cap=None):
    """Decimal text -> DTYPE values, each token parsed directly into DTYPE."""
    if isinstance(text, str):
        text = text.encode()
    if cap is None:
        cap = text.count(b",") + text.count(b"\n") + 2
    out = np.zeros(cap, dtype)
    m = getattr(lib(), "orc_parse_text_" + _sfx(dtype))(text, len(text), _ptr(out), cap)
    return out[:m]


_MIX = None


def mix64(ids):
    """Vectorised orc_mix64 (splitmix64 finalizer) over uint32/uint64 ids."""
    z = np.asarray(ids, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def csr_hashes(counts, ids):
    """Per-point order-independent hash of a CSR table (sum of mix64(id))."""
    counts = np.asarray(counts, dtype=np.int64)
    h = np.zeros(len(counts), np.uint64)
    if len(ids) == 0:
        return h
    m = mix64(ids)
    offs = np.zeros(len(counts) + 1, np.int64)
    np.cumsum(counts, out=offs[1:])
    nz = counts > 0
    with np.errstate(over="ignore"):
        cs = np.concatenate([[np.uint64(0)], np.cumsum(m, dtype=np.uint64)])
        h = cs[offs[1:]] - cs[offs[:-1]]
    h[~nz] = 0
    return h


# ------------------------------------------------------------ CPU baselines --
# The reference's CPU self-join paths restated (cpu_baselines.c), for timing
# beside the GPU: unfused filter arithmetic, so boundary pairs may differ
# from the grid contract.

def cpu_bruteforce(data, eps, qids, nthreads=1):
    """Counts of the queries qids against all points; returns (counts, seconds)."""
    data = _prep(data)
    n, d = data.shape
    q = np.ascontiguousarray(qids, dtype=np.uint32)
    counts = np.zeros(len(q), np.uint32)
    sec = C.c_double(0)
    rc = getattr(lib(), "cpu_bruteforce_" + _sfx(data.dtype))(_ptr(data), n, d, _scalar(data.dtype, eps), _ptr(q),
                                                              len(q), int(nthreads), _ptr(counts), C.byref(sec))
    if rc != 0:
        raise RuntimeError("cpu_bruteforce failed: %d" % rc)
    return counts, sec.value


def cpu_rtree(data, eps, qids):
    """Sequential R-tree: builds over all points, queries qids.
    Returns (counts, build_seconds, query_seconds)."""
    data = _prep(data)
    n, d = data.shape
    q = np.ascontiguousarray(qids, dtype=np.uint32)
    counts = np.zeros(len(q), np.uint32)
    tb, tq = C.c_double(0), C.c_double(0)
    rc = getattr(lib(), "cpu_rtree_" + _sfx(data.dtype))(_ptr(data), n, d, _scalar(data.dtype, eps), _ptr(q), len(q),
                                                         _ptr(counts), C.byref(tb), C.byref(tq))
    if rc != 0:
        raise RuntimeError("cpu_rtree failed: %d" % rc)
    return counts, tb.value, tq.value
