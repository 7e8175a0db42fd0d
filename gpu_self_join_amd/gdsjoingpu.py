"""Python drop-in for the reference's gdsjoingpu.py (gdsjoingpu.py:146-232).

    numNeighbors, neighborTable = gdsjoin(DATASET, EPSILON, NDIM, DTYPE, verbose=False)

DATASET is a flat sequence of N*NDIM values (a 2-D array is ravelled);
DTYPE is "float" or "double".  Returns uint32 numNeighbors[N] and the
concatenated uint32 neighbor ids of every point in point order, as the
reference does (gdsjoingpu.py:197,222,232).  The calls go through the same
two-call C ABI: GDSJoinPy (6-argument layout, main.cu:360) and
copyResultIntoPythonArray (main.cu:507).

Differences from the reference, on purpose:
  * np.asarray replaces np.asfarray (removed in NumPy 2).
  * The gpuContext argument the C signature declares is passed (gpu=...).
  * The library is loaded from this package's directory, not os.getcwd().
  * Errors raise GdsjError instead of printing and returning zeros.
  * NDIM, DTYPE and the params.h options are runtime values (params=...).
"""
import contextlib
import ctypes as C
import os
import sys

import numpy as np

from . import _lib


@contextlib.contextmanager
def _hide_c_stdout():
    """Silences C-level stdout like the reference's HideOutput (gdsjoingpu.py:95-125)."""
    try:
        fd = sys.stdout.fileno()
    except (AttributeError, OSError, ValueError):
        yield
        return
    sys.stdout.flush()
    saved = os.dup(fd)
    devnull = os.open(os.devnull, os.O_WRONLY)
    try:
        os.dup2(devnull, fd)
        yield
    finally:
        os.dup2(saved, fd)
        os.close(saved)
        os.close(devnull)


def gdsjoin(DATASET, EPSILON, NDIM, DTYPE, verbose=False, params=None, gpu=0):
    if verbose:
        return gdsjoinmain(DATASET, EPSILON, NDIM, DTYPE, params=params, gpu=gpu)
    with _hide_c_stdout():
        return gdsjoinmain(DATASET, EPSILON, NDIM, DTYPE, params=params, gpu=gpu)


def gdsjoinmain(DATASET, EPSILON, NDIM, DTYPE, params=None, gpu=0):
    L = _lib.load()
    if DTYPE not in ("float", "double"):
        raise ValueError('DTYPE must be "float" or "double"')
    NDIM = int(NDIM)
    if NDIM <= 0:
        raise ValueError("NDIM must be positive")
    npdt = np.float32 if DTYPE == "float" else np.float64
    src = np.asarray(DATASET)
    if src.dtype in (np.float32, np.float64):
        # floating input: a direct cast is the same as through float64
        data = np.ascontiguousarray(src.ravel(), dtype=npdt)
    else:
        data = np.ascontiguousarray(np.asarray(src, dtype=np.float64).ravel().astype(npdt))
    if data.size % NDIM:
        raise ValueError("len(DATASET)=%d is not a multiple of NDIM=%d" % (data.size, NDIM))
    num_points = data.size // NDIM
    if params is None:
        params = _lib.default_params()
    L.gdsj_set_legacy_params(C.byref(params))
    counts = np.zeros(num_points, dtype=np.uint32)
    if DTYPE == "float":
        L.GDSJoinPy(data.ctypes.data, num_points, C.c_float(np.float32(EPSILON)), NDIM, int(gpu),
                    counts.ctypes.data)
    else:
        L.GDSJoinPyDouble(data.ctypes.data, num_points, C.c_double(float(EPSILON)), NDIM, int(gpu),
                          counts.ctypes.data)
    rc = L.gdsj_legacy_status()
    if rc != _lib.GDSJ_OK:
        raise _lib.GdsjError(rc, _lib.last_error())
    total = int(counts.sum(dtype=np.uint64))
    if total >= 2 ** 32:
        raise _lib.GdsjError(_lib.GDSJ_ERANGE, "table of %d ids exceeds the uint32 count of the C ABI" % total)
    table = np.empty(total, dtype=np.uint32)
    L.copyResultIntoPythonArray(table.ctypes.data, total)
    rc = L.gdsj_legacy_status()
    if rc != _lib.GDSJ_OK:
        raise _lib.GdsjError(rc, _lib.last_error())
    return counts, table
