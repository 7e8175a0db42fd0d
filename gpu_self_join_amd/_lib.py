"""ctypes binding of libgdsjoin.so (the C ABI declared in include/gdsjoin.h).

The library is the only compute path: when it is missing this module raises
instead of falling back to anything else.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libgdsjoin.so"

GDSJ_OK = 0
GDSJ_EINVAL = -1
GDSJ_ERANGE = -2
GDSJ_EOOM = -3
GDSJ_EDEVICE = -4
GDSJ_EOVERFLOW = -5
GDSJ_ESTATE = -6
GDSJ_EUNSUPPORTED = -7

DTYPE_F32 = 0
DTYPE_F64 = 1

FILTER_PLAIN, FILTER_GRAM, FILTER_AUTO, FILTER_GRAM_BF16 = 0, 1, 2, 3
EST_EXACT, EST_LEARNED, EST_SAMPLE = 0, 1, 2

# every symbol include/gdsjoin.h declares
EXPORTS = [
    "gdsj_params_default", "gdsj_create", "gdsj_destroy", "gdsj_join_device", "gdsj_fill_device",
    "gdsj_join_host", "gdsj_result_copy", "gdsj_get_stats", "gdsj_eps_threshold_f32",
    "gdsj_eps_threshold_f64", "gdsj_last_error", "GDSJoinPy", "GDSJoinPyDouble",
    "copyResultIntoPythonArray", "gdsj_set_legacy_params", "gdsj_legacy_status",
    "gdsj_mlp_param_count", "gdsj_mlp_feature_count", "gdsj_mlp_forward", "gdsj_mlp_grad",
    "gdsj_model_get", "gdsj_model_set",
    "gdsj_batch_create", "gdsj_batch_destroy", "gdsj_batch_autotune", "gdsj_batch_run", "gdsj_batch_get_stats",
    "gdsj_csv_load", "gdsj_bruteforce_device", "gdsj_bruteforce_host", "gdsj_rejoin_device",
    "gdsj_bruteforce_device_noalloc", "gdsj_bruteforce_rows_host",
]

NOALLOC_SCRATCH_BYTES = 32 << 20  # GDSJ_NOALLOC_SCRATCH_BYTES

ILP_AUTO = 0xFFFFFFFF


class GdsjParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("ilp", C.c_uint32), ("shortcircuit", C.c_uint32),
                ("reorder", C.c_uint32), ("stamp", C.c_uint32), ("queryreorder", C.c_uint32),
                ("filter", C.c_uint32), ("estimator", C.c_uint32), ("sort_neighbors", C.c_uint32),
                ("verbose", C.c_uint32), ("tile_queries", C.c_uint32), ("schedule", C.c_uint32),
                ("prune", C.c_uint32), ("train", C.c_uint32), ("filter_stats", C.c_uint32),
                ("hitmask", C.c_uint32), ("ilp_fold", C.c_uint32),
                ("vfilter", C.c_uint32), ("sort", C.c_uint32),
                ("cpu_sum", C.c_uint32), ("mfma", C.c_uint32),
                ("ffilter", C.c_uint32), ("index", C.c_uint32),
                ("index_dims", C.c_uint32), ("gram_axes", C.c_uint32)]


class GdsjStats(C.Structure):
    _fields_ = [("total_pairs", C.c_uint64), ("candidates", C.c_uint64), ("n_nonempty_cells", C.c_uint64),
                ("total_cells", C.c_uint64), ("n_tiles", C.c_uint64), ("n_runs", C.c_uint64),
                ("filter_rechecks", C.c_uint64), ("filter_hits", C.c_uint64), ("est_predicted", C.c_uint64),
                ("est_capacity", C.c_uint64), ("est_overflow_tiles", C.c_uint64), ("est_tiles", C.c_uint64),
                ("est_mean_rel_err", C.c_double), ("est_max_rel_err", C.c_double), ("filter_max_err", C.c_double),
                ("ms_grid", C.c_double), ("ms_estimate", C.c_double), ("ms_count", C.c_double),
                ("ms_scan", C.c_double), ("ms_compact", C.c_double), ("ms_fill", C.c_double),
                ("ms_sort", C.c_double), ("ms_train", C.c_double), ("ms_total", C.c_double),
                ("perm", C.c_uint32 * 64), ("mfma_tiles", C.c_uint64),
                ("index_used", C.c_uint64), ("index_risky", C.c_uint64),
                ("gram_axes", C.c_uint64)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "perm"}
        d["perm"] = list(self.perm)
        return d


class GdsjInstance(C.Structure):
    _fields_ = [("d_data", C.c_void_p), ("n", C.c_uint32), ("dims", C.c_uint32), ("epsilon", C.c_double)]


class GdsjBatchTuning(C.Structure):
    _fields_ = [("tile_queries", C.c_uint32), ("waves_per_block", C.c_uint32), ("ilp", C.c_uint32),
                ("lpt", C.c_uint32), ("blocks", C.c_uint32), ("reg_budget", C.c_uint32), ("ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class GdsjBatchStats(C.Structure):
    _fields_ = [("total_pairs", C.c_uint64), ("candidates", C.c_uint64), ("n_points", C.c_uint64),
                ("n_tiles", C.c_uint64), ("est_predicted", C.c_uint64), ("est_capacity", C.c_uint64),
                ("est_overflow_tiles", C.c_uint64), ("slotted", C.c_uint32), ("graph_replays", C.c_uint32),
                ("graph_captures", C.c_uint32), ("reserved", C.c_uint32),
                ("ms_build", C.c_double), ("ms_tune", C.c_double), ("ms_train", C.c_double),
                ("ms_last_run", C.c_double), ("tuning", GdsjBatchTuning)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "tuning"}
        d["tuning"] = self.tuning.as_dict()
        return d


class GdsjError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gdsjoin error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib_path():
    return os.path.join(_HERE, LIB_NAME)


def load():
    """Loads libgdsjoin.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise ImportError("%s not built: run __graft_entry__.build() or make -C gpu_self_join_amd/csrc" % path)
    L = C.CDLL(path)
    P = C.c_void_p
    u32, u64, i32 = C.c_uint32, C.c_uint64, C.c_int
    L.gdsj_params_default.argtypes = [C.POINTER(GdsjParams)]
    L.gdsj_create.argtypes = [i32, C.POINTER(P)]
    L.gdsj_destroy.argtypes = [P]
    L.gdsj_join_device.argtypes = [P, P, u32, u32, i32, C.c_double, C.POINTER(GdsjParams), P, P, P, u64,
                                   C.POINTER(u64), P]
    L.gdsj_fill_device.argtypes = [P, P, u64]
    L.gdsj_join_host.argtypes = [P, P, u32, u32, i32, C.c_double, C.POINTER(GdsjParams), P, C.POINTER(u64)]
    L.gdsj_result_copy.argtypes = [P, P, u64]
    L.gdsj_get_stats.argtypes = [P, C.POINTER(GdsjStats)]
    L.gdsj_eps_threshold_f32.argtypes = [C.c_float]
    L.gdsj_eps_threshold_f32.restype = C.c_float
    L.gdsj_eps_threshold_f64.argtypes = [C.c_double]
    L.gdsj_eps_threshold_f64.restype = C.c_double
    L.gdsj_last_error.restype = C.c_char_p
    L.GDSJoinPy.argtypes = [P, C.c_uint, C.c_float, C.c_uint, C.c_int, P]
    L.GDSJoinPy.restype = None
    L.GDSJoinPyDouble.argtypes = [P, C.c_uint, C.c_double, C.c_uint, C.c_int, P]
    L.GDSJoinPyDouble.restype = None
    L.copyResultIntoPythonArray.argtypes = [P, C.c_uint]
    L.copyResultIntoPythonArray.restype = None
    L.gdsj_set_legacy_params.argtypes = [C.POINTER(GdsjParams)]
    L.gdsj_legacy_status.restype = C.c_int
    L.gdsj_mlp_param_count.restype = u32
    L.gdsj_mlp_feature_count.restype = u32
    L.gdsj_mlp_forward.argtypes = [P, P, u32, P]
    L.gdsj_mlp_grad.argtypes = [P, P, P, u32, P, P]
    L.gdsj_model_get.argtypes = [P, P, u32, C.POINTER(C.c_double)]
    L.gdsj_model_set.argtypes = [P, P, u32, C.c_double]
    L.gdsj_batch_create.argtypes = [i32, C.POINTER(GdsjInstance), u32, i32, C.POINTER(GdsjParams), C.POINTER(P)]
    L.gdsj_batch_destroy.argtypes = [P]
    L.gdsj_batch_autotune.argtypes = [P, C.POINTER(GdsjBatchTuning)]
    L.gdsj_batch_run.argtypes = [P, P, P, P, u64, C.POINTER(u64), P]
    L.gdsj_batch_get_stats.argtypes = [P, C.POINTER(GdsjBatchStats)]
    L.gdsj_csv_load.argtypes = [C.c_char_p, i32, P, u64, C.POINTER(u64)]
    L.gdsj_bruteforce_device.argtypes = [P, P, u32, u32, i32, C.c_double, C.POINTER(GdsjParams), P, P, P, u64,
                                         C.POINTER(u64), P]
    L.gdsj_rejoin_device.argtypes = [P, P, P, P, u64, C.POINTER(u64), P]
    L.gdsj_bruteforce_rows_host.argtypes = [P, P, u64]
    L.gdsj_bruteforce_device_noalloc.argtypes = [P, u32, u32, i32, C.c_double, C.POINTER(GdsjParams), P, P, P,
                                                 u64, C.POINTER(u64), P]
    L.gdsj_bruteforce_host.argtypes = [P, P, u32, u32, i32, C.c_double, C.POINTER(GdsjParams), P,
                                       C.POINTER(u64)]
    _lib = L
    return L


def last_error():
    m = load().gdsj_last_error()
    return m.decode() if m else ""


def check(rc):
    if rc != GDSJ_OK:
        raise GdsjError(rc, last_error())
    return rc


def csv_load(path, dtype="float"):
    """Values of a text dataset read as the reference's importNDDataset reads them (flat array)."""
    import numpy as np
    dt = DTYPE_F64 if dtype in ("double", np.float64) else DTYPE_F32
    npdt = np.float64 if dt == DTYPE_F64 else np.float32
    L = load()
    n = C.c_uint64(0)
    check(L.gdsj_csv_load(os.fsencode(path), dt, None, 0, C.byref(n)))
    out = np.empty(n.value, dtype=npdt)
    check(L.gdsj_csv_load(os.fsencode(path), dt, out.ctypes.data if n.value else None, n.value, C.byref(n)))
    return out


def default_params(**kw):
    p = GdsjParams()
    load().gdsj_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError("unknown gdsj_params field %r" % k)
        setattr(p, k, int(v))
    return p
