"""Handle-based self-join on device-resident torch tensors.

    sj = SelfJoin(device=0)
    res = sj.join(points, eps, k=2, ilp=0)   # points: (N, D) float32/float64 cuda tensor
    res.counts, res.offsets, res.ids, res.total

Wraps gdsj_create / gdsj_join_device / gdsj_fill_device (include/gdsjoin.h).
torch provides device memory and the stream; every kernel is in
libgdsjoin.so.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class JoinResult:
    counts: object   # uint32 [N] (torch.int32 storage, reinterpret as uint32)
    offsets: object  # int64 [N+1] CSR offsets
    ids: object      # int32 [total] neighbor ids (uint32 values)
    total: int
    stats: dict


def _torch():
    import torch
    return torch


class SelfJoin:
    def __init__(self, device=0):
        self.L = _lib.load()
        h = C.c_void_p()
        _lib.check(self.L.gdsj_create(int(device), C.byref(h)))
        self.h = h
        self.device = int(device)

    def close(self):
        if self.h:
            self.L.gdsj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        st = _lib.GdsjStats()
        _lib.check(self.L.gdsj_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    def join(self, points, eps, params=None, out_ids=None, counts=None, offsets=None, stream=None, **kw):
        """Runs the join on `points` (cuda tensor, (N, D), float32 or float64).

        out_ids / counts / offsets may be preallocated (int32 [cap], int32 [N],
        int64 [N+1]); when out_ids is too small the table is completed into a
        fresh tensor without recomputing the grid or counts."""
        torch = _torch()
        if points.dim() != 2 or not points.is_cuda:
            raise ValueError("points must be a 2-D cuda tensor")
        points = points.contiguous()
        n, d = points.shape
        if points.dtype == torch.float32:
            dt = _lib.DTYPE_F32
            eps = float(np.float32(eps))
        elif points.dtype == torch.float64:
            dt = _lib.DTYPE_F64
            eps = float(eps)
        else:
            raise TypeError("points must be float32 or float64")
        if params is None:
            params = _lib.default_params(**kw)
        elif kw:
            raise TypeError("pass either params or keyword fields")
        dev = points.device
        if counts is None:
            counts = torch.empty(n, dtype=torch.int32, device=dev)
        if offsets is None:
            offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
        cap = 0 if out_ids is None else out_ids.numel()
        ids_ptr = None if out_ids is None else out_ids.data_ptr()
        total = C.c_uint64(0)
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        rc = self.L.gdsj_join_device(self.h, points.data_ptr(), n, d, dt, eps, C.byref(params),
                                     counts.data_ptr(), offsets.data_ptr(), ids_ptr, cap, C.byref(total),
                                     stream)
        tot = int(total.value)
        if rc == _lib.GDSJ_EOVERFLOW or (out_ids is None and rc == _lib.GDSJ_OK):
            out_ids = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
            _lib.check(self.L.gdsj_fill_device(self.h, out_ids.data_ptr(), out_ids.numel()))
        else:
            _lib.check(rc)
        return JoinResult(counts=counts, offsets=offsets, ids=out_ids[:tot], total=tot, stats=self.stats())
