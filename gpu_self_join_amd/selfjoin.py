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

    def rejoin(self, out_ids=None, counts=None, offsets=None, stream=None, n=None):
        """Runs the last join again over the grid it built (the points must
        be unchanged since): the counting or slotted pass, scan and fill.
        n: the number of points of that join (for the output tensors)."""
        torch = _torch()
        dev = torch.device("cuda", self.device)
        if counts is None or offsets is None:
            if n is None:
                raise ValueError("pass counts and offsets, or n")
            counts = torch.empty(n, dtype=torch.int32, device=dev) if counts is None else counts
            offsets = torch.empty(n + 1, dtype=torch.int64, device=dev) if offsets is None else offsets
        cap = 0 if out_ids is None else out_ids.numel()
        ids_ptr = None if out_ids is None else out_ids.data_ptr()
        total = C.c_uint64(0)
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        rc = self.L.gdsj_rejoin_device(self.h, counts.data_ptr(), offsets.data_ptr(), ids_ptr, cap, C.byref(total),
                                       stream)
        tot = int(total.value)
        if rc == _lib.GDSJ_EOVERFLOW or (out_ids is None and rc == _lib.GDSJ_OK):
            out_ids = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
            _lib.check(self.L.gdsj_fill_device(self.h, out_ids.data_ptr(), out_ids.numel()))
        else:
            _lib.check(rc)
        return JoinResult(counts=counts, offsets=offsets, ids=out_ids[:tot], total=tot, stats=self.stats())

    def bruteforce(self, points, eps, params=None, stream=None, **kw):
        """All-pairs join of `points` (the reference's search mode 1,
        GPU.cu:1285-1480): every ordered pair tested with the join's
        arithmetic for params.ilp / shortcircuit, dims in input order.
        Returns a JoinResult whose rows are ascending (stats: none)."""
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
        counts = torch.empty(n, dtype=torch.int32, device=dev)
        offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
        total = C.c_uint64(0)
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(self.L.gdsj_bruteforce_device(self.h, points.data_ptr(), n, d, dt, eps, C.byref(params),
                                                 counts.data_ptr(), offsets.data_ptr(), None, 0, C.byref(total),
                                                 stream))
        tot = int(total.value)
        ids = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
        _lib.check(self.L.gdsj_bruteforce_device(self.h, points.data_ptr(), n, d, dt, eps, C.byref(params),
                                                 counts.data_ptr(), offsets.data_ptr(), ids.data_ptr(), ids.numel(),
                                                 C.byref(total), stream))
        return JoinResult(counts=counts, offsets=offsets, ids=ids[:tot], total=tot, stats={})


@dataclass
class BatchResult:
    counts: object     # uint32 [sum n] (torch.int32 storage)
    offsets: object    # int64 [sum n + 1], positions in ids
    ids: object        # int32 [total]: instance-local neighbor ids
    row_base: list     # row_base[i] = first row of instance i (len ninst + 1)
    total: int
    stats: dict

    def instance(self, i):
        """(counts, offsets rebased to 0, ids) of instance i, as torch views."""
        a, b = self.row_base[i], self.row_base[i + 1]
        off = self.offsets[a:b + 1]
        lo = int(off[0].item())
        hi = int(off[-1].item())
        return self.counts[a:b], off - lo, self.ids[lo:hi]


class BatchJoin:
    """Many independent self-joins run as one job (gdsj_batch_*).

        bj = BatchJoin([x0, x1, ...], [eps0, eps1, ...], ilp=_lib.ILP_AUTO)
        bj.autotune()
        res = bj.run()          # res.instance(i) -> (counts, offsets, ids)

    The point tensors (cuda, (n_i, d_i), all float32 or all float64) must stay
    alive and unchanged while the batch exists: the grids are built from them
    once, in the constructor."""

    def __init__(self, points, eps, device=0, params=None, **kw):
        torch = _torch()
        self.L = _lib.load()
        if len(points) != len(eps):
            raise ValueError("one epsilon per instance")
        self.points = [p.contiguous() for p in points]
        dts = {p.dtype for p in self.points}
        if len(dts) > 1:
            raise TypeError("all instances must share one dtype")
        dtype = dts.pop() if dts else torch.float32
        if dtype == torch.float32:
            self.dt = _lib.DTYPE_F32
            eps = [float(np.float32(e)) for e in eps]
        elif dtype == torch.float64:
            self.dt = _lib.DTYPE_F64
            eps = [float(e) for e in eps]
        else:
            raise TypeError("points must be float32 or float64")
        if params is None:
            params = _lib.default_params(**kw)
        elif kw:
            raise TypeError("pass either params or keyword fields")
        self.params = params
        m = len(self.points)
        arr = (_lib.GdsjInstance * max(m, 1))()
        self.row_base = [0]
        for i, (p, e) in enumerate(zip(self.points, eps)):
            if p.dim() != 2 or not p.is_cuda:
                raise ValueError("instance %d: points must be a 2-D cuda tensor" % i)
            arr[i].d_data = p.data_ptr()
            arr[i].n = p.shape[0]
            arr[i].dims = p.shape[1]
            arr[i].epsilon = e
            self.row_base.append(self.row_base[-1] + p.shape[0])
        self.device = torch.device("cuda", int(device))
        h = C.c_void_p()
        _lib.check(self.L.gdsj_batch_create(int(device), arr, m, self.dt, C.byref(params), C.byref(h)))
        self.h = h
        self.npts = self.row_base[-1]
        self._ids = self._counts = self._offsets = None

    def close(self):
        if self.h:
            self.L.gdsj_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self):
        t = _lib.GdsjBatchTuning()
        _lib.check(self.L.gdsj_batch_autotune(self.h, C.byref(t)))
        return t.as_dict()

    def stats(self):
        st = _lib.GdsjBatchStats()
        _lib.check(self.L.gdsj_batch_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    def run(self, out_ids=None, counts=None, offsets=None, stream=None):
        """Runs every join; grows the id buffer when the result does not fit.

        Without explicit outputs the batch's own buffers are reused, so the
        captured graph is replayed; each run overwrites the previous result."""
        torch = _torch()
        dev = self.device
        if counts is None:
            if self._counts is None:
                self._counts = torch.empty(max(self.npts, 1), dtype=torch.int32, device=dev)
            counts = self._counts
        if offsets is None:
            if self._offsets is None:
                self._offsets = torch.empty(self.npts + 1, dtype=torch.int64, device=dev)
            offsets = self._offsets
        if out_ids is None:
            if self._ids is None:
                self._ids = torch.empty(1, dtype=torch.int32, device=dev)
            out_ids = self._ids
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        total = C.c_uint64(0)
        rc = self.L.gdsj_batch_run(self.h, counts.data_ptr(), offsets.data_ptr(), out_ids.data_ptr(),
                                   out_ids.numel(), C.byref(total), stream)
        tot = int(total.value)
        if rc == _lib.GDSJ_EOVERFLOW:
            out_ids = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
            self._ids = out_ids
            rc = self.L.gdsj_batch_run(self.h, counts.data_ptr(), offsets.data_ptr(), out_ids.data_ptr(),
                                       out_ids.numel(), C.byref(total), stream)
            tot = int(total.value)
        _lib.check(rc)
        return BatchResult(counts=counts[:self.npts], offsets=offsets, ids=out_ids[:tot], row_base=self.row_base,
                           total=tot, stats=self.stats())
