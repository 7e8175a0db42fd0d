"""The allocation-free all-pairs join (gdsj_bruteforce_device_noalloc),
called through the C ABI, against the allocating entry point
(gdsj_bruteforce_device): identical counts, offsets and rows on two calls in
a row, the second with a smaller input (the shared scratch array keeps the
first call's contents), and a clear error for inputs beyond its scratch."""
import ctypes as C

import numpy as np
import pytest

from gpu_self_join_amd import _lib

pytestmark = pytest.mark.gpu


def _noalloc(x, eps, cap, **kw):
    import torch
    L = _lib.load()
    pts = torch.from_numpy(x).cuda()
    n, d = x.shape
    counts = torch.empty(n, dtype=torch.int32, device="cuda")
    offsets = torch.empty(n + 1, dtype=torch.int64, device="cuda")
    ids = torch.empty(max(cap, 1), dtype=torch.int32, device="cuda")
    total = C.c_uint64(0)
    dt = _lib.DTYPE_F64 if x.dtype == np.float64 else _lib.DTYPE_F32
    p = _lib.default_params(**kw)
    torch.cuda.synchronize()
    rc = L.gdsj_bruteforce_device_noalloc(pts.data_ptr(), n, d, dt, float(eps), C.byref(p), counts.data_ptr(),
                                          offsets.data_ptr(), ids.data_ptr(), cap, C.byref(total), None)
    return rc, counts.cpu().numpy().view(np.uint32), offsets.cpu().numpy(), ids.cpu().numpy().view(np.uint32), \
        total.value


def _reference(x, eps, **kw):
    import torch

    from gpu_self_join_amd import SelfJoin
    sj = SelfJoin(0)
    try:
        r = sj.bruteforce(torch.from_numpy(x).cuda(), eps, **kw)
        return (r.counts.cpu().numpy().view(np.uint32), r.offsets.cpu().numpy(),
                r.ids.cpu().numpy().view(np.uint32), r.total)
    finally:
        sj.close()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_noalloc_equals_allocating_entry_point_twice(gpu, dt):
    rng = np.random.default_rng(41)
    big = (rng.random((6000, 5)) * 10.0).astype(dt)
    small = (rng.random((1700, 5)) * 10.0).astype(dt)
    eps = float(np.asarray(1.1, dtype=dt))
    kw = dict(ilp=8, shortcircuit=1)
    for x in (big, small):
        c0, o0, i0, t0 = _reference(x, eps, **kw)
        rc, c1, o1, i1, t1 = _noalloc(x, eps, int(t0), **kw)
        print("noalloc n=%d total=%d rc=%d" % (len(x), t1, rc), flush=True)
        assert rc == _lib.GDSJ_OK, _lib.last_error()
        assert t1 == t0
        assert np.array_equal(c1, c0) and np.array_equal(o1, o0)
        assert np.array_equal(i1[:t1], i0)


def test_noalloc_rejects_inputs_beyond_its_scratch(gpu):
    L = _lib.load()
    n, d = 200_000, 64  # 51 MB of float SoA copy > 32 MiB
    assert n * d * 4 > _lib.NOALLOC_SCRATCH_BYTES
    total = C.c_uint64(0)
    p = _lib.default_params()
    dummy = C.c_void_p(1)  # never dereferenced: the size check comes first
    rc = L.gdsj_bruteforce_device_noalloc(dummy, n, d, _lib.DTYPE_F32, 1.0, C.byref(p), dummy, dummy, None, 0,
                                          C.byref(total), None)
    assert rc == _lib.GDSJ_EINVAL
    assert "scratch" in _lib.last_error()
