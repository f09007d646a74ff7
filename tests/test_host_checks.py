"""CPU tests of the Python layer's device-argument checks: a device array of
the wrong dtype or element count is rejected with ValueError before any C
call, because the engine reads and writes device memory by the counts it is
given.  The arrays here are stand-ins that hold no device memory; every case
raises before a pointer would be used, so no GPU is touched."""
import numpy as np
import pytest

import lbvh_amd
from lbvh_amd import hip


class FakeDevice(hip.DeviceArray):
    """A DeviceArray look-alike with a dtype and shape but no memory."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = 1 << 40  # never dereferenced: every case raises first

    def close(self):
        self.ptr = None


def _ctx():
    c = object.__new__(lbvh_amd.Context)
    c.handle = None
    c.device = 0
    return c


def _bvh(ctx, n):
    return lbvh_amd.Bvh(ctx, None, n, 63, 1)


def test_build_rejects_wrong_dtype():
    ctx = _ctx()
    with pytest.raises(ValueError, match="float32"):
        ctx.build_triangles(FakeDevice((100, 12), np.float64))
    with pytest.raises(ValueError, match="multiple of 12"):
        ctx.build_triangles(FakeDevice((100 * 12 + 5,), np.float32))
    with pytest.raises(ValueError, match="float32"):
        ctx.build_aabbs(FakeDevice((100, 6), np.int32))


def test_build_f64_rejects_float32_boxes():
    """A float32 buffer passed as FP64 boxes would make the kernels read
    twice its bytes."""
    ctx = _ctx()
    with pytest.raises(ValueError, match="float64"):
        ctx.build_aabbs_f64(FakeDevice((1000, 6), np.float32))
    with pytest.raises(ValueError, match="64 \\(n - 1\\) bytes"):
        ctx.build_aabbs_f64(FakeDevice((1000, 6), np.float64),
                            out=FakeDevice((998,), lbvh_amd.NODE64_DTYPE))


def test_refit_rejects_other_counts():
    ctx = _ctx()
    b = _bvh(ctx, 1000)
    with pytest.raises(ValueError, match="expected 12000"):
        ctx.refit_triangles(b, FakeDevice((999, 12), np.float32))
    with pytest.raises(ValueError, match="float32"):
        ctx.refit_triangles(b, FakeDevice((1000, 12), np.float64))


def test_trace_rejects_short_outputs():
    ctx = _ctx()
    b = _bvh(ctx, 1000)
    rays = FakeDevice((500, 8), np.float32)
    ok = (FakeDevice(500, np.uint32), FakeDevice(500, np.float32), FakeDevice((500, 2), np.float32))
    with pytest.raises(ValueError, match="prim"):
        ctx.trace(b, rays, out=(FakeDevice(499, np.uint32), ok[1], ok[2]))
    with pytest.raises(ValueError, match="t:"):
        ctx.trace(b, rays, out=(ok[0], FakeDevice(500, np.float64), ok[2]))
    with pytest.raises(ValueError, match="uv"):
        ctx.trace(b, rays, out=(ok[0], ok[1], FakeDevice(500, np.float32)))
    with pytest.raises(ValueError, match="rays"):
        ctx.trace(b, FakeDevice((500, 8), np.float16), out=ok)


def test_batch_rejects_bad_device_arguments():
    ctx = _ctx()
    off = np.array([0, 100, 300], np.uint64)
    with pytest.raises(ValueError, match="expected 3600"):
        ctx.build_batch(FakeDevice((299, 12), np.float32), off)
    batch = lbvh_amd.Batch(ctx, None, off, 1)
    with pytest.raises(ValueError, match="expected 3600"):
        batch.refit(FakeDevice((301, 12), np.float32))
    with pytest.raises(ValueError, match="expected 3600"):
        batch.rebuild(FakeDevice((300, 12), np.float32).rows(0, 299))
    with pytest.raises(ValueError, match="ray offsets"):
        batch.trace(FakeDevice((10, 8), np.float32), FakeDevice(2, np.uint64))
    batch.handle = None


def test_server_rejects_bad_device_request():
    srv = object.__new__(lbvh_amd.Server)
    srv.handle = None
    srv._live = {}
    rays = FakeDevice((64, 8), np.float32)
    with pytest.raises(ValueError, match="triangles"):
        srv.submit(rays, 0, FakeDevice((100, 12), np.float64))
    with pytest.raises(ValueError, match="prim"):
        srv.submit(rays, 0, FakeDevice((100, 12), np.float32),
                   out=(FakeDevice(63, np.uint32), FakeDevice(64, np.float32), None))
