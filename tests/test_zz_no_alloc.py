"""GPU test of the allocation-free build, lbvh_build_triangles_noalloc,
through the C ABI: its node array equals lbvh_build_triangles_device's for
the same triangles and options, bit for bit, on two calls in a row (the
second smaller, so the scratch array holds the first call's contents), for
both code widths and leaf orders, up to the maximum size; sizes beyond the
maximum and misaligned outputs are rejected on the host."""
import ctypes

import numpy as np
import pytest

import lbvh_amd
from lbvh_amd import hip
from lbvh_amd import workload as W

pytestmark = pytest.mark.gpu

MAX = 262144


def _noalloc(ctx, d_tri, n, opts, d_nodes):
    return lbvh_amd.lib().lbvh_build_triangles_noalloc(ctx.handle, ctypes.c_void_p(d_tri.ptr), n,
                                                      ctypes.byref(opts),
                                                      ctypes.c_void_p(d_nodes.ptr))


def _device_build_nodes(ctx, d_tri, n, bits, leaf):
    out = ctypes.c_void_p()
    opts = lbvh_amd._make_opts(bits, leaf, False)
    ctx._check(lbvh_amd.lib().lbvh_build_triangles_device(ctx.handle, ctypes.c_void_p(d_tri.ptr),
                                                         n, ctypes.byref(opts),
                                                         ctypes.byref(out)))
    b = lbvh_amd.Bvh(ctx, out, n, bits, leaf)
    nodes = b.nodes()
    b.destroy()
    return nodes


@pytest.mark.parametrize("bits,leaf", [(30, 0), (63, 1), (63, 0), (30, 1)])
def test_noalloc_equals_device_build_twice(ctx, bits, leaf):
    sizes = [MAX, 70001] if (bits, leaf) == (63, 1) else [200000, 4097]
    d_nodes = hip.DeviceArray((MAX - 1,), lbvh_amd.NODE_DTYPE)
    for k, n in enumerate(sizes):
        print("noalloc build bits=%d leaf=%d call=%d n=%d" % (bits, leaf, k, n), flush=True)
        tri = W.make_triangles(W.cluster_vertices(4000 + 17 * k + bits + leaf, n))
        d_tri = hip.DeviceArray.from_host(tri)
        opts = lbvh_amd._make_opts(bits, leaf, False)
        ctx._check(_noalloc(ctx, d_tri, n, opts, d_nodes))
        ctx.synchronize()
        got = d_nodes.numpy()[:n - 1]
        want = _device_build_nodes(ctx, d_tri, n, bits, leaf)
        assert got.tobytes() == want.tobytes(), "call %d (n=%d)" % (k, n)
        d_tri.close()
    d_nodes.close()


def test_noalloc_rejects_sizes_and_alignment(ctx):
    d_tri = hip.DeviceArray((1024, 12), np.float32)
    d_nodes = hip.DeviceArray((1023,), lbvh_amd.NODE_DTYPE)
    opts = lbvh_amd._make_opts(30, 0, False)
    lib = lbvh_amd.lib()
    # more than the maximum: rejected before any launch (the buffer is not read)
    assert _noalloc(ctx, d_tri, MAX + 1, opts, d_nodes) == 1
    assert b"LBVH_NOALLOC_MAX_PRIMS" in lib.lbvh_ctx_last_error(ctx.handle)
    assert _noalloc(ctx, d_tri, 1, opts, d_nodes) == 1
    assert lib.lbvh_build_triangles_noalloc(ctx.handle, ctypes.c_void_p(d_tri.ptr + 4), 100,
                                            ctypes.byref(opts), ctypes.c_void_p(d_nodes.ptr)) == 1
    assert lib.lbvh_build_triangles_noalloc(ctx.handle, ctypes.c_void_p(d_tri.ptr), 100,
                                            ctypes.byref(opts),
                                            ctypes.c_void_p(d_nodes.ptr + 16)) == 1
    keep = lbvh_amd._make_opts(30, 0, True)
    assert _noalloc(ctx, d_tri, 100, keep, d_nodes) == 1
    onesweep = lbvh_amd._make_opts(30, 0, False, sort_strategy=lbvh_amd.SORT_ONESWEEP)
    assert _noalloc(ctx, d_tri, 100, onesweep, d_nodes) == 1
    d_tri.close()
    d_nodes.close()
