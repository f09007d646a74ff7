"""GPU parity of lbvh::builder<double> (lbvh_build_aabbs_f64 and its device
variant) through the C ABI: node<double> arrays bit-exact against the golden
vectors written by the real builder<double> and against the oracle on
seeded inputs (ties, clusters, flat axes, 1e6-scale and tiny extents, both
leaf orders), device-resident input, alternation with FP32 builds on one
context, and argument errors."""
import numpy as np
import pytest

import lbvh_amd
import oracle
from conftest import f64_golden_cases

pytestmark = pytest.mark.gpu

CASES = f64_golden_cases()
IDS = [c[0] for c in CASES]


def _boxes(kind, prims):
    return oracle.triangle_aabb_f64(prims) if kind == "triangles" else prims


def _scene(dist, n, seed):
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        lo = rng.random((n, 3))
        ext = rng.random((n, 3)) * 0.01
    elif dist == "clustered":
        centres = rng.normal(size=(32, 3)) * 10.0
        lo = centres[rng.integers(0, 32, size=n)] + rng.normal(size=(n, 3)) * 0.05
        ext = rng.random((n, 3)) * 1e-3
    elif dist == "duplicates":
        base = rng.random((max(n // 16, 1), 3))
        lo = np.repeat(base, 16, axis=0)[:n]
        if lo.shape[0] < n:
            lo = np.concatenate([lo, base[: n - lo.shape[0]]])
        ext = np.full((n, 3), 1e-3)
    elif dist == "flat":
        lo = rng.random((n, 3))
        lo[:, 1] = -2.5
        ext = rng.random((n, 3)) * 0.01
        ext[:, 1] = 0.0
    elif dist == "huge":
        lo = rng.uniform(-1e6, 1e6, size=(n, 3))
        ext = rng.random((n, 3)) * 1e3
    else:  # tiny: extents far below one Morton cell of FP32 codes
        lo = 0.5 + rng.random((n, 3)) * 1e-9
        ext = rng.random((n, 3)) * 1e-12
    return np.ascontiguousarray(np.concatenate([lo, lo + ext], axis=1))


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
def test_f64_build_matches_golden(ctx, case, leaf):
    _, kind, prims, want = case
    got = ctx.build_aabbs_f64(_boxes(kind, prims), leaf)
    assert np.array_equal(got.view(np.uint64).reshape(-1, 8), want[leaf])


@pytest.mark.parametrize("n", [2, 3, 4, 255, 256, 257, 511, 512, 513, 1000, 1025, 2049, 4097, 65537,
                               300001])
@pytest.mark.parametrize("dist", ["uniform", "clustered", "duplicates", "flat", "huge", "tiny"])
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
def test_f64_build_matches_oracle(ctx, n, dist, leaf):
    print("f64 build n=%d dist=%s leaf=%d" % (n, dist, leaf), flush=True)
    boxes = _scene(dist, n, 1000 + n)
    want, _, _ = oracle.build_f64(boxes, leaf)
    got = ctx.build_aabbs_f64(boxes, leaf)
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("strategy,digits,tile", [(1, 8, 0), (2, 8, 0), (2, 10, 0), (3, 8, 0),
                                                  (3, 10, 4096)])
def test_f64_every_sort_configuration_matches_oracle(ctx, strategy, digits, tile):
    """Each sort strategy (onesweep, reduce-then-scan, hybrid) and digit
    width gives the same tree: the sort is stable by primitive id."""
    print("f64 sort strategy=%d digits=%d tile=%d" % (strategy, digits, tile), flush=True)
    boxes = _scene("duplicates", 250007, 77)
    want, _, _ = oracle.build_f64(boxes, 1)
    got = ctx.build_aabbs_f64(boxes, 1, sort_strategy=strategy, sort_digit_bits=digits,
                              sort_tile=tile)
    assert got.tobytes() == want.tobytes()


def test_f64_tuner_times_candidates_on_a_fresh_context():
    c = lbvh_amd.Context(0)
    boxes = _scene("clustered", 150001, 78)
    want, _, _ = oracle.build_f64(boxes, 0)
    assert c.build_aabbs_f64(boxes, 0).tobytes() == want.tobytes()
    log = c.tune_log()
    assert len(log) >= 4 and sum(r["picked"] for r in log) >= 1
    assert all(r["key_bits"] == 60 for r in log)
    c.close()


def test_f64_large_matches_oracle(ctx):
    n = 4_000_037
    print("f64 build n=%d clustered" % n, flush=True)
    boxes = _scene("clustered", n, 7)
    want, _, _ = oracle.build_f64(boxes, 1)
    got = ctx.build_aabbs_f64(boxes, 1)
    assert got.tobytes() == want.tobytes()


def test_f64_device_input_matches_oracle(ctx):
    from lbvh_amd.hip import DeviceArray
    n = 2_000_003
    print("f64 device build n=%d" % n, flush=True)
    boxes = _scene("uniform", n, 9)
    want, _, _ = oracle.build_f64(boxes, 0)
    d = DeviceArray.from_host(boxes)
    out = ctx.build_aabbs_f64(d, 0)
    ctx.synchronize()
    assert out.numpy()[: n - 1].tobytes() == want.tobytes()
    # a second build into the same output buffer
    boxes2 = _scene("huge", n, 10)
    d.copy_from(boxes2)
    ctx.build_aabbs_f64(d, 1, out=out)
    ctx.synchronize()
    want2, _, _ = oracle.build_f64(boxes2, 1)
    assert out.numpy()[: n - 1].tobytes() == want2.tobytes()
    d.close()
    out.close()


F64_TORCH = r"""
import sys, numpy as np
import torch  # first: the engine then shares torch's HIP runtime
sys.path.insert(0, %r)
import lbvh_amd, oracle
rng = np.random.default_rng(4)
lo = rng.random((300001, 3))
boxes = np.concatenate([lo, lo + rng.random((300001, 3)) * 1e-3], axis=1)
ctx = lbvh_amd.Context(0)
ctx.set_stream(torch.cuda.current_stream())
out = ctx.build_aabbs_f64(torch.from_numpy(boxes).cuda(), 1)
torch.cuda.synchronize()
want, _, _ = oracle.build_f64(boxes, 1)
assert out.cpu().numpy().view(oracle.NODE64_DTYPE)[:300000].tobytes() == want.tobytes()
print("f64 torch ok")
"""


def test_f64_torch_tensor_input():
    """A torch float64 CUDA tensor as device input (fresh process, torch
    imported first: one HIP runtime per process)."""
    import subprocess
    import sys
    from conftest import ROOT
    r = subprocess.run([sys.executable, "-c", F64_TORCH % ROOT], capture_output=True, text=True,
                       timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "f64 torch ok" in r.stdout


def test_f64_alternates_with_fp32_builds(ctx):
    """FP64 and FP32 builds share the context workspace; each stays exact
    whatever ran before it, larger or smaller."""
    from lbvh_amd import workload as W
    tri = W.make_triangles(W.cluster_vertices(5, 120000))
    want32, _, _, _ = oracle.build(tri, 63, 1)
    for n in (200001, 5003, 90001):
        boxes = _scene("clustered", n, n)
        want, _, _ = oracle.build_f64(boxes, 1)
        assert ctx.build_aabbs_f64(boxes, 1).tobytes() == want.tobytes()
        b32 = ctx.build_triangles(tri, code_bits=63, leaf_order=1)
        assert b32.nodes().tobytes() == want32.tobytes()
        b32.destroy()


def test_f64_rejects_bad_arguments(ctx):
    with pytest.raises(lbvh_amd.LbvhError) as e:
        ctx.build_aabbs_f64(np.zeros((1, 6)))
    assert e.value.status == 1
    with pytest.raises(lbvh_amd.LbvhError):
        ctx.build_aabbs_f64(np.zeros((0, 6)))
    with pytest.raises(lbvh_amd.LbvhError):
        ctx.build_aabbs_f64(np.zeros((10, 6)), leaf_order=7)
    # boxes 8 bytes past a 16-byte boundary: refused before any launch
    from lbvh_amd.hip import DeviceArray
    d = DeviceArray.from_host(np.zeros(6 * 100 + 1))
    with pytest.raises(lbvh_amd.LbvhError) as e:
        ctx.build_aabbs_f64(d.rows(1, 601), 0)
    assert e.value.status == 1
    d.close()
    # the context stays usable
    boxes = _scene("uniform", 1000, 3)
    want, _, _ = oracle.build_f64(boxes, 0)
    assert ctx.build_aabbs_f64(boxes, 0).tobytes() == want.tobytes()


# ---- traverser<double> ------------------------------------------------------------

MISS64 = np.uint64(2**64 - 1)


# scale, offset of the 1e6- and 1e-6-scale copies of the uniform scene
SCALED = {"huge": (2.0**20, 3.0e6), "tiny": (2.0**-20, 0.0)}


def _trace_scenes():
    from lbvh_amd import workload as W
    u = W.uniform_vertices(21, 6000, 0.04).astype(np.float64)
    out = {
        "uniform": W.triangles_f64(u),
        "clustered": W.triangles_f64(W.cluster_vertices(22, 6000, half_extent=0.02)),
        # every triangle identical: all codes equal, equal distances everywhere
        "all_equal": W.triangles_f64(np.tile(W.uniform_vertices(23, 1, 0.05), (300, 1))),
    }
    for k, (sc, off) in SCALED.items():
        out[k] = W.triangles_f64(u * sc + off)
    return out


def _rays_f64(tri, seed, m=4000, scene=None):
    """Primary rays of a 48x48 image plus rays from inside the scene's box in
    random directions, some with a tmax that ends before the first hit."""
    from lbvh_amd import workload as W
    sc, off = SCALED.get(scene, (1.0, 0.0))
    boxes = W.triangle_boxes_f64(tri)
    lo, hi = boxes[:, :3].min(axis=0), boxes[:, 3:].max(axis=0)
    inner = W.box_rays(seed, lo.astype(np.float32), hi.astype(np.float32), m).astype(np.float64)
    inner[:, 6] *= sc
    inner[::3, 7] = 0.05 * sc
    prim = W.primary_rays(48, 48).astype(np.float64)
    prim[:, 0:3] = prim[:, 0:3] * sc + off
    prim[:, 6] *= sc
    return np.ascontiguousarray(np.concatenate([prim, inner]))


def W_boxes(tri):
    from lbvh_amd import workload as W
    return W.triangle_boxes_f64(tri)


def _same_f64_hits(got, want, what):
    p, t, uv = got
    wp, wt, wuv = want
    assert np.array_equal(p, wp), "%s: ids differ at %d rays" % (what, int((p != wp).sum()))
    assert np.array_equal(t.view(np.uint64), wt.view(np.uint64)), "%s: t differs" % what
    if uv is not None:
        assert np.array_equal(uv.view(np.uint64), wuv.view(np.uint64)), "%s: uv differs" % what


@pytest.mark.parametrize("scene", ["uniform", "clustered", "all_equal", "huge", "tiny"])
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
@pytest.mark.parametrize("mode", [0, 1], ids=["first", "lowid"])
def test_f64_trace_matches_oracle(ctx, scene, leaf, mode):
    """The GPU's FP64 traversal of the GPU's FP64 tree: ids, t and uv bit for
    bit the oracle's traverser<double> (pinned to the compiled reference's),
    also on 1e6- and 1e-6-scale copies of the uniform scene."""
    tri = _trace_scenes()[scene]
    nodes = ctx.build_aabbs_f64(W_boxes(tri), leaf)
    rays = _rays_f64(tri, 31, scene=scene)
    got = ctx.trace_f64(nodes, tri, rays, mode)
    want = oracle.trace_f64(nodes, tri, rays, mode)
    _same_f64_hits(got, want, "%s leaf %d mode %d" % (scene, leaf, mode))
    assert (got[0] != MISS64).sum() > 100
    assert (got[0] == MISS64).sum() > 100
    assert np.all(got[1][got[0] == MISS64] == np.inf)
    assert np.all(got[2][got[0] == MISS64] == 0.0)


def test_f64_trace_device_full_image(ctx):
    """The 1,048,576-triangle clustered scene (seed 2) widened to FP64: GPU
    build and traversal of the full 1920x1080 primary set with device-
    resident nodes, triangles and rays, against the oracle in both modes."""
    from lbvh_amd import hip
    from lbvh_amd import workload as W
    v, tri32 = W.scene_clustered(2, 1 << 20)
    tri = W.triangles_f64(v)
    rays = W.primary_rays(1920, 1080).astype(np.float64)
    print("f64 trace: build 1M, trace 1920x1080", flush=True)
    d_tri = hip.DeviceArray.from_host(tri)
    d_boxes = hip.DeviceArray.from_host(W.triangle_boxes_f64(tri))
    d_rays = hip.DeviceArray.from_host(rays)
    nodes = ctx.build_aabbs_f64(d_boxes, 1)
    for mode in (0, 1):
        p, t, uv = ctx.trace_f64(nodes, d_tri, d_rays, mode)
        ctx.synchronize()
        got = (p.numpy(), t.numpy(), uv.numpy().reshape(-1, 2))
        host_nodes = nodes.numpy()[: tri.shape[0] - 1]
        print("f64 trace: oracle, mode %d" % mode, flush=True)
        want = oracle.trace_f64(host_nodes, tri, rays, mode)
        _same_f64_hits(got, want, "1M mode %d" % mode)
        assert (got[0] != MISS64).sum() > 100000  # 148,745 of the primary rays hit
        for a in (p, t, uv):
            a.close()
    for a in (d_tri, d_boxes, d_rays, nodes):
        a.close()


def test_f64_trace_without_uv_and_into_given_outputs(ctx):
    from lbvh_amd import hip
    tri = _trace_scenes()["uniform"]
    nodes = ctx.build_aabbs_f64(W_boxes(tri), 1)
    rays = _rays_f64(tri, 32)
    want = oracle.trace_f64(nodes, tri, rays, 0)
    p, t, uv = ctx.trace_f64(nodes, tri, rays, 0, with_uv=False)
    assert uv is None
    _same_f64_hits((p, t, None), want, "host without uv")
    m = rays.shape[0]
    d_nodes = hip.DeviceArray.from_host(nodes)
    d_tri = hip.DeviceArray.from_host(tri)
    d_rays = hip.DeviceArray.from_host(rays)
    out = (hip.DeviceArray(m + 5, np.uint64), hip.DeviceArray(m + 5, np.float64),
           hip.DeviceArray((m + 5, 2), np.float64))
    # twice into the same outputs: the second call overwrites every ray
    for _ in range(2):
        ctx.trace_f64(d_nodes, d_tri, d_rays, 0, out=out)
        ctx.synchronize()
        _same_f64_hits((out[0].numpy()[:m], out[1].numpy()[:m], out[2].numpy()[:m]), want,
                       "device into given outputs")
    for a in out + (d_nodes, d_tri, d_rays):
        a.close()


def test_f64_trace_rejects_bad_arguments(ctx):
    from lbvh_amd import hip
    tri = _trace_scenes()["uniform"][:100]
    nodes = ctx.build_aabbs_f64(W_boxes(tri), 1)
    rays = _rays_f64(tri, 33, 100)
    with pytest.raises(lbvh_amd.LbvhError) as e:
        ctx.trace_f64(nodes, tri, rays, 2)  # no watertight FP64 intersector
    assert e.value.status == 1
    with pytest.raises(ValueError):
        ctx.trace_f64(nodes[:50], tri, rays, 0)
    # rays 8 bytes past a 16-byte boundary: refused before any launch
    d_nodes = hip.DeviceArray.from_host(nodes)
    d_tri = hip.DeviceArray.from_host(tri)
    d_rays = hip.DeviceArray.from_host(np.concatenate([np.zeros(1), rays.reshape(-1)]))
    with pytest.raises(lbvh_amd.LbvhError) as e:
        ctx.trace_f64(d_nodes, d_tri, d_rays.rows(1, 1 + rays.size), 0)
    assert e.value.status == 1
    for a in (d_nodes, d_tri, d_rays):
        a.close()
    # the context stays usable; zero rays is a no-op
    want = oracle.trace_f64(nodes, tri, rays, 1)
    _same_f64_hits(ctx.trace_f64(nodes, tri, rays, 1), want, "after errors")
    p, t, uv = ctx.trace_f64(nodes, tri, rays[:0], 1)
    assert p.shape == (0,)


@pytest.mark.parametrize("n", [2, 3, 5, 64])
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
def test_f64_trace_edge_rays_match_oracle(ctx, n, leaf):
    """Tiny trees and degenerate rays (axis-aligned, in a triangle's plane,
    NaN and infinite components, tmin > tmax, zero direction): the GPU's hits
    equal the oracle's bit for bit, NaN results compared as bits."""
    from lbvh_amd import workload as W
    from test_oracle_f64 import _edge_rays64
    tri = W.triangles_f64(W.uniform_vertices(70 + n, n, 0.2))
    nodes = ctx.build_aabbs_f64(W_boxes(tri), leaf)
    rays = _edge_rays64(tri)
    for mode in (0, 1):
        got = ctx.trace_f64(nodes, tri, rays, mode)
        _same_f64_hits(got, oracle.trace_f64(nodes, tri, rays, mode), "n %d mode %d" % (n, mode))
