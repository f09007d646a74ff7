"""GPU tests of the request server: every request's hits equal the plain
path's result for the same request served alone and in order (bit-exact ids,
t and uv), while requests overlap and share traversal launches; admission
under a memory cap evicts idle scenes; bad requests are rejected."""
import numpy as np
import pytest

import lbvh_amd
from lbvh_amd import workload as W

pytestmark = pytest.mark.gpu


def _scene(seed, n):
    return W.make_triangles(W.cluster_vertices(seed, n))


def _rays(seed, m):
    """Primary rays of a small pinhole camera, then a random subset."""
    r = W.primary_rays(96, 96)
    rng = np.random.default_rng(seed)
    return r[rng.integers(0, r.shape[0], m)]


def _plan(sizes=(3000, 50000, 200000, 1000, 120000, 7000), steps=40, max_rays=9000):
    """(kind, scene index, triangles or None, rays) in arrival order."""
    tris = [_scene(500 + k, n) for k, n in enumerate(sizes)]
    rng = np.random.default_rng(7)
    plan = []
    made = []
    for step in range(steps):
        ray_n = int(rng.integers(1, max_rays))
        rays = _rays(step, ray_n)
        if len(made) < len(sizes) and (step % 4 == 0 or not made):
            k = len(made)
            plan.append(("new", k, tris[k], rays))
            made.append(k)
        elif step % 5 == 3:
            k = made[int(rng.integers(0, len(made)))]
            v = W.perturb(900 + step, W.cluster_vertices(500 + k, sizes[k]), 2e-4)
            plan.append(("refit", k, W.make_triangles(v), rays))
        else:
            k = made[int(rng.integers(0, len(made)))]
            plan.append(("trace", k, None, rays))
    return plan


def _plain_results(ctx, plan, code_bits=63, leaf_order=1, mode=lbvh_amd.ISECT_REFERENCE):
    trees = {}
    out = []
    for kind, k, tri, rays in plan:
        if kind == "new":
            trees[k] = ctx.build_triangles(tri, code_bits=code_bits, leaf_order=leaf_order)
        elif kind == "refit":
            ctx.refit_triangles(trees[k], tri)
        out.append(ctx.trace(trees[k], rays, mode))
    for b in trees.values():
        b.destroy()
    return out


def test_server_results_equal_plain_path_in_order(ctx):
    plan = _plan()
    expect = _plain_results(ctx, plan)
    srv = lbvh_amd.Server(0)
    ids = {}
    tickets = []
    for kind, k, tri, rays in plan:
        if kind == "new":
            t, sid = srv.submit(rays, 0, tri)
            ids[k] = sid
        elif kind == "refit":
            t, _ = srv.submit(rays, ids[k], tri)
        else:
            t, _ = srv.submit(rays, ids[k])
        tickets.append(t)
    for i, t in enumerate(tickets):
        p, tt, uv, info = srv.wait(t)
        ep, et, euv = expect[i]
        assert np.array_equal(p, ep), "request %d (%s) ids" % (i, plan[i][0])
        assert np.array_equal(tt.view(np.uint32), et.view(np.uint32)), "request %d t" % i
        assert np.array_equal(uv.view(np.uint32), euv.view(np.uint32)), "request %d uv" % i
        assert info["done_us"] >= info["ready_us"] >= info["admit_us"] >= info["submit_us"]
        assert info["coalesced_with"] >= 1
    st = srv.stats()
    assert st["completed"] == len(plan) and st["failed"] == 0
    assert st["resident_scenes"] == 6
    assert st["launches"] <= len(plan)
    srv.close()


@pytest.mark.parametrize("code_bits,leaf_order,mode", [(30, 1, 2), (63, 1, 1), (30, 0, 0)],
                         ids=["b30-pid-watertight", "b63-pid-lowid", "b30-ref-reference"])
def test_server_options_equal_plain_path(ctx, code_bits, leaf_order, mode):
    """A server built with other code widths, leaf orders and intersector
    modes gives, request by request, the plain path's hits with the same
    options (reference leaves on unsorted triangles make traversal visit most
    of each tree, so that plan is smaller)."""
    print("server options bits=%d leaf=%d mode=%d" % (code_bits, leaf_order, mode), flush=True)
    if leaf_order == 0:
        plan = _plan(sizes=(3000, 20000, 1000, 9000), steps=16, max_rays=2000)
    else:
        plan = _plan(steps=24)
    expect = _plain_results(ctx, plan, code_bits, leaf_order, mode)
    srv = lbvh_amd.Server(0, code_bits=code_bits, leaf_order=leaf_order, isect_mode=mode)
    ids = {}
    tickets = []
    for kind, k, tri, rays in plan:
        if kind == "new":
            t, sid = srv.submit(rays, 0, tri)
            ids[k] = sid
        elif kind == "refit":
            t, _ = srv.submit(rays, ids[k], tri)
        else:
            t, _ = srv.submit(rays, ids[k])
        tickets.append(t)
    for i, t in enumerate(tickets):
        p, tt, uv, _ = srv.wait(t)
        ep, et, euv = expect[i]
        assert np.array_equal(p, ep), "request %d (%s) ids" % (i, plan[i][0])
        assert np.array_equal(tt.view(np.uint32), et.view(np.uint32)), "request %d t" % i
        assert np.array_equal(uv.view(np.uint32), euv.view(np.uint32)), "request %d uv" % i
    st = srv.stats()
    assert st["completed"] == len(plan) and st["failed"] == 0
    srv.close()


def test_server_coalesces_concurrent_ray_batches(ctx):
    srv = lbvh_amd.Server(0)
    a_tri, b_tri = _scene(11, 40000), _scene(12, 60000)
    t, a = srv.submit(_rays(1, 100), 0, a_tri)
    srv.wait(t)
    t, b = srv.submit(_rays(2, 100), 0, b_tri)
    srv.wait(t)
    ta = ctx.build_triangles(a_tri, code_bits=63, leaf_order=1)
    tb = ctx.build_triangles(b_tri, code_bits=63, leaf_order=1)
    burst = [(a if i % 2 == 0 else b, _rays(100 + i, 2000 + 37 * i)) for i in range(24)]
    tickets = [srv.submit(r, sid)[0] for sid, r in burst]
    for (sid, r), t in zip(burst, tickets):
        p, tt, uv, info = srv.wait(t)
        ep, et, euv = ctx.trace(ta if sid == a else tb, r, lbvh_amd.ISECT_REFERENCE)
        assert np.array_equal(p, ep)
        assert np.array_equal(tt.view(np.uint32), et.view(np.uint32))
    st = srv.stats()
    assert st["max_coalesced"] >= 2, st
    assert st["launches"] < 2 + len(burst)
    ta.destroy()
    tb.destroy()
    srv.close()


def test_server_memory_cap_evicts_idle_scenes(ctx):
    cap = 40 << 20
    srv = lbvh_amd.Server(0, memory_cap_bytes=cap, lanes=2)
    ids = []
    for k in range(10):  # about 5 MB each (triangles + tree): more than the cap holds
        t, sid = srv.submit(_rays(k, 500), 0, _scene(300 + k, 50000))
        srv.wait(t)
        ids.append(sid)
    st = srv.stats()
    assert st["evictions"] > 0
    assert st["memory_peak_bytes"] <= cap
    # the first scene is gone: naming it is an error
    with pytest.raises(lbvh_amd.LbvhError) as e:
        srv.submit(_rays(9, 10), ids[0])
    assert e.value.status == 1
    # the last scene is resident and still answers correctly
    ref = ctx.build_triangles(_scene(309, 50000), code_bits=63, leaf_order=1)
    r = _rays(10, 3000)
    p, tt, _, _ = srv.wait(srv.submit(r, ids[-1])[0])
    ep, et, _ = ctx.trace(ref, r, 0)
    assert np.array_equal(p, ep) and np.array_equal(tt.view(np.uint32), et.view(np.uint32))
    ref.destroy()
    # a request larger than the cap fails cleanly
    t, _ = srv.submit(_rays(11, 10), 0, _scene(99, 400000))
    with pytest.raises(lbvh_amd.LbvhError) as e:
        srv.wait(t)
    assert e.value.status == 2
    srv.close()


def test_server_rejects_bad_requests():
    srv = lbvh_amd.Server(0)
    with pytest.raises(lbvh_amd.LbvhError):
        srv.submit(_rays(1, 10), 0, None)  # a new scene without triangles
    with pytest.raises(lbvh_amd.LbvhError):
        srv.submit(_rays(1, 10), 12345)  # unknown scene
    t, sid = srv.submit(_rays(1, 10), 0, _scene(1, 2000))
    srv.wait(t)
    with pytest.raises(lbvh_amd.LbvhError):
        srv.submit(_rays(1, 10), sid, _scene(1, 1999))  # refit with another count
    # a request with no rays still builds its scene
    t, sid2 = srv.submit(np.zeros((0, 8), np.float32), 0, _scene(2, 3000))
    p, tt, uv, info = srv.wait(t)
    assert p.shape == (0,)
    assert srv.stats()["resident_scenes"] == 2
    srv.close()


def test_server_device_requests_equal_device_plain_path(ctx):
    """Device-resident requests (no staging): same plan, same results as the
    plain device entry points one at a time."""
    from lbvh_amd import hip
    plan = _plan()[:24]
    expect = _plain_results(ctx, plan)
    srv = lbvh_amd.Server(0)
    ids, tickets, keep = {}, [], []
    for kind, k, tri, rays in plan:
        dr = hip.DeviceArray.from_host(rays)
        dt = hip.DeviceArray.from_host(tri) if tri is not None else None
        keep.append((dr, dt))
        if kind == "new":
            t, sid = srv.submit(dr, 0, dt)
            ids[k] = sid
        elif kind == "refit":
            t, _ = srv.submit(dr, ids[k], dt)
        else:
            t, _ = srv.submit(dr, ids[k])
        tickets.append(t)
    for i, t in enumerate(tickets):
        p, tt, uv, info = srv.wait(t)
        ep, et, euv = expect[i]
        assert np.array_equal(p.numpy(), ep), "request %d" % i
        assert np.array_equal(tt.numpy().view(np.uint32), et.view(np.uint32)), "request %d" % i
        assert np.array_equal(uv.numpy().view(np.uint32), euv.view(np.uint32)), "request %d" % i
    srv.close()


def test_server_host_refit_of_device_scene_leaves_borrowed_buffer(ctx):
    """A scene created from device memory borrows that memory and only reads
    it.  A later refit with host data must go to a buffer of the server's
    own, leaving the caller's device triangles untouched, and its results
    must equal the plain path's for the same two requests."""
    from lbvh_amd import hip
    tri0 = _scene(77, 30000)
    tri1 = W.make_triangles(W.perturb(78, W.cluster_vertices(77, 30000), 2e-4))
    r0, r1 = _rays(1, 3000), _rays(2, 3000)
    plain = ctx.build_triangles(tri0, code_bits=63, leaf_order=1)
    e0 = ctx.trace(plain, r0, lbvh_amd.ISECT_REFERENCE)
    ctx.refit_triangles(plain, tri1)
    e1 = ctx.trace(plain, r1, lbvh_amd.ISECT_REFERENCE)
    plain.destroy()
    srv = lbvh_amd.Server(0)
    d_tri0, d_r0 = hip.DeviceArray.from_host(tri0), hip.DeviceArray.from_host(r0)
    t, sid = srv.submit(d_r0, 0, d_tri0)
    p, tt, uv, _ = srv.wait(t)
    assert np.array_equal(p.numpy(), e0[0])
    assert np.array_equal(tt.numpy().view(np.uint32), e0[1].view(np.uint32))
    t, _ = srv.submit(r1, sid, tri1)  # host rays and host triangles: a host refit
    p, tt, uv, _ = srv.wait(t)
    assert np.array_equal(p, e1[0])
    assert np.array_equal(tt.view(np.uint32), e1[1].view(np.uint32))
    assert np.array_equal(uv.view(np.uint32), e1[2].view(np.uint32))
    assert d_tri0.numpy().tobytes() == tri0.tobytes(), "the borrowed device buffer was written"
    srv.close()
