"""Fixed workloads to run under rocprofv3 (kernel traces and counters).

  python3 -m lbvh_amd.profile_run single    # 1,048,576-triangle clustered scene,
                                            # 63-bit codes, primitive-id leaves:
                                            # 5 builds, 3 x (primary + bounce) traces
  python3 -m lbvh_amd.profile_run large     # 16,777,216 uniform triangles (seed 3):
                                            # 3 builds each with 30- and 63-bit
                                            # codes, 3 refits after a perturbation
  python3 -m lbvh_amd.profile_run batched   # 1,024 scenes log-uniform [64, 262144]
                                            # (seed 4): 3 rebuilds + 3 batched traces
  python3 -m lbvh_amd.profile_run f64 [n]   # builder<double>: n (default 16,777,216)
                                            # uniform FP64 boxes (seed 1), 5 builds each
                                            # with reference and primitive-id leaves
  python3 -m lbvh_amd.profile_run copy      # calibration: an elementwise kernel that
                                            # reads 512 MiB and writes 512 MiB (torch),
                                            # 5 times

Each prints one line per step so that a long run shows progress.
"""
import sys
import time



def _log(msg):
    print("profile_run: " + msg, flush=True)


def single():
    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W
    _, tri = W.scene_clustered(2, 1 << 20)
    rays = W.primary_rays(1920, 1080)
    ctx = lbvh_amd.Context(0)
    d_tri = hip.DeviceArray.from_host(tri)
    d_rays = hip.DeviceArray.from_host(rays)
    b = ctx.build_triangles(d_tri, code_bits=63, leaf_order=1)
    p, t, _ = ctx.trace(b, d_rays, 0)
    ctx.synchronize()
    bounce, _ = W.bounce_rays(5, tri, rays, p.numpy(), t.numpy())
    d_bounce = hip.DeviceArray.from_host(bounce)
    b.destroy()
    for i in range(5):
        b = ctx.build_triangles(d_tri, code_bits=63, leaf_order=1)
        ctx.synchronize()
        _log("build %d" % i)
        if i < 3:
            ctx.trace(b, d_rays, 0)
            ctx.trace(b, d_bounce, 0)
            ctx.synchronize()
            _log("trace %d (%d primary, %d bounce rays)" % (i, rays.shape[0], bounce.shape[0]))
        b.destroy()
    ctx.close()


def large():
    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W
    v, tri = W.scene_uniform(3, 1 << 24, 5e-4)
    ctx = lbvh_amd.Context(0)
    d_tri = hip.DeviceArray.from_host(tri)
    for bits in (30, 63):
        for i in range(3):
            b = ctx.build_triangles(d_tri, code_bits=bits, leaf_order=0)
            ctx.synchronize()
            _log("build %d-bit %d" % (bits, i))
            if bits == 30 and i == 2:
                tri2 = W.make_triangles(W.perturb(33, v, 1e-4))
                d2 = hip.DeviceArray.from_host(tri2)
                for k in range(3):
                    ctx.refit_triangles(b, d2)
                    ctx.synchronize()
                    _log("refit %d" % k)
            b.destroy()
    ctx.close()


def batched():
    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W
    sizes = W.batched_scene_sizes(4, 1024, 64, 262144)
    tri, off = W.batched_scenes(4, sizes=sizes)
    rays, roff = W.batched_box_rays(4, tri, off, 1024)
    ctx = lbvh_amd.Context(0)
    d_tri = hip.DeviceArray.from_host(tri)
    d_rays = hip.DeviceArray.from_host(rays)
    d_roff = hip.DeviceArray.from_host(roff)
    b = ctx.build_batch(d_tri, off, leaf_order=1)
    ctx.synchronize()
    for i in range(3):
        b.rebuild(graph=True)
        b.trace(d_rays, d_roff, 0)
        ctx.synchronize()
        _log("rebuild + trace %d (%d triangles, %d rays)" % (i, int(off[-1]), rays.shape[0]))
    b.destroy()
    ctx.close()


def f64():
    import numpy as np
    import lbvh_amd
    from lbvh_amd import hip
    n = int(sys.argv[2]) if len(sys.argv) > 2 else 1 << 24
    rng = np.random.default_rng(1)
    lo = rng.random((n, 3))
    boxes = np.concatenate([lo, lo + rng.random((n, 3)) * 5e-4], axis=1)
    del lo
    ctx = lbvh_amd.Context(0)
    d = hip.DeviceArray.from_host(boxes)
    del boxes
    out = hip.DeviceArray((n - 1,), lbvh_amd.NODE64_DTYPE)
    for leaf in (0, 1):
        for i in range(5):
            ctx.build_aabbs_f64(d, leaf, out=out)
            ctx.synchronize()
            _log("f64 build leaf order %d, %d" % (leaf, i))
    ctx.close()


def copy():
    import torch
    a = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for i in range(5):
        torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        _log("copy %d (%d bytes read, %d written)" % (i, a.numel() * 4, b.numel() * 4))


if __name__ == "__main__":
    mode = sys.argv[1] if len(sys.argv) > 1 else "single"
    t0 = time.time()
    {"single": single, "large": large, "batched": batched, "f64": f64, "copy": copy}[mode]()
    _log("%s done in %.1f s" % (mode, time.time() - t0))
