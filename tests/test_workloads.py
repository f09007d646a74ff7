"""CPU tests of the synthetic workload generators that feed the benchmark
and the GPU tests: the batched scenes, their box-surface rays and the serving
request stream are deterministic and have the sizes and fractions they
state."""
import math

import numpy as np

from lbvh_amd import workload as W


def _vertices(tri):
    a = tri[:, 0:3]
    return a, a - tri[:, 3:6], a + tri[:, 6:9]


def test_batched_scene_sizes_are_loguniform_in_range():
    s = W.batched_scene_sizes(4, 4096, 64, 262144)
    assert s.shape == (4096,) and s.min() >= 64 and s.max() <= 262144
    # log-uniform: the median of log(size) sits near the midpoint of the log range
    mid = 0.5 * (math.log(64) + math.log(262144))
    assert abs(np.median(np.log(s.astype(np.float64))) - mid) < 0.2
    # E[n] = (hi - lo) / ln(hi / lo) for a log-uniform size
    expect = (262144 - 64) / math.log(262144 / 64)
    assert abs(s.mean() / expect - 1.0) < 0.05
    assert np.array_equal(s, W.batched_scene_sizes(4, 4096, 64, 262144))


def test_batched_scenes_layout_and_determinism():
    sizes = [2, 3, 100, 1000]
    tri, off = W.batched_scenes(seed=7, sizes=sizes)
    assert off.tolist() == [0, 2, 5, 105, 1105]
    assert tri.shape == (1105, 12) and tri.dtype == np.float32
    # every scene is a uniform scene in its own unit cube: centres in [0,1)^3,
    # vertices within the 0.01 offset
    for s in range(len(sizes)):
        a, b, c = _vertices(tri[off[s]:off[s + 1]])
        cen = (a + b + c) / 3.0
        assert cen.min() > -0.011 and cen.max() < 1.011
    # a scene equals the single-scene generator with its own seed
    v = W.uniform_vertices((7 << 32) | 2, 100, 0.01)
    assert np.array_equal(tri[5:105], W.make_triangles(v))
    tri2, off2 = W.batched_scenes(seed=7, sizes=sizes)
    assert np.array_equal(tri, tri2) and np.array_equal(off, off2)


def test_batched_box_rays_start_on_each_scene_box():
    tri, off = W.batched_scenes(seed=9, sizes=[50, 400, 3000])
    rays, roff = W.batched_box_rays(9, tri, off, per_scene=256)
    assert rays.shape == (768, 8) and roff.tolist() == [0, 256, 512, 768]
    for s in range(3):
        a, b, c = _vertices(tri[off[s]:off[s + 1]])
        lo = np.minimum(np.minimum(a, b), c).min(axis=0)
        hi = np.maximum(np.maximum(a, b), c).max(axis=0)
        r = rays[roff[s]:roff[s + 1]]
        o, d = r[:, 0:3], r[:, 3:6]
        eps = 1e-5
        assert np.all(o >= lo - eps) and np.all(o <= hi + eps)
        # on the surface: at least one coordinate on a face
        on_face = np.isclose(o, lo, atol=eps) | np.isclose(o, hi, atol=eps)
        assert np.all(on_face.any(axis=1))
        assert np.allclose(np.linalg.norm(d.astype(np.float64), axis=1), 1.0, atol=1e-5)
        assert np.all(r[:, 6] == np.float32(1e-4)) and np.all(r[:, 7] == np.float32(1e30))


def test_serving_stream_mix_and_sizes():
    st = W.serving_stream(seed=5, count=4000, rate_hz=1000.0)
    kinds = [r["kind"] for r in st]
    assert kinds[0] == "new"
    n = len(st)
    refit = kinds.count("refit") / n
    trace = kinds.count("trace") / n
    assert abs(refit - 0.10) < 0.02 and abs(trace - 0.30) < 0.03
    t = np.array([r["arrival_s"] for r in st])
    assert np.all(np.diff(t) > 0)
    # Poisson arrivals: mean gap 1 / rate
    assert abs(np.diff(t).mean() * 1000.0 - 1.0) < 0.05
    new = [r for r in st if r["kind"] == "new"]
    assert min(r["n_tri"] for r in new) >= 1024 and max(r["n_tri"] for r in new) <= 4194304
    assert all(4096 <= r["n_rays"] <= 2097152 for r in st)
    # requests on a scene come after its creation and carry its size
    made = {}
    for r in st:
        if r["kind"] == "new":
            assert r["scene"] == len(made)
            made[r["scene"]] = r["n_tri"]
        else:
            assert r["scene"] in made and r["n_tri"] == made[r["scene"]]
    assert st == W.serving_stream(seed=5, count=4000, rate_hz=1000.0)


def test_bench_traversal_touched_bytes_constant():
    """The traversal roofline's byte count: bench.py states the bytes the
    closest-hit traversal touches on its workload (popped nodes x 32 +
    box-tested children x 32 + tested triangles x 48, an upper-bound proxy
    for its memory traffic); the instrumented CPU traversal of the same tree
    must count the same.  The tree is the oracle's, which the GPU build equals bit for bit."""
    import importlib.util
    import os
    import oracle
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    _, tri = W.scene_clustered(2, 1 << 20)
    nodes, _, _, _ = oracle.build(tri, 63, oracle.LEAF_PRIMITIVE_ID)
    rays = W.primary_rays(1920, 1080)
    p, t, _, c1 = oracle.trace(nodes, tri, rays, oracle.ISECT_REFERENCE, counters=True)
    bounce, _ = W.bounce_rays(5, tri, rays, p, t)
    _, _, _, c2 = oracle.trace(nodes, tri, bounce, oracle.ISECT_REFERENCE, counters=True)
    touched = sum(int(c[0]) * 32 + int(c[1]) * 32 + int(c[2]) * 48 for c in (c1, c2))
    assert touched == bench.TRACE_TOUCHED_BYTES_PER_STEP


def _load_bench():
    import importlib.util
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    return bench


def test_bench_fp64_distance_error_matches_oracle():
    """bench.py reports the hit distance's error against FP64 with its own
    numpy evaluation of the plane distance; it must agree with the oracle's
    FP64 evaluation of the same formula, and a planted error of k units of
    the FP32 spacing must read back as k ulp."""
    import oracle
    bench = _load_bench()
    _, tri = W.scene_uniform(1, 65536, 0.01)
    nodes, _, _, _ = oracle.build(tri, 30, oracle.LEAF_PRIMITIVE_ID)
    rays = W.primary_rays(128, 128)
    p, t, _ = oracle.trace(nodes, tri, rays, oracle.ISECT_REFERENCE)
    hit = p != 0xFFFFFFFF
    assert hit.sum() > 1000
    t64 = oracle.t_fp64(tri, rays, p)
    rel = np.abs(t[hit] - t64[hit]) / np.abs(t64[hit])
    e = bench.t_error_vs_fp64(tri, rays, p, t)
    assert e["hits"] == int(hit.sum())
    assert math.isclose(e["max_rel"], float("%.4g" % rel.max()), rel_tol=1e-3)
    assert math.isclose(e["mean_rel"], float("%.4g" % rel.mean()), rel_tol=1e-3)
    # plant: the FP32 value rounded from t64, moved by 3 spacings
    t_exact = t64.astype(np.float32)
    moved = t_exact.copy()
    for _ in range(3):
        moved = np.nextafter(moved, np.float32(np.inf))
    e3 = bench.t_error_vs_fp64(tri, rays, p, np.where(hit, moved, t))
    e0 = bench.t_error_vs_fp64(tri, rays, p, np.where(hit, t_exact, t))
    assert e0["max_ulp"] <= 0.5 + 1e-6
    assert 2.5 <= e3["max_ulp"] <= 3.5 + 1e-6 and e3["mean_ulp"] >= 2.5


def _py_traverse(boxes, left, right, tris, ray):
    """traverser::operator() (lbvh.h:1874-1942) with triangle_intersector
    (lbvh.h:2055-2075) and the octant box test (lbvh.h:991-1010, 1068-1094),
    restated from the reference in Python with float32 arithmetic (every
    operation rounds to float32, none fused).  Returns (prim, t, popped,
    box tests, triangle tests): popped counts the entries not pruned."""
    f = np.float32
    o = [f(x) for x in ray[0:3]]
    d = [f(x) for x in ray[3:6]]
    rtmin, rtmax = f(ray[6]), f(ray[7])
    rcp = [f(1) / d[k] for k in range(3)]
    ipos = [(-o[k]) * rcp[k] for k in range(3)]
    octs = [0 if d[0] > 0 else 3, 1 if d[1] > 0 else 4, 2 if d[2] > 0 else 5]
    inv = [3 - octs[0], 5 - octs[1], 7 - octs[2]]
    mx = lambda a, b: a if a > b else b
    mn = lambda a, b: a if a < b else b

    def dot(a, b):
        out = f(0)
        for k in range(3):
            out = out + a[k] * b[k]
        return out

    def isect(i):
        p0, e0, e1, nrm = tris[i, 0:3], tris[i, 3:6], tris[i, 6:9], tris[i, 9:12]
        c = [p0[k] - o[k] for k in range(3)]
        r = [d[1] * c[2] - d[2] * c[1], d[2] * c[0] - d[0] * c[2], d[0] * c[1] - d[1] * c[0]]
        inv_det = f(1) / dot(nrm, d)
        u = dot(r, e1) * inv_det
        v = dot(r, e0) * inv_det
        w = f(1) - u - v
        if u >= 0 and v >= 0 and w >= 0:
            t = dot(nrm, c) * inv_det
            if rtmin <= t <= rtmax:
                return t
        return None

    def box(i):
        b = boxes[i]
        tn = [b[octs[k]] * rcp[k] + ipos[k] for k in range(3)]
        tf = [b[inv[k]] * rcp[k] + ipos[k] for k in range(3)]
        tmin = mx(mx(tn[0], tn[1]), tn[2])
        tmax = mn(mn(tf[0], tf[1]), tf[2])
        return tmax >= mx(f(0), tmin), tmin

    leaf = 0x80000000
    stack = [(0, f(np.inf))]
    best, best_t = 0xFFFFFFFF, f(np.inf)
    popped = nbox = ntri = 0
    while stack:
        ni, etmin = stack.pop()
        if best_t < etmin:
            continue
        popped += 1
        hits = []
        for ch in (int(left[ni]), int(right[ni])):
            if ch & leaf:
                ntri += 1
                t = isect(ch & ~leaf)
                if t is not None and t < best_t:
                    best, best_t = ch & ~leaf, t
                hits.append((False, None))
            else:
                nbox += 1
                hits.append(box(ch))
        (lh, lt), (rh, rt) = hits
        entries = []
        if lh and rh:
            entries = [(int(right[ni]), rt), (int(left[ni]), lt)] if lt < rt else \
                      [(int(left[ni]), lt), (int(right[ni]), rt)]
        elif lh:
            entries = [(int(left[ni]), lt)]
        elif rh:
            entries = [(int(right[ni]), rt)]
        for e in entries:
            if len(stack) < 128:
                stack.append(e)
    return best, best_t, popped, nbox, ntri


def test_touched_byte_counts_match_an_independent_traversal():
    """The traversal roofline's byte count rests on the oracle's per-ray
    counters (popped nodes, box tests, triangle tests); a second restatement
    of the reference traverser, written in Python, counts the same for every
    ray of a sample of primary and bounce rays on the 1M benchmark tree, and
    finds the same hits."""
    import oracle
    _, tri = W.scene_clustered(2, 1 << 20)
    nodes, _, _, _ = oracle.build(tri, 63, oracle.LEAF_PRIMITIVE_ID)
    rays = W.primary_rays(1920, 1080)
    p, t, _ = oracle.trace(nodes, tri, rays, oracle.ISECT_REFERENCE)
    bounce, _ = W.bounce_rays(5, tri, rays, p, t)
    rng = np.random.default_rng(77)
    hit = np.flatnonzero(p != 0xFFFFFFFF)
    sample = np.concatenate([rays[rng.choice(rays.shape[0], 16, replace=False)],
                             rays[rng.choice(hit, 16, replace=False)],
                             bounce[rng.choice(bounce.shape[0], 32, replace=False)]])
    boxes = nodes["box"]
    left, right = nodes["left"], nodes["right"]
    total = 0
    for ray in sample:
        op, ot, _, c = oracle.trace(nodes, tri, ray[None, :], oracle.ISECT_REFERENCE, threads=1,
                                    counters=True)
        bp, bt, popped, nbox, ntri = _py_traverse(boxes, left, right, tri, ray)
        assert (int(op[0]), float(ot[0])) == (bp, float(bt))
        assert (int(c[0]), int(c[1]), int(c[2])) == (popped, nbox, ntri)
        total += popped + nbox + ntri
    assert total > 64 * 50
