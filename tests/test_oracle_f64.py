"""CPU tests that pin the builder<double> restatement of the oracle.

Checked against golden node arrays written by the real lbvh::builder<double>
(tests/golden/f64.npz, over triangle<double> and over plain aabb<double>),
the compiled reference where it is present, and what the mathematics fixes:
numpy min/max for the converter and the centroid bounds, a per-bit
interleave of the quantised centres, the binary radix tree by definition,
and subtree min/max for the fit.
"""
import numpy as np
import pytest

import oracle
from conftest import f64_golden_cases

CASES = f64_golden_cases()
IDS = [c[0] for c in CASES]
BIT63 = np.uint64(1 << 63)


def _boxes(kind, prims):
    return oracle.triangle_aabb_f64(prims) if kind == "triangles" else prims


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
def test_f64_build_matches_golden(case, leaf):
    _, kind, prims, want = case
    nodes, _, _ = oracle.build_f64(_boxes(kind, prims), leaf)
    assert np.array_equal(nodes.view(np.uint64).reshape(-1, 8), want[leaf])


def test_f64_golden_covers_ties_and_both_input_kinds():
    kinds = {c[1] for c in CASES}
    assert kinds == {"triangles", "aabbs"}
    names = {c[0] for c in CASES}
    assert {"triangles:all_equal", "triangles:dup_blocks", "aabbs:points"} <= names


def test_f64_triangle_converter_is_min_max():
    rng = np.random.default_rng(5)
    tri = rng.normal(size=(3000, 12)) * 100.0
    got = oracle.triangle_aabb_f64(tri)
    p0 = tri[:, 0:3]
    p1 = p0 - tri[:, 3:6]
    p2 = p0 + tri[:, 6:9]
    assert np.array_equal(got[:, 0:3], np.minimum(np.minimum(p0, p1), p2))
    assert np.array_equal(got[:, 3:6], np.maximum(np.maximum(p0, p1), p2))


def _interleave_naive(q):
    out = np.zeros(q.shape[0], dtype=np.uint64)
    for b in range(20):
        for axis, sh in ((0, 2), (1, 1), (2, 0)):
            bit = (q[:, axis] >> np.uint64(b)) & np.uint64(1)
            out |= bit << np.uint64(3 * b + sh)
    return out


@pytest.mark.parametrize("scale", [1.0, 1e-7, 3e5])
def test_f64_codes_are_interleaved_quantised_centres(scale):
    rng = np.random.default_rng(8)
    lo = rng.normal(size=(4000, 3)) * scale
    boxes = np.concatenate([lo, lo + rng.random((4000, 3)) * scale], axis=1)
    codes, bounds = oracle.morton_codes_f64(boxes)
    c = (boxes[:, 0:3] + boxes[:, 3:6]) * 0.5
    assert np.array_equal(bounds[0:3], c.min(axis=0))
    assert np.array_equal(bounds[3:6], c.max(axis=0))
    # every lane value lands in [0, 2^20): the top code bits stay clear
    q = np.floor((c - bounds[0:3]) * (1048576.0 / ((bounds[3:6] - bounds[0:3]) + 1.0)))
    assert q.min() >= 0 and q.max() < 1048576
    assert np.array_equal(codes, _interleave_naive(q.astype(np.uint64)))
    assert int(codes.max()) < (1 << 60)


def _radix_tree(codes):
    """Binary radix tree over (code, index) keys, numbered as Karras (the
    internal node of range [lo, hi] is hi as a left child, lo as a right
    child or the root); leaves flagged with bit 63."""
    n = len(codes)
    keys = [(int(c) << 32) | i for i, c in enumerate(codes)]

    def delta(a, b):
        return 96 - (keys[a] ^ keys[b]).bit_length()

    nodes = {}

    def rec(lo, hi, index):
        d = delta(lo, hi)
        split = max(s for s in range(lo, hi) if delta(lo, s) > d) if hi - lo > 1 else lo
        left = split | (1 << 63) if split == lo else split
        right = (split + 1) | (1 << 63) if split + 1 == hi else split + 1
        nodes[index] = (left, right)
        if split != lo:
            rec(lo, split, split)
        if split + 1 != hi:
            rec(split + 1, hi, split + 1)

    rec(0, n - 1, 0)
    return nodes


@pytest.mark.parametrize("seed", range(6))
def test_f64_karras_matches_definition(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 300))
    span = int(rng.choice([4, 64, 1 << 40]))
    codes = np.sort(rng.integers(0, span, size=n).astype(np.uint64)) << np.uint64(rng.integers(0, 20))
    codes = np.sort(codes)
    nodes = oracle.karras_f64(codes)
    want = _radix_tree(codes)
    for i in range(n - 1):
        assert (int(nodes[i]["left"]), int(nodes[i]["right"])) == want[i], (i, n)


def _subtree(nodes, boxes):
    n1 = len(nodes)
    out = np.zeros((n1, 6))

    def rec(i):
        parts = []
        for ch in (int(nodes[i]["left"]), int(nodes[i]["right"])):
            if ch >> 63:
                parts.append(boxes[ch & ((1 << 63) - 1)])
            else:
                rec(ch)
                parts.append(out[ch])
        out[i, 0:3] = np.minimum(parts[0][0:3], parts[1][0:3])
        out[i, 3:6] = np.maximum(parts[0][3:6], parts[1][3:6])

    import sys
    sys.setrecursionlimit(20000)
    rec(0)
    return out


@pytest.mark.parametrize("leaf", [0, 1])
def test_f64_fit_is_subtree_min_max(leaf):
    rng = np.random.default_rng(12)
    lo = rng.random((2500, 3))
    boxes = np.concatenate([lo, lo + rng.random((2500, 3)) * 0.01], axis=1)
    nodes, _, _ = oracle.build_f64(boxes, leaf)
    assert np.array_equal(nodes["box"], _subtree(nodes, boxes))


@pytest.mark.parametrize("leaf", [0, 1])
def test_f64_structure(leaf):
    """Root unreferenced, every other internal node and every leaf referenced
    once (lbvh_test.cpp:369-446); primitive-id leaves are a permutation of
    the positions that sorts the codes stably."""
    rng = np.random.default_rng(13)
    lo = rng.random((3000, 3)) * 1e3
    boxes = np.concatenate([lo, lo + 1.0], axis=1)
    nodes, sc, ids = oracle.build_f64(boxes, leaf)
    m = len(nodes)
    internal = np.zeros(m, np.int64)
    leaves = np.zeros(m + 1, np.int64)
    for ch in (nodes["left"], nodes["right"]):
        is_leaf = (ch & BIT63) != 0
        np.add.at(internal, ch[~is_leaf].astype(np.int64), 1)
        np.add.at(leaves, (ch[is_leaf] & ~BIT63).astype(np.int64), 1)
    assert internal[0] == 0 and np.all(internal[1:] == 1) and np.all(leaves == 1)
    codes, _ = oracle.morton_codes_f64(boxes)
    assert np.array_equal(ids, np.argsort(codes, kind="stable").astype(np.uint32))
    assert np.array_equal(sc, codes[ids])


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not present")
@pytest.mark.parametrize("leaf", [0, 1])
def test_f64_oracle_equals_reference_large(leaf):
    rng = np.random.default_rng(21)
    n = 60000
    c = rng.normal(size=(n, 3)) * rng.choice([0.01, 1.0, 100.0], size=(n, 1))
    v = np.concatenate([c + rng.normal(size=(n, 3)) * 1e-3 for _ in range(3)], axis=1)
    a, b, cc = v[:, 0:3], v[:, 3:6], v[:, 6:9]
    tri = np.concatenate([a, a - b, cc - a, np.zeros((n, 3))], axis=1)
    want = oracle.ref_build_f64(tri, "triangles", leaf)
    got, _, _ = oracle.build_f64(oracle.triangle_aabb_f64(tri), leaf)
    assert got.tobytes() == want.tobytes()


# ---- traverser<double> ------------------------------------------------------------

def tri96_of(v):
    from lbvh_amd import workload as W
    return W.triangles_f64(v)


def _f64_scenes():
    from lbvh_amd import workload as W
    out = {
        "uniform": W.uniform_vertices(21, 6000, 0.01),
        "clustered": W.cluster_vertices(22, 6000, half_extent=0.02),
        # every triangle identical: all codes equal, ties everywhere
        "all_equal": np.tile(W.uniform_vertices(23, 1, 0.05), (300, 1)),
    }
    scenes = {k: tri96_of(v) for k, v in out.items()}
    # 1e6-scale and 1e-6-scale copies of the uniform scene (FP64 vertices
    # scaled exactly by powers of two): the arithmetic must hold at both ends
    u = np.asarray(out["uniform"], dtype=np.float64)
    scenes["huge"] = tri96_of(u * 2.0**20 + 3.0e6)
    scenes["tiny"] = tri96_of(u * 2.0**-20)
    return scenes


def _rays64(w=48, h=48, scene=None):
    from lbvh_amd import workload as W
    r = W.primary_rays(w, h).astype(np.float64)
    if scene == "huge":
        r[:, 0:3] = r[:, 0:3] * 2.0**20 + 3.0e6
        r[:, 6] *= 2.0**20
    elif scene == "tiny":
        r[:, 0:3] *= 2.0**-20
        r[:, 6] *= 2.0**-20
    return r


@pytest.mark.skipif(not oracle.ref_available(), reason="reference driver not built")
@pytest.mark.parametrize("scene", ["uniform", "clustered", "all_equal", "huge", "tiny"])
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
@pytest.mark.parametrize("mode", [0, 1], ids=["first", "lowid"])
def test_f64_trace_matches_compiled_reference(scene, leaf, mode):
    tri = _f64_scenes()[scene]
    nodes, _, _ = oracle.build_f64(oracle.triangle_aabb_f64(tri), leaf)
    rays = _rays64(scene=scene)
    p, t, uv = oracle.trace_f64(nodes, tri, rays, mode)
    rp, rt, ruv = oracle.ref_trace_f64(nodes, tri, rays, mode)
    assert np.array_equal(p, rp)
    assert np.array_equal(t.view(np.uint64), rt.view(np.uint64))
    assert np.array_equal(uv.view(np.uint64), ruv.view(np.uint64))
    assert (p != np.uint64(2**64 - 1)).sum() > 0


@pytest.mark.parametrize("scene", ["uniform", "clustered"])
def test_f64_trace_equals_brute_force(scene):
    """The closest hit does not depend on the tree: lowest-id ties against
    every triangle tested in turn."""
    tri = _f64_scenes()[scene]
    nodes, _, _ = oracle.build_f64(oracle.triangle_aabb_f64(tri), 1)
    rays = _rays64()
    p, t, _ = oracle.trace_f64(nodes, tri, rays, 1)
    bp, bt = oracle.brute_closest_f64(tri, rays)
    assert np.array_equal(p, bp)
    assert np.array_equal(t.view(np.uint64), bt.view(np.uint64))
    assert (p != np.uint64(2**64 - 1)).sum() > 100


@pytest.mark.parametrize("scene", ["uniform", "clustered"])
def test_f64_trace_same_closest_hit_in_both_leaf_orders(scene):
    """With reference leaves a leaf names its sorted position and both the
    fit and the traverser read primitives[position] (lbvh.h:1855, 1904-1905):
    every triangle is still a leaf exactly once under a box fitted from it,
    so the closest hit (lowest id on ties) equals the primitive-id tree's and
    brute force."""
    tri = _f64_scenes()[scene]
    rays = _rays64()
    got = []
    for leaf in (0, 1):
        nodes, _, _ = oracle.build_f64(oracle.triangle_aabb_f64(tri), leaf)
        values = np.concatenate([nodes["left"], nodes["right"]])
        leaves = values[values >= BIT63] - BIT63
        assert np.array_equal(np.sort(leaves), np.arange(len(tri), dtype=np.uint64))
        got.append(oracle.trace_f64(nodes, tri, rays, 1))
    assert np.array_equal(got[0][0], got[1][0])
    assert np.array_equal(got[0][1].view(np.uint64), got[1][1].view(np.uint64))


def _edge_rays64(tri):
    """Axis-aligned rays (zero direction components: infinite reciprocals),
    rays along a triangle's plane, NaN and infinite components, an empty
    [tmin, tmax] interval and a zero direction, aimed at the scene."""
    rng = np.random.default_rng(61)
    lo = tri[:, 0:3].min(axis=0) - 0.1
    hi = tri[:, 0:3].max(axis=0) + 0.1
    out = []
    for k in range(3):
        for sgn in (1.0, -1.0):
            for _ in range(8):
                o = rng.uniform(lo, hi)
                o[k] = lo[k] if sgn > 0 else hi[k]
                d = np.zeros(3)
                d[k] = sgn
                out.append(np.concatenate([o, d, [0.0, np.inf]]))
    p0, e1 = tri[0, 0:3], tri[0, 6:9]
    out.append(np.concatenate([p0 - e1, e1, [0.0, np.inf]]))       # in the plane of a triangle
    c = (lo + hi) / 2
    out.append(np.concatenate([c, [np.nan, 1.0, 0.0], [0.0, np.inf]]))
    out.append(np.concatenate([[np.nan, 0.5, 0.5], [0.0, 0.0, 1.0], [0.0, np.inf]]))
    out.append(np.concatenate([c - [0, 0, 2], [0.0, 0.0, np.inf], [0.0, np.inf]]))
    out.append(np.concatenate([c - [0, 0, 2], [0.0, 0.0, 1.0], [5.0, 1.0]]))  # tmin > tmax
    out.append(np.concatenate([c, [0.0, 0.0, 0.0], [0.0, np.inf]]))          # zero direction
    return np.ascontiguousarray(np.array(out, dtype=np.float64))


@pytest.mark.skipif(not oracle.ref_available(), reason="reference driver not built")
@pytest.mark.parametrize("n", [2, 3, 5, 64])
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
def test_f64_trace_edge_rays_match_compiled_reference(n, leaf):
    """Tiny trees and degenerate rays: the oracle's traverser<double> equals
    the compiled reference's bit for bit (NaN results compared as bits)."""
    from lbvh_amd import workload as W
    tri = tri96_of(W.uniform_vertices(70 + n, n, 0.2))
    nodes, _, _ = oracle.build_f64(oracle.triangle_aabb_f64(tri), leaf)
    rays = _edge_rays64(tri)
    for mode in (0, 1):
        p, t, uv = oracle.trace_f64(nodes, tri, rays, mode)
        rp, rt, ruv = oracle.ref_trace_f64(nodes, tri, rays, mode)
        assert np.array_equal(p, rp)
        assert np.array_equal(t.view(np.uint64), rt.view(np.uint64))
        assert np.array_equal(uv.view(np.uint64), ruv.view(np.uint64))
