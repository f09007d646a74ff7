"""CPU tests that pin the oracle (oracle/lbvh_oracle.c).

Each function of the restatement is checked against something other than
itself: golden vectors written by the real tay10r/lbvh (tests/golden), the
compiled reference where it is present (oracle/_ref), and what the
mathematics fixes (per-bit interleave, a library stable sort, the radix tree
by definition, subtree min/max, brute-force closest hit, FP64 distances).
"""
import numpy as np
import pytest

import oracle
from lbvh_amd import workload as W
from conftest import golden_cases, nodes_u32

CASES = golden_cases()
IDS = [c[0] for c in CASES]
VARIANTS = [(30, 0), (30, 1), (63, 0), (63, 1)]


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("bits,leaf", VARIANTS, ids=["b30_ref", "b30_pid", "b63_ref", "b63_pid"])
def test_build_matches_golden(case, bits, leaf):
    name, rec = case
    tri = rec["tri"]
    nodes, bounds, sc, ids = oracle.build(tri, bits, leaf)
    key = "b%d_l%d" % (bits, leaf)
    assert np.array_equal(nodes_u32(nodes), rec[key + "_nodes"]), "node array differs"
    assert np.array_equal(sc, rec[key + "_sorted"]), "sorted codes differ"
    assert np.array_equal(bounds.view(np.uint32), rec["bounds"].view(np.uint32))
    assert np.array_equal(oracle.morton_codes(tri, bits), rec["codes%d" % bits])


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
@pytest.mark.parametrize("mode", [0, 1], ids=["first", "lowid"])
def test_trace_matches_golden(case, leaf, mode):
    name, rec = case
    key = "b30_l%d" % leaf
    nodes = rec[key + "_nodes"].view(oracle.NODE_DTYPE).reshape(-1)
    p, t, uv = oracle.trace(nodes, rec["tri"], rec["rays"], mode)
    assert np.array_equal(p, rec["%s_m%d_prim" % (key, mode)])
    assert np.array_equal(t.view(np.uint32), rec["%s_m%d_t" % (key, mode)].view(np.uint32))
    hit = p != 0xFFFFFFFF
    assert np.array_equal(uv[hit].view(np.uint32), rec["%s_m%d_uv" % (key, mode)][hit].view(np.uint32))


# ---- closed forms and library routines ------------------------------------

def _interleave_naive(x, y, z, bits):
    code = 0
    for b in range(bits):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


@pytest.mark.parametrize("bits", [30, 63])
def test_morton_encoding_per_bit(bits):
    """Codes equal a per-bit interleave of trunc((c - min) * D / (ext + 1))
    computed independently in float32 numpy arithmetic."""
    v = W.uniform_vertices(5, 3000, 0.01)
    tri = W.make_triangles(v)
    codes = oracle.morton_codes(tri, bits)
    box = oracle.triangle_aabb(tri)
    c = (box[:, 0:3] + box[:, 3:6]) * np.float32(0.5)
    lo = c.min(axis=0)
    ext = c.max(axis=0) - lo
    dom = np.float32(1024.0 if bits == 30 else 1048576.0)
    scale = dom / (ext + np.float32(1.0))
    q = ((c - lo) * scale).astype(np.float32)
    qi = q.astype(np.uint64)
    lane_bits = 10 if bits == 30 else 20
    for i in range(0, len(tri), 7):
        expect = _interleave_naive(int(qi[i, 0]), int(qi[i, 1]), int(qi[i, 2]), lane_bits)
        assert int(codes[i]) == expect


def test_morton_lane_values_exhaustive():
    """Every 10-bit lane value, and every 20-bit value on each axis, through
    the oracle's encoder: place centres so that q takes that value exactly."""
    # a scene whose centroid extent is 1023 on x: scale = 1024/1024 = 1, q = c - min
    n = 1024
    xs = np.arange(n, dtype=np.float32)
    tri = np.zeros((n, 12), dtype=np.float32)
    tri[:, 0] = xs  # p0 = (x, 0, 0), zero-size triangle: box centre = p0
    codes = oracle.morton_codes(tri, 30)
    for x in range(n):
        assert int(codes[x]) == _interleave_naive(x, 0, 0, 10)
    # 20-bit lanes: extent 2^20 - 1 gives scale = 1 exactly
    m = 1 << 20
    tri = np.zeros((m, 12), dtype=np.float32)
    tri[:, 1] = np.arange(m, dtype=np.float32)  # y axis
    codes = oracle.morton_codes(tri, 63)
    ref = np.zeros(m, dtype=np.uint64)
    yv = np.arange(m, dtype=np.uint64)
    for b in range(20):
        ref |= ((yv >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + 1)
    assert np.array_equal(codes, ref)


@pytest.mark.parametrize("bits", [30, 63])
def test_sort_is_numpy_stable_sort(bits):
    rng = np.random.default_rng(3)
    dt = np.uint32 if bits == 30 else np.uint64
    hi = (1 << 30) if bits == 30 else (1 << 60)
    codes = rng.integers(0, 1 << 12, size=20000).astype(dt) * dt(hi >> 12)  # many duplicates
    sc, ids = oracle.sort_pairs(codes, bits)
    order = np.argsort(codes, kind="stable")
    assert np.array_equal(ids, order.astype(np.uint32))
    assert np.array_equal(sc, codes[order])


def _karras_bruteforce(codes, code_bits):
    """Binary radix tree over index-augmented keys, numbered as Karras
    (internal node of range [lo, hi] is hi if it is a left child, lo if a
    right child or the root)."""
    n = len(codes)
    w = 32 if code_bits == 30 else 64
    keys = [(int(c) << 32) | i for i, c in enumerate(codes)]
    width = w + 32

    def delta(a, b):
        x = keys[a] ^ keys[b]
        return width - x.bit_length()

    nodes = {}

    def rec(lo, hi, index):
        d = delta(lo, hi)
        split = max(s for s in range(lo, hi) if delta(lo, s) > d) if hi - lo > 1 else lo
        left = split | 0x80000000 if split == lo else split
        right = (split + 1) | 0x80000000 if split + 1 == hi else split + 1
        nodes[index] = (left, right)
        if split != lo:
            rec(lo, split, split)
        if split + 1 != hi:
            rec(split + 1, hi, split + 1)

    rec(0, n - 1, 0)
    return nodes


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("seed", range(6))
def test_karras_matches_definition(bits, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 300))
    dt = np.uint32 if bits == 30 else np.uint64
    span = int(rng.choice([4, 64, 1 << 20]))
    codes = np.sort(rng.integers(0, span, size=n).astype(dt))
    if bits == 63:
        codes = codes * dt(1 << 30)
    nodes = oracle.karras(codes, bits)
    bf = _karras_bruteforce(codes, bits)
    assert len(bf) == n - 1
    for i in range(n - 1):
        assert (int(nodes[i]["left"]), int(nodes[i]["right"])) == bf[i], (i, n)


def _leaf_boxes_numpy(tri):
    p0 = tri[:, 0:3]
    p1 = p0 - tri[:, 3:6]
    p2 = p0 + tri[:, 6:9]
    lo = np.minimum(np.minimum(p0, p1), p2)
    hi = np.maximum(np.maximum(p0, p1), p2)
    return lo, hi


def _subtree_boxes(nodes, lo, hi):
    n1 = len(nodes)
    out_lo = np.zeros((n1, 3), np.float32)
    out_hi = np.zeros((n1, 3), np.float32)

    def rec(i):
        bl, bh = [], []
        for ch in (int(nodes[i]["left"]), int(nodes[i]["right"])):
            if ch & 0x80000000:
                bl.append(lo[ch & 0x7FFFFFFF])
                bh.append(hi[ch & 0x7FFFFFFF])
            else:
                rec(ch)
                bl.append(out_lo[ch])
                bh.append(out_hi[ch])
        out_lo[i] = np.minimum(bl[0], bl[1])
        out_hi[i] = np.maximum(bh[0], bh[1])

    import sys
    sys.setrecursionlimit(10000)
    rec(0)
    return out_lo, out_hi


@pytest.mark.parametrize("leaf", [0, 1])
def test_fit_is_subtree_min_max(leaf):
    v = W.uniform_vertices(9, 2000, 0.01)
    tri = W.make_triangles(v)
    nodes, _, _, _ = oracle.build(tri, 30, leaf)
    lo, hi = _leaf_boxes_numpy(tri)
    elo, ehi = _subtree_boxes(nodes, lo, hi)
    assert np.array_equal(nodes["box"][:, 0:3], elo)
    assert np.array_equal(nodes["box"][:, 3:6], ehi)


def check_bvh(nodes):
    """lbvh_test.cpp:369-446: root referenced 0 times, every other internal
    node once, every leaf once; lbvh_test.cpp:458-518: a parent's volume is
    at least each internal child's."""
    m = len(nodes)
    left = nodes["left"].astype(np.int64)
    right = nodes["right"].astype(np.int64)
    internal_refs = np.zeros(m, np.int64)
    leaf_refs = np.zeros(m + 1, np.int64)
    for ch in (left, right):
        is_leaf = (ch & 0x80000000) != 0
        np.add.at(internal_refs, ch[~is_leaf], 1)
        np.add.at(leaf_refs, ch[is_leaf] & 0x7FFFFFFF, 1)
    assert internal_refs[0] == 0
    assert np.all(internal_refs[1:] == 1)
    assert np.all(leaf_refs == 1)
    size = nodes["box"][:, 3:6] - nodes["box"][:, 0:3]
    vol = size[:, 0] * size[:, 1] * size[:, 2]
    for ch in (left, right):
        internal = (ch & 0x80000000) == 0
        assert np.all(vol[internal] >= vol[ch[internal]])


@pytest.mark.parametrize("bits,leaf", VARIANTS)
def test_structural_invariants(bits, leaf):
    v = W.cluster_vertices(4, 5000)
    nodes, _, _, _ = oracle.build(W.make_triangles(v), bits, leaf)
    check_bvh(nodes)


@pytest.mark.parametrize("mode", [1, 2], ids=["lowid", "watertight"])
def test_trace_equals_brute_force(mode):
    v = W.uniform_vertices(1, 20000, 0.01)
    tri = W.make_triangles(v)
    rays = W.primary_rays(48, 48)
    nodes, _, _, _ = oracle.build(tri, 30, 1)
    p, t, _ = oracle.trace(nodes, tri, rays, mode)
    bp, bt = oracle.brute_closest(tri, rays, mode)
    assert np.array_equal(p, bp)
    assert np.array_equal(t.view(np.uint32), bt.view(np.uint32))
    assert np.sum(p != 0xFFFFFFFF) > 100


def test_watertight_close_to_fp64():
    v = W.uniform_vertices(1, 20000, 0.01)
    tri = W.make_triangles(v)
    rays = W.primary_rays(48, 48)
    nodes, _, _, _ = oracle.build(tri, 30, 1)
    for mode in (0, 2):
        p, t, _ = oracle.trace(nodes, tri, rays, mode)
        hit = p != 0xFFFFFFFF
        t64 = oracle.t_fp64(tri, rays, p)
        rel = np.abs(t[hit] - t64[hit]) / np.abs(t64[hit])
        assert rel.max() < 1e-5


def test_watertight_uv_meaning_matches_reference():
    """u and v of both intersectors weight the same vertices (b and c)."""
    v = W.uniform_vertices(1, 20000, 0.01)
    tri = W.make_triangles(v)
    rays = W.primary_rays(48, 48)
    nodes, _, _, _ = oracle.build(tri, 30, 1)
    p1, t1, uv1 = oracle.trace(nodes, tri, rays, 1)
    p2, t2, uv2 = oracle.trace(nodes, tri, rays, 2)
    same = (p1 == p2) & (p1 != 0xFFFFFFFF)
    assert same.sum() > 100
    assert np.abs(uv1[same] - uv2[same]).max() < 1e-3


# ---- the compiled reference, where present ---------------------------------

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference driver not built here")


@needs_ref
@pytest.mark.parametrize("bits,leaf", VARIANTS)
def test_oracle_equals_reference_clustered(bits, leaf):
    v = W.cluster_vertices(8, 30000)
    tri = W.make_triangles(v)
    nodes, bounds, sc, _ = oracle.build(tri, bits, leaf)
    rn, rsc, rb = oracle.ref_build(tri, bits, leaf)
    assert nodes.tobytes() == rn.tobytes()
    assert np.array_equal(sc, rsc)


@needs_ref
@pytest.mark.parametrize("mode", [0, 1])
def test_oracle_trace_equals_reference(mode):
    v = W.cluster_vertices(8, 30000)
    tri = W.make_triangles(v)
    nodes, _, _, _ = oracle.build(tri, 30, 1)
    rays = W.primary_rays(64, 64)
    p, t, uv = oracle.trace(nodes, tri, rays, mode)
    rp, rt, ruv = oracle.ref_trace(nodes, tri, rays, mode)
    assert np.array_equal(p, rp)
    assert np.array_equal(t.view(np.uint32), rt.view(np.uint32))


# ---- full-size workloads through hashes ------------------------------------

def test_uniform65536_hashes(hashes):
    e = hashes["uniform65536_seed1"]
    v, tri = W.scene_uniform(1, 65536, 0.01)
    assert oracle.fnv1a64(tri) == e["tri"], "scene generator drifted"
    rays = W.primary_rays(256, 256)
    assert oracle.fnv1a64(rays) == e["rays"]
    for leaf in (0, 1):
        nodes, bounds, sc, _ = oracle.build(tri, 30, leaf)
        assert oracle.fnv1a64(nodes) == e["b30_l%d_nodes" % leaf]
        assert oracle.fnv1a64(sc) == e["b30_l%d_sorted" % leaf]
    nodes, _, _, _ = oracle.build(tri, 30, 1)
    p, t, _ = oracle.trace(nodes, tri, rays, 0)
    assert oracle.fnv1a64(np.concatenate([p, t.view(np.uint32)])) == e["b30_l1_m0_hits"]


@pytest.mark.slow
def test_uniform16m_hashes(hashes):
    e = hashes["uniform16m_seed3"]
    v, tri = W.scene_uniform(3, 1 << 24, 5e-4)
    del v
    assert oracle.fnv1a64(tri) == e["tri"], "scene generator drifted"
    nodes, bounds, sc, _ = oracle.build(tri, 30, 0)
    assert oracle.fnv1a64(nodes) == e["b30_l0_nodes"]
    assert oracle.fnv1a64(sc) == e["b30_l0_sorted"]


def edge_rays32(tri):
    """FP32 degenerate rays aimed at the scene: axis-aligned (infinite
    reciprocals), in a triangle's plane, NaN and infinite components, an
    empty [tmin, tmax] interval and a zero direction."""
    rng = np.random.default_rng(62)
    v = np.asarray(tri, dtype=np.float64)
    lo = v[:, 0:3].min(axis=0) - 0.1
    hi = v[:, 0:3].max(axis=0) + 0.1
    out = []
    for k in range(3):
        for sgn in (1.0, -1.0):
            for _ in range(8):
                o = rng.uniform(lo, hi)
                o[k] = lo[k] if sgn > 0 else hi[k]
                d = np.zeros(3)
                d[k] = sgn
                out.append(np.concatenate([o, d, [0.0, np.inf]]))
    p0, e1 = v[0, 0:3], v[0, 6:9]
    out.append(np.concatenate([p0 - e1, e1, [0.0, np.inf]]))
    c = (lo + hi) / 2
    out.append(np.concatenate([c, [np.nan, 1.0, 0.0], [0.0, np.inf]]))
    out.append(np.concatenate([[np.nan, 0.5, 0.5], [0.0, 0.0, 1.0], [0.0, np.inf]]))
    out.append(np.concatenate([c - [0, 0, 2], [0.0, 0.0, np.inf], [0.0, np.inf]]))
    out.append(np.concatenate([c - [0, 0, 2], [0.0, 0.0, 1.0], [5.0, 1.0]]))
    out.append(np.concatenate([c, [0.0, 0.0, 0.0], [0.0, np.inf]]))
    return np.ascontiguousarray(np.array(out, dtype=np.float32))


@needs_ref
@pytest.mark.parametrize("n", [2, 3, 5, 64])
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
def test_oracle_trace_edge_rays_equal_reference(n, leaf):
    """Tiny trees and degenerate rays: the oracle's traversal equals the
    compiled reference's bit for bit in both tie modes."""
    tri = W.make_triangles(W.uniform_vertices(80 + n, n, 0.2))
    nodes, _, _, _ = oracle.build(tri, 30, leaf)
    rays = edge_rays32(tri)
    for mode in (0, 1):
        p, t, uv = oracle.trace(nodes, tri, rays, mode)
        rp, rt, ruv = oracle.ref_trace(nodes, tri, rays, mode)
        assert np.array_equal(p, rp)
        assert np.array_equal(t.view(np.uint32), rt.view(np.uint32))
        hit = p != 0xFFFFFFFF
        assert np.array_equal(uv[hit].view(np.uint32), ruv[hit].view(np.uint32))
