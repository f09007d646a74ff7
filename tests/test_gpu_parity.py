"""GPU parity: the HIP engine, called through its C ABI, against the golden
vectors written by the real tay10r/lbvh and against the oracle.

Bar: bit-exact Morton codes, sorted order, topology and boxes; identical
hit/miss, primitive ids, t bits and uv bits for the reference intersector.
The watertight intersector is compared with the oracle's restatement of the
same test with a tolerance of 2 ulp on t (ids identical)."""
import numpy as np
import pytest

import lbvh_amd
import oracle
from lbvh_amd import workload as W
from conftest import golden_cases, nodes_u32

pytestmark = pytest.mark.gpu

CASES = golden_cases()
IDS = [c[0] for c in CASES]
VARIANTS = [(30, 0), (30, 1), (63, 0), (63, 1)]
VIDS = ["b30_ref", "b30_pid", "b63_ref", "b63_pid"]


def ulp_diff(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("bits,leaf", VARIANTS, ids=VIDS)
def test_build_matches_golden(ctx, case, bits, leaf):
    name, rec = case
    tri = rec["tri"]
    bvh = ctx.build_triangles(tri, code_bits=bits, leaf_order=leaf, keep_codes=True)
    key = "b%d_l%d" % (bits, leaf)
    nodes = bvh.nodes()
    assert np.array_equal(nodes_u32(nodes), rec[key + "_nodes"]), "node array differs"
    sc, ids, bounds = bvh.codes()
    assert np.array_equal(sc, rec[key + "_sorted"]), "sorted codes differ"
    assert np.array_equal(bounds.view(np.uint32), rec["bounds"].view(np.uint32))
    # stable (code, id) order
    order = np.argsort(rec["codes%d" % bits], kind="stable").astype(np.uint32)
    assert np.array_equal(ids, order)
    bvh.destroy()


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
@pytest.mark.parametrize("mode", [0, 1], ids=["first", "lowid"])
def test_trace_matches_golden(ctx, case, leaf, mode):
    name, rec = case
    bvh = ctx.build_triangles(rec["tri"], code_bits=30, leaf_order=leaf)
    p, t, uv = ctx.trace(bvh, rec["rays"], mode)
    key = "b30_l%d_m%d" % (leaf, mode)
    assert np.array_equal(p, rec[key + "_prim"])
    assert np.array_equal(t.view(np.uint32), rec[key + "_t"].view(np.uint32))
    hit = p != 0xFFFFFFFF
    assert np.array_equal(uv[hit].view(np.uint32), rec[key + "_uv"][hit].view(np.uint32))
    assert np.all(uv[~hit] == 0)
    bvh.destroy()


@pytest.mark.parametrize("nrays", [4096, 600000], ids=["rays4096", "rays600000"])
@pytest.mark.parametrize("mode", [0, 1], ids=["first", "lowid"])
def test_trace_deep_stacks_match_oracle(ctx, nrays, mode):
    """20,000 large, mutually overlapping triangles: most rays hit both
    children at many levels, so traversal stacks run far past the entries
    kept in registers and LDS into the private array (hundreds of pops per
    ray).  A dispatch that leaves most of the device idle and one that fills
    it, against the oracle, bit for bit."""
    print("deep-stack trace nrays=%d mode=%d" % (nrays, mode), flush=True)
    rng = np.random.default_rng(3)
    n = 20000
    c = rng.random((n, 3)) * 0.2 + 0.4
    v = np.concatenate([c + rng.normal(size=(n, 3)) * 0.3 for _ in range(3)], axis=1)
    tri = W.make_triangles(v.astype(np.float32))
    rays = W.box_rays(5, np.full(3, -0.5, np.float32), np.full(3, 1.5, np.float32), nrays)
    bvh = ctx.build_triangles(tri, code_bits=63, leaf_order=1)
    nodes, _, _, _ = oracle.build(tri, 63, 1)
    assert bvh.nodes().tobytes() == nodes.tobytes()
    p, t, uv = ctx.trace(bvh, rays, mode)
    op, ot, ouv = oracle.trace(nodes, tri, rays, mode)
    assert np.array_equal(p, op)
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    hit = p != 0xFFFFFFFF
    assert np.array_equal(uv[hit].view(np.uint32), ouv[hit].view(np.uint32))
    bvh.destroy()


@pytest.mark.parametrize("n", [2, 3, 5, 33, 4097, 70001])
@pytest.mark.parametrize("bits,leaf", VARIANTS, ids=VIDS)
def test_build_matches_oracle_ragged(ctx, n, bits, leaf):
    v = W.cluster_vertices(1000 + n, n)
    tri = W.make_triangles(v)
    bvh = ctx.build_triangles(tri, code_bits=bits, leaf_order=leaf, keep_codes=True)
    on, ob, osc, oids = oracle.build(tri, bits, leaf)
    assert bvh.nodes().tobytes() == on.tobytes()
    sc, ids, bounds = bvh.codes()
    assert np.array_equal(sc, osc)
    assert np.array_equal(ids, oids)


@pytest.mark.parametrize("slots", ["1", "200"])
@pytest.mark.parametrize("n", [9000, 70001, 300001])
@pytest.mark.parametrize("bits,leaf", [(30, 0), (63, 1)], ids=["b30_ref", "b63_pid"])
def test_split_tree_merges_match_oracle(ctx, monkeypatch, slots, n, bits, leaf):
    """The upper tree levels merge a group's lists 8 at a time (then the
    partial lists) when they hold more elements than a workgroup's slots.
    LBVH_TREE_GROUP_SLOTS lowers that limit so ordinary scenes take the
    split path (512-position level-0 workgroups, 64 per group): 18 level-0
    workgroups in one level-1 group (n = 9000), a partial last group
    (70001), ten level-1 workgroups under one top group (300001)."""
    monkeypatch.setenv("LBVH_TREE_GROUP_SLOTS", slots)
    v = W.cluster_vertices(2000 + n, n)
    tri = W.make_triangles(v)
    bvh = ctx.build_triangles(tri, code_bits=bits, leaf_order=leaf, keep_codes=True)
    on, ob, osc, oids = oracle.build(tri, bits, leaf)
    assert bvh.nodes().tobytes() == on.tobytes()
    sc, ids, bounds = bvh.codes()
    assert np.array_equal(sc, osc)
    assert np.array_equal(ids, oids)
    # the refit starts from the flags and links the tree kernels leave
    moved = W.make_triangles(W.perturb(7, v, 1e-4))
    ctx.refit_triangles(bvh, moved)
    assert bvh.nodes().tobytes() == oracle.fit(on, moved).tobytes()


SORT_CONFIGS = [(1, 8, 1024), (1, 8, 2048), (1, 8, 4096), (1, 10, 2048), (1, 10, 4096),
                (2, 8, 2048), (2, 8, 4096), (2, 10, 2048), (2, 10, 4096),
                (3, 8, 2048), (3, 8, 4096), (3, 10, 4096)]
SORT_NAMES = {1: "onesweep", 2: "rts", 3: "hybrid"}


@pytest.mark.parametrize("strategy,digit_bits,tile", SORT_CONFIGS,
                         ids=["%s-d%d-t%d" % (SORT_NAMES[c[0]], c[1], c[2])
                              for c in SORT_CONFIGS])
@pytest.mark.parametrize("bits", [30, 63])
def test_sort_configurations_agree(ctx, strategy, digit_bits, tile, bits):
    """Every sort strategy at every digit width and tile size gives the
    oracle's stable (code, id) order and tree."""
    v = W.cluster_vertices(77, 300000)
    tri = W.make_triangles(v)
    bvh = ctx.build_triangles(tri, code_bits=bits, leaf_order=1, keep_codes=True,
                              sort_digit_bits=digit_bits, sort_tile=tile, sort_strategy=strategy)
    on, _, osc, oids = oracle.build(tri, bits, 1)
    sc, ids, _ = bvh.codes()
    assert np.array_equal(sc, osc)
    assert np.array_equal(ids, oids)
    assert bvh.nodes().tobytes() == on.tobytes()


def _hybrid_edge_scene(kind, n, seed):
    """Key distributions that stress the hybrid sort's buckets."""
    rng = np.random.default_rng(seed)
    if kind == "one-cluster-and-outlier":
        # every code but one shares its top bits: one bucket of n - 1 keys,
        # sorted by the oversize-range path
        v = rng.random((n, 3, 3), dtype=np.float32) * np.float32(1e-3)
        v[-1] += np.float32(1e3)
    elif kind == "duplicates":
        # a handful of distinct triangles repeated: long runs of equal codes
        base = rng.random((7, 3, 3), dtype=np.float32)
        v = base[rng.integers(0, 7, n)]
    elif kind == "two-clusters":
        v = rng.random((n, 3, 3), dtype=np.float32) * np.float32(1e-2)
        v[n // 3:] += np.float32(5.0)
    else:  # "shells": clustered, buckets of every size
        v = W.cluster_vertices(seed, n)
    return W.make_triangles(v.reshape(n, 9))


@pytest.mark.parametrize("kind", ["one-cluster-and-outlier", "duplicates", "two-clusters",
                                  "shells"])
@pytest.mark.parametrize("n", [2, 3, 2049, 4097, 9000, 250001])
@pytest.mark.parametrize("bits", [30, 63])
def test_hybrid_sort_edge_distributions(ctx, kind, n, bits):
    """The hybrid sort on skewed keys (one bucket holding nearly every key,
    runs of equal codes, ragged sizes) gives the oracle's order and tree."""
    tri = _hybrid_edge_scene(kind, n, 600 + n)
    bvh = ctx.build_triangles(tri, code_bits=bits, leaf_order=1, keep_codes=True,
                              sort_strategy=lbvh_amd.SORT_HYBRID)
    assert bvh.info().sort_strategy == 3
    on, _, osc, oids = oracle.build(tri, bits, 1)
    sc, ids, _ = bvh.codes()
    assert np.array_equal(sc, osc)
    assert np.array_equal(ids, oids)
    assert bvh.nodes().tobytes() == on.tobytes()
    bvh.destroy()


def test_rejects_bad_arguments(ctx):
    tri = W.make_triangles(W.uniform_vertices(1, 10, 0.01))
    with pytest.raises(lbvh_amd.LbvhError) as e:
        ctx.build_triangles(tri[:1])
    assert e.value.status == 1
    with pytest.raises(lbvh_amd.LbvhError):
        ctx.build_triangles(tri, code_bits=31)
    with pytest.raises(lbvh_amd.LbvhError):
        ctx.build_triangles(tri, sort_digit_bits=7)


def test_uniform65536_hashes(ctx, hashes):
    e = hashes["uniform65536_seed1"]
    v, tri = W.scene_uniform(1, 65536, 0.01)
    assert oracle.fnv1a64(tri) == e["tri"]
    rays = W.primary_rays(256, 256)
    for leaf in (0, 1):
        bvh = ctx.build_triangles(tri, code_bits=30, leaf_order=leaf, keep_codes=True)
        assert oracle.fnv1a64(bvh.nodes()) == e["b30_l%d_nodes" % leaf]
        sc, _, _ = bvh.codes()
        assert oracle.fnv1a64(sc) == e["b30_l%d_sorted" % leaf]
        for mode in (0, 1):
            p, t, _ = ctx.trace(bvh, rays, mode)
            assert oracle.fnv1a64(np.concatenate([p, t.view(np.uint32)])) == \
                e["b30_l%d_m%d_hits" % (leaf, mode)]
        bvh.destroy()


def test_clustered1m_hashes(ctx, hashes):
    e = hashes["clustered1m_seed2"]
    v, tri = W.scene_clustered(2, 1 << 20)
    if oracle.fnv1a64(tri) != e["tri"]:
        pytest.skip("libm differs from the fixture host: compared in-run only")
    for bits in (30, 63):
        for leaf in (0, 1):
            bvh = ctx.build_triangles(tri, code_bits=bits, leaf_order=leaf, keep_codes=True)
            assert oracle.fnv1a64(bvh.nodes()) == e["b%d_l%d_nodes" % (bits, leaf)]
            sc, _, _ = bvh.codes()
            assert oracle.fnv1a64(sc) == e["b%d_l%d_sorted" % (bits, leaf)]
            bvh.destroy()


def test_uniform16m_hashes(ctx, hashes):
    e = hashes["uniform16m_seed3"]
    v, tri = W.scene_uniform(3, 1 << 24, 5e-4)
    del v
    assert oracle.fnv1a64(tri) == e["tri"]
    for bits in (30, 63):
        for leaf in (0, 1):
            bvh = ctx.build_triangles(tri, code_bits=bits, leaf_order=leaf, keep_codes=True)
            assert oracle.fnv1a64(bvh.nodes()) == e["b%d_l%d_nodes" % (bits, leaf)]
            sc, _, _ = bvh.codes()
            assert oracle.fnv1a64(sc) == e["b%d_l%d_sorted" % (bits, leaf)]
            bvh.destroy()


def test_clustered1m_primary_and_bounce_rays(ctx):
    """The 1,048,576-triangle clustered scene with 63-bit codes: primary rays
    and one diffuse bounce per hit, against the oracle traversal."""
    v, tri = W.scene_clustered(2, 1 << 20)
    bvh = ctx.build_triangles(tri, code_bits=63, leaf_order=1)
    nodes = bvh.nodes()
    rays = W.primary_rays(480, 270)
    for mode in (0, 1):
        p, t, uv = ctx.trace(bvh, rays, mode)
        op, ot, ouv = oracle.trace(nodes, tri, rays, mode)
        assert np.array_equal(p, op)
        assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
        assert np.array_equal(uv.view(np.uint32), ouv.view(np.uint32))
    b, src = W.bounce_rays(5, tri, rays, p, t)
    assert len(b) > 1000
    for mode in (0, 1):
        p2, t2, _ = ctx.trace(bvh, b, mode)
        op2, ot2, _ = oracle.trace(nodes, tri, b, mode)
        assert np.array_equal(p2, op2)
        assert np.array_equal(t2.view(np.uint32), ot2.view(np.uint32))


def test_watertight_matches_oracle(ctx):
    v, tri = W.scene_uniform(1, 65536, 0.01)
    bvh = ctx.build_triangles(tri, code_bits=30, leaf_order=1)
    nodes = bvh.nodes()
    rays = W.primary_rays(256, 256)
    p, t, uv = ctx.trace(bvh, rays, 2)
    op, ot, ouv = oracle.trace(nodes, tri, rays, 2)
    assert np.array_equal(p, op)
    hit = p != 0xFFFFFFFF
    assert ulp_diff(t[hit], ot[hit]).max() <= 2  # tolerance: 2 ulp on t
    t64 = oracle.t_fp64(tri, rays, p)
    rel = np.abs(t[hit] - t64[hit]) / np.abs(t64[hit])
    print("watertight t: max rel err vs FP64 %.3e over %d hits" % (rel.max(), hit.sum()))
    assert rel.max() < 1e-5


@pytest.mark.parametrize("leaf", [0, 1])
def test_refit_equals_oracle_fit(ctx, leaf):
    v, tri = W.scene_uniform(3, 200000, 5e-4)
    bvh = ctx.build_triangles(tri, code_bits=30, leaf_order=leaf)
    before = bvh.nodes()
    v2 = W.perturb(33, v, 1e-4)
    tri2 = W.make_triangles(v2)
    ctx.refit_triangles(bvh, tri2)
    after = bvh.nodes()
    assert np.array_equal(after["left"], before["left"])
    assert np.array_equal(after["right"], before["right"])
    expect = oracle.fit(before, tri2)
    assert after.tobytes() == expect.tobytes()
    # a second refit back to the original data restores the original boxes
    ctx.refit_triangles(bvh, tri)
    assert bvh.nodes().tobytes() == before.tobytes()


def test_refits_alternating_between_trees(ctx):
    """A context keeps the treelet list of the tree it last refitted: refits
    of A, B, A again and B again each equal the oracle's fit."""
    va, ta = W.scene_uniform(5, 150000, 5e-4)
    vb, tb = W.scene_clustered(6, 90000)
    a = ctx.build_triangles(ta, code_bits=63, leaf_order=1)
    b = ctx.build_triangles(tb, code_bits=30, leaf_order=0)
    na, nb = a.nodes(), b.nodes()
    for k in range(2):
        ma = W.make_triangles(W.perturb(40 + k, va, 1e-4))
        ctx.refit_triangles(a, ma)
        assert a.nodes().tobytes() == oracle.fit(na, ma).tobytes(), "A, round %d" % k
        mb = W.make_triangles(W.perturb(50 + k, vb.reshape(-1, 9), 1e-4))
        ctx.refit_triangles(b, mb)
        assert b.nodes().tobytes() == oracle.fit(nb, mb).tobytes(), "B, round %d" % k
    a.destroy()
    b.destroy()


@pytest.mark.parametrize("n", [2, 3, 5, 1023, 1024, 1025, 2049, 5000])
@pytest.mark.parametrize("bits,leaf", [(30, 0), (63, 1)], ids=["b30_ref", "b63_pid"])
def test_refit_small_and_treelet_edge_sizes(ctx, n, bits, leaf):
    """Refits of trees that are one treelet (n <= 1,024: the root is node 0)
    or a few treelets under a short top tree equal the oracle's fit."""
    v = W.cluster_vertices(3000 + n, n)
    bvh = ctx.build_triangles(W.make_triangles(v), code_bits=bits, leaf_order=leaf)
    before = bvh.nodes()
    moved = W.make_triangles(W.perturb(11, v, 1e-4))
    ctx.refit_triangles(bvh, moved)
    assert bvh.nodes().tobytes() == oracle.fit(before, moved).tobytes()
    bvh.destroy()


@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
def test_refit_with_nan_vertices_equals_oracle_fit(ctx, leaf):
    """A refit whose new vertices hold NaNs (the topology is kept, so no
    code is computed from them): treelets with a NaN box are fitted in tree
    order, and every box equals the oracle's fit bit for bit."""
    v, tri = W.scene_uniform(8, 120000, 5e-4)
    bvh = ctx.build_triangles(tri, code_bits=30, leaf_order=leaf)
    before = bvh.nodes()
    v2 = W.perturb(9, v, 1e-4)
    rng = np.random.default_rng(10)
    bad = rng.choice(v2.shape[0], 40, replace=False)
    v2[bad, rng.integers(0, 9, 40)] = np.float32(np.nan)
    tri2 = W.make_triangles(v2)
    ctx.refit_triangles(bvh, tri2)
    assert bvh.nodes().tobytes() == oracle.fit(before, tri2).tobytes()


def test_device_path_equals_host_path(ctx):
    from lbvh_amd import hip
    v, tri = W.scene_uniform(1, 65536, 0.01)
    rays = W.primary_rays(128, 128)
    host = ctx.build_triangles(tri, code_bits=63, leaf_order=1)
    hp, ht, huv = ctx.trace(host, rays, 1)
    dtri = hip.DeviceArray.from_host(tri)
    drays = hip.DeviceArray.from_host(rays)
    dev = ctx.build_triangles(dtri, code_bits=63, leaf_order=1)
    dp, dt, duv = ctx.trace(dev, drays, 1)
    ctx.synchronize()
    assert dev.nodes().tobytes() == host.nodes().tobytes()
    assert np.array_equal(dp.numpy(), hp)
    assert np.array_equal(dt.numpy().view(np.uint32), ht.view(np.uint32))
    assert np.array_equal(duv.numpy().view(np.uint32), huv.view(np.uint32))
    # device refit from a device buffer, then trace again
    v2 = W.perturb(2, v, 1e-4)
    tri2 = W.make_triangles(v2)
    ctx.refit_triangles(dev, hip.DeviceArray.from_host(tri2))
    ctx.synchronize()
    assert dev.nodes().tobytes() == oracle.fit(host.nodes(), tri2).tobytes()


TORCH_INTEROP = r"""
import sys, numpy as np
import torch  # first: the engine then shares torch's HIP runtime
sys.path.insert(0, %r)
import lbvh_amd, oracle
from lbvh_amd import workload as W
v, tri = W.scene_uniform(1, 20000, 0.01)
rays = W.primary_rays(64, 64)
ctx = lbvh_amd.Context(0)
ctx.set_stream(torch.cuda.current_stream())
b = ctx.build_triangles(torch.from_numpy(tri).cuda(), code_bits=30, leaf_order=1)
prim = torch.empty(rays.shape[0], dtype=torch.int32, device="cuda")
t = torch.empty(rays.shape[0], dtype=torch.float32, device="cuda")
uv = torch.empty((rays.shape[0], 2), dtype=torch.float32, device="cuda")
ctx.trace(b, torch.from_numpy(rays).cuda(), 0, out=(prim, t, uv))
torch.cuda.synchronize()
on, _, _, _ = oracle.build(tri, 30, 1)
op, ot, _ = oracle.trace(on, tri, rays, 0)
assert b.nodes().tobytes() == on.tobytes()
assert np.array_equal(prim.cpu().numpy().view(np.uint32), op)
assert np.array_equal(t.cpu().numpy().view(np.uint32), ot.view(np.uint32))
print("torch interop ok")
"""


def test_torch_tensors_share_the_runtime(tmp_path):
    """torch CUDA tensors work as device arrays when torch is imported first
    (run in a fresh process: one HIP runtime per process)."""
    import subprocess
    import sys
    from conftest import ROOT
    r = subprocess.run([sys.executable, "-c", TORCH_INTEROP % ROOT], capture_output=True, text=True,
                       timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "torch interop ok" in r.stdout


def test_cpp_dropin_conformance(tmp_path):
    """The C++ drop-in header (include/lbvh_mi355x.hpp) used with the
    reference's call shapes: builder over triangles and over a user primitive
    with a lambda converter, host traverser, device tree; nodes and hits
    against the oracle."""
    import os
    import subprocess
    from lbvh_amd import build as B
    if not os.path.exists(B.DROPIN_BIN):
        pytest.skip("drop-in conformance program not built")
    out = tmp_path / "dropin.bin"
    r = subprocess.run([B.DROPIN_BIN, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:8], np.uint64)[0])
    off = 8
    tri = np.frombuffer(raw[off:off + 48 * n], np.float32).reshape(n, 12)
    off += 48 * n
    nodes = np.frombuffer(raw[off:off + 32 * (n - 1)], oracle.NODE_DTYPE)
    off += 32 * (n - 1)
    m = int(np.frombuffer(raw[off:off + 8], np.uint64)[0])
    off += 8
    rays = np.frombuffer(raw[off:off + 32 * m], np.float32).reshape(m, 8)
    off += 32 * m
    hp = np.frombuffer(raw[off:off + 4 * m], np.uint32)
    off += 4 * m
    ht = np.frombuffer(raw[off:off + 4 * m], np.float32)
    on, _, _, _ = oracle.build(tri, 30, 0)
    assert nodes.tobytes() == on.tobytes()
    op, ot, _ = oracle.trace(on, tri, rays, 0)
    assert np.array_equal(hp, op)
    assert np.array_equal(ht.view(np.uint32), ot.view(np.uint32))
    # builder<double> over triangle<double> (its traversal is checked against
    # brute force inside the program)
    off += 4 * m + 64
    n64 = int(np.frombuffer(raw[off:off + 8], np.uint64)[0])
    off += 8
    tri64 = np.frombuffer(raw[off:off + 96 * n64], np.float64).reshape(n64, 12)
    off += 96 * n64
    nodes64 = np.frombuffer(raw[off:off + 64 * (n64 - 1)], oracle.NODE64_DTYPE)
    off += 64 * (n64 - 1)
    nodes64_pid = np.frombuffer(raw[off:off + 64 * (n64 - 1)], oracle.NODE64_DTYPE)
    assert off + 64 * (n64 - 1) == len(raw)
    boxes64 = oracle.triangle_aabb_f64(tri64)
    want, _, _ = oracle.build_f64(boxes64, 0)
    assert nodes64.tobytes() == want.tobytes()
    want_pid, _, _ = oracle.build_f64(boxes64, 1)
    assert nodes64_pid.tobytes() == want_pid.tobytes()


def test_aabb_build_equals_triangle_build(ctx):
    v, tri = W.scene_uniform(4, 50000, 0.01)
    boxes = oracle.triangle_aabb(tri)
    a = ctx.build_triangles(tri, code_bits=30, leaf_order=1)
    b = ctx.build_aabbs(boxes, code_bits=30, leaf_order=1)
    assert a.nodes().tobytes() == b.nodes().tobytes()


def test_phase_timing_reports_four_phases(ctx):
    v, tri = W.scene_uniform(1, 65536, 0.01)
    ctx.set_timing(True)
    bvh = ctx.build_triangles(tri, code_bits=30)
    ms = ctx.phase_ms()
    ctx.set_timing(False)
    assert len(ms) == 4 and all(x >= 0 for x in ms)


def test_autotuner_times_candidates_and_keeps_results_exact():
    """A fresh context tunes the sort on the first build of a size bucket:
    every candidate is timed, one is kept, and the tree equals the oracle's
    whatever configuration was picked."""
    c = lbvh_amd.Context(0)
    v = W.cluster_vertices(91, 200000)
    tri = W.make_triangles(v)
    bvh = c.build_triangles(tri, code_bits=63, leaf_order=1, keep_codes=True)
    log = c.tune_log()
    # 4 configurations each of reduce-then-scan, onesweep and the hybrid
    assert len(log) == 12 and all(r["ms"] > 0 for r in log)
    picked = [r for r in log if r["picked"]]
    assert len(picked) == 1
    inf = bvh.info()
    assert inf.sort_digit_bits == picked[0]["digit_bits"]
    assert inf.sort_tile == picked[0]["tile"]
    on, _, osc, oids = oracle.build(tri, 63, 1)
    assert bvh.nodes().tobytes() == on.tobytes()
    sc, ids, _ = bvh.codes()
    assert np.array_equal(sc, osc) and np.array_equal(ids, oids)
    # the same bucket again: no new tuning
    bvh2 = c.build_triangles(tri, code_bits=63, leaf_order=1)
    assert len(c.tune_log()) == 12
    assert bvh2.info().sort_tile == inf.sort_tile
    # a batch of the same size tunes its own candidates (no onesweep)
    sizes = W.loguniform_sizes(5, 64, 1000, 8000)
    btri, off = W.batched_scenes(seed=5, sizes=sizes)
    b = c.build_batch(btri, off, leaf_order=1)
    blog = [r for r in c.tune_log() if r["batch"]]
    assert len(blog) == 8 and sum(r["picked"] for r in blog) == 1
    for s in (0, 31, 63):
        on, _, _, _ = oracle.build(btri[int(off[s]):int(off[s + 1])], 30, 1)
        assert b.nodes(s).tobytes() == on.tobytes()
    # explicit options bypass the tuner
    c.set_autotune(False)
    bvh3 = c.build_triangles(W.make_triangles(W.cluster_vertices(92, 1 << 19)), code_bits=30)
    assert len(c.tune_log()) == 20
    assert bvh3.info().sort_strategy in (1, 2, 3)
    for x in (bvh, bvh2, bvh3, b):
        x.destroy()
    c.close()


def test_traversal_tuner_records_each_mode_and_keeps_results_exact():
    """The first traversal of a (mode, ray-count) bucket times every
    traversal configuration (rays as given or in key order, x 8, 6 or 4
    waves per SIMD) on its own rays; each mode is tuned apart, exactly one
    configuration is kept per bucket, and the hits equal the oracle's on the
    tuning call and on the calls that reuse the choice."""
    c = lbvh_amd.Context(0)
    tri = W.make_triangles(W.cluster_vertices(93, 100000))
    bvh = c.build_triangles(tri, code_bits=63, leaf_order=1)
    on, _, _, _ = oracle.build(tri, 63, 1)
    rays = W.primary_rays(256, 256)
    for mode in (0, 2):
        want = oracle.trace(on, tri, rays, mode)
        for _ in range(2):
            p, t, uv = c.trace(bvh, rays, mode)
            assert np.array_equal(p, want[0])
            if mode == 0:
                assert np.array_equal(t.view(np.uint32), want[1].view(np.uint32))
                assert np.array_equal(uv.view(np.uint32), want[2].view(np.uint32))
    log = [r for r in c.tune_log() if r["kind"] == "traversal"]
    assert len(log) == 12
    for mode in (0, 2):
        mine = [r for r in log if r["mode"] == mode]
        assert len(mine) == 6 and all(r["ms"] > 0 and r["log2_n"] == 16 for r in mine)
        assert sum(r["picked"] for r in mine) == 1
        assert sorted((r["reorder"], r["waves_per_simd"]) for r in mine) == \
            sorted((ro, w) for ro in (False, True) for w in (8, 6, 4))
    bvh.destroy()
    c.close()


@pytest.mark.parametrize("n", [2, 3, 5, 64])
@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
def test_trace_edge_rays_match_oracle(ctx, n, leaf):
    """Tiny trees and degenerate rays (axis-aligned, in a triangle's plane,
    NaN and infinite components, tmin > tmax, zero direction) in every
    intersector mode: ids, t and uv equal the oracle's bit for bit."""
    from test_oracle import edge_rays32
    tri = W.make_triangles(W.uniform_vertices(80 + n, n, 0.2))
    bvh = ctx.build_triangles(tri, code_bits=30, leaf_order=leaf)
    on, _, _, _ = oracle.build(tri, 30, leaf)
    assert bvh.nodes().tobytes() == on.tobytes()
    rays = edge_rays32(tri)
    for mode in (0, 1, 2):
        p, t, uv = ctx.trace(bvh, rays, mode)
        op, ot, ouv = oracle.trace(on, tri, rays, mode)
        assert np.array_equal(p, op), "mode %d ids" % mode
        assert np.array_equal(t.view(np.uint32), ot.view(np.uint32)), "mode %d t" % mode
        assert np.array_equal(uv.view(np.uint32), ouv.view(np.uint32)), "mode %d uv" % mode
    bvh.destroy()
