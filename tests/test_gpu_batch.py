"""GPU parity of scene batches: every scene of a segmented build, called
through the C ABI, equals the oracle's (and the single-scene engine path's)
build of that scene alone -- codes, sorted ids, bounds, node arrays bit for
bit -- and batched traversal gives each scene's hits exactly.  Rebuilds and
refits replayed from a HIP graph are checked on every replay."""
import numpy as np
import pytest

import lbvh_amd
import oracle
from lbvh_amd import hip
from lbvh_amd import workload as W

pytestmark = pytest.mark.gpu

# edge sizes (pairs, treelet boundaries at 1024 leaves, single-block and
# multi-block scenes) and a random spread
EDGE_SIZES = [2, 3, 5, 64, 1023, 1024, 1025, 2048, 4097, 70001]


def _mixed_batch(seed=11, extra=40, clustered_every=3):
    """Uniform scenes of the edge sizes plus random sizes; every third
    scene is clustered (many duplicate codes)."""
    rng = np.random.default_rng(seed)
    sizes = EDGE_SIZES + [int(x) for x in np.exp(rng.uniform(np.log(2), np.log(20000), extra))]
    sizes = [max(2, s) for s in sizes]
    rng.shuffle(sizes)
    tri, off = W.batched_scenes(seed=seed, sizes=sizes)
    for s in range(0, len(sizes), clustered_every):
        m = sizes[s]
        tri[int(off[s]):int(off[s + 1])] = W.make_triangles(W.cluster_vertices(seed * 1000 + s, m))
    return tri, off


def _scene(tri, off, s):
    return tri[int(off[s]):int(off[s + 1])]


def _aimed_rays(seed, tri, off, per_scene):
    """Box-surface origins (as the batched workload) with directions aimed at
    the centroid of a random triangle of the scene, so most rays hit."""
    rays, roff = W.batched_box_rays(seed, tri, off, per_scene)
    rng = np.random.default_rng(seed)
    for s in range(len(off) - 1):
        ts = _scene(tri, off, s)
        sl = slice(int(roff[s]), int(roff[s + 1]))
        pick = rng.integers(0, ts.shape[0], int(roff[s + 1] - roff[s]))
        a = ts[pick, 0:3].astype(np.float64)
        c = (3 * a - ts[pick, 3:6] + ts[pick, 6:9]) / 3.0  # (a + b + c) / 3
        d = c - rays[sl, 0:3]
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays[sl, 3:6] = d.astype(np.float32)
    return rays, roff


@pytest.fixture(scope="module")
def mixed():
    return _mixed_batch()


@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
def test_batch_scenes_equal_oracle(ctx, mixed, leaf):
    tri, off = mixed
    b = ctx.build_batch(tri, off, leaf_order=leaf, keep_codes=True)
    info = b.info()
    assert info.n_scenes == len(off) - 1 and info.n_prims == int(off[-1])
    for s in range(b.n_scenes):
        ts = _scene(tri, off, s)
        on, ob, osc, oids = oracle.build(ts, 30, leaf)
        codes, ids, bounds = b.codes(s)
        assert np.array_equal(bounds.view(np.uint32), ob.view(np.uint32)), "scene %d bounds" % s
        assert np.array_equal(codes, osc), "scene %d codes" % s
        assert np.array_equal(ids, oids), "scene %d ids" % s
        assert b.nodes(s).tobytes() == on.tobytes(), "scene %d nodes" % s
    b.destroy()


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["first", "lowid", "watertight"])
def test_batch_trace_equals_single_scene_trace(ctx, mixed, mode):
    tri, off = mixed
    rays, roff = W.batched_box_rays(5, tri, off, per_scene=256)
    aimed, aoff = _aimed_rays(7, tri, off, per_scene=256)
    rays = np.concatenate([rays, aimed])
    roff = np.concatenate([roff[:-1] * 2, roff[-1:] * 2])
    # interleave per scene: scene s gets its 256 random and 256 aimed rays
    nseg = len(off) - 1
    rays = np.stack([rays[:nseg * 256].reshape(nseg, 256, 8),
                     rays[nseg * 256:].reshape(nseg, 256, 8)], axis=1).reshape(-1, 8)
    b = ctx.build_batch(tri, off, leaf_order=1)
    p, t, uv = b.trace(rays, roff, mode)
    hits = 0
    for s in range(b.n_scenes):
        ts = _scene(tri, off, s)
        rs = rays[int(roff[s]):int(roff[s + 1])]
        single = ctx.build_triangles(ts, code_bits=30, leaf_order=1)
        sp, st, suv = ctx.trace(single, rs, mode)
        sl = slice(int(roff[s]), int(roff[s + 1]))
        assert np.array_equal(p[sl], sp), "scene %d ids" % s
        assert np.array_equal(t[sl].view(np.uint32), st.view(np.uint32)), "scene %d t" % s
        assert np.array_equal(uv[sl].view(np.uint32), suv.view(np.uint32)), "scene %d uv" % s
        hits += int(np.count_nonzero(sp != lbvh_amd.MISS))
        single.destroy()
    assert hits > len(rays) // 3
    b.destroy()


def test_batch_trace_equals_oracle(ctx, mixed):
    tri, off = mixed
    rays, roff = _aimed_rays(6, tri, off, per_scene=128)
    b = ctx.build_batch(tri, off, leaf_order=0)
    p, t, uv = b.trace(rays, roff, lbvh_amd.ISECT_REFERENCE_LOWID)
    for s in range(b.n_scenes):
        ts = _scene(tri, off, s)
        on, _, _, _ = oracle.build(ts, 30, 0)
        rs = rays[int(roff[s]):int(roff[s + 1])]
        op, ot, ouv = oracle.trace(on, ts, rs, lbvh_amd.ISECT_REFERENCE_LOWID)
        sl = slice(int(roff[s]), int(roff[s + 1]))
        assert np.array_equal(p[sl], op), "scene %d" % s
        assert np.array_equal(t[sl].view(np.uint32), ot.view(np.uint32)), "scene %d" % s
    b.destroy()


def test_batch_uneven_ray_counts(ctx, mixed):
    """Scenes with no rays and scenes with many share one launch."""
    tri, off = mixed
    nseg = len(off) - 1
    counts = np.array([(s * 37) % 5 * 50 for s in range(nseg)], dtype=np.uint64)
    roff = np.zeros(nseg + 1, dtype=np.uint64)
    np.cumsum(counts, out=roff[1:])
    full, full_off = _aimed_rays(8, tri, off, per_scene=200)
    rays = np.concatenate([full[int(full_off[s]):int(full_off[s]) + int(counts[s])]
                           for s in range(nseg)])
    b = ctx.build_batch(tri, off, leaf_order=1)
    p, t, _ = b.trace(rays, roff, 1)
    for s in range(nseg):
        if counts[s] == 0:
            continue
        single = ctx.build_triangles(_scene(tri, off, s), code_bits=30, leaf_order=1)
        sp, st, _ = ctx.trace(single, rays[int(roff[s]):int(roff[s + 1])], 1)
        assert np.array_equal(p[int(roff[s]):int(roff[s + 1])], sp)
        single.destroy()
    b.destroy()


@pytest.mark.parametrize("leaf", [0, 1], ids=["ref", "pid"])
def test_batch_refit_equals_oracle_fit(ctx, mixed, leaf):
    tri, off = mixed
    b = ctx.build_batch(hip.DeviceArray.from_host(tri), off, leaf_order=leaf)
    before = [b.nodes(s) for s in range(b.n_scenes)]
    tri2 = tri.copy()
    tri2[:, 0:3] += np.float32(3e-5) * np.sin(np.arange(tri.shape[0], dtype=np.float32))[:, None]
    d2 = hip.DeviceArray.from_host(tri2)
    b.refit(d2)
    ctx.synchronize()
    for s in range(b.n_scenes):
        after = b.nodes(s)
        assert np.array_equal(after["left"], before[s]["left"])
        assert np.array_equal(after["right"], before[s]["right"])
        # leaf values are sorted positions (REFERENCE) or primitive ids
        # (PRIMITIVE_ID); the fit reads primitives[leaf value] either way
        expect = oracle.fit(before[s], _scene(tri2, off, s))
        assert after.tobytes() == expect.tobytes(), "scene %d" % s
    b.destroy()


def test_batch_graph_rebuild_and_refit_replays(ctx, mixed):
    """The captured rebuild and refit give the eager result on every replay,
    with eager engine work and host synchronization between replays."""
    tri, off = mixed
    tri_b = tri.copy()
    tri_b[:, 0:3] += np.float32(0.25)
    d = hip.DeviceArray.from_host(tri)
    b = ctx.build_batch(d, off, leaf_order=1, keep_codes=True)
    ctx.synchronize()
    eager_a = [b.nodes(s).tobytes() for s in range(b.n_scenes)]
    expect_b = [oracle.build(_scene(tri_b, off, s), 30, 1)[0].tobytes()
                for s in range(b.n_scenes)]
    for replay in range(4):
        use_b = replay % 2 == 1
        d.copy_from(tri_b if use_b else tri)
        b.rebuild(graph=True)
        ctx.synchronize()
        want = expect_b if use_b else eager_a
        for s in range(b.n_scenes):
            assert b.nodes(s).tobytes() == want[s], "replay %d scene %d" % (replay, s)
        # eager work between replays: a single-scene build and trace
        single = ctx.build_triangles(_scene(tri, off, 0), code_bits=30)
        ctx.trace(single, W.primary_rays(8, 8))
        single.destroy()
    # refit replays: topology of A, boxes of the current data
    d.copy_from(tri)
    b.rebuild()
    ctx.synchronize()
    nodes_a = [b.nodes(s) for s in range(b.n_scenes)]
    for replay in range(3):
        src = tri_b if replay % 2 == 0 else tri
        d.copy_from(src)
        b.refit(graph=True)
        ctx.synchronize()
        for s in range(b.n_scenes):
            expect = oracle.fit(nodes_a[s], _scene(src, off, s))
            assert b.nodes(s).tobytes() == expect.tobytes(), "refit replay %d scene %d" % (replay, s)
    b.destroy()


def test_batch_sort_configurations_agree(ctx, mixed):
    tri, off = mixed
    ref = None
    for strategy, digit_bits, tile in [(2, 8, 2048), (2, 8, 4096), (2, 10, 2048), (2, 10, 4096),
                                       (3, 8, 4096), (3, 10, 4096)]:
        b = ctx.build_batch(tri, off, leaf_order=1, sort_digit_bits=digit_bits,
                            sort_tile=tile, sort_strategy=strategy)
        inf = b.info()
        assert inf.sort_digit_bits == digit_bits and inf.sort_tile == tile
        assert inf.sort_strategy == strategy
        got = b"".join(b.nodes(s).tobytes() for s in range(b.n_scenes))
        if ref is None:
            ref = got
        assert got == ref
        b.destroy()


@pytest.mark.parametrize("tile", [2048, 4096])
@pytest.mark.parametrize("leaf,keep", [(0, False), (0, True), (1, True)],
                         ids=["ref-keys-only", "ref-keep-codes", "pid"])
def test_batch_segmented_hybrid_equals_oracle(ctx, tile, leaf, keep):
    """The hybrid sort with 10-bit digits on 49 scenes (36-bit keys, global
    digits = the scene bits and code bits 20-29) runs as one pass that sorts
    each scene by code bits 20-29 over chunks of 16,384 keys that never cross
    a scene (scenes of 1 to 7 chunks, chunk-sized and one-past scenes, and
    many tiny ones), then the LDS bucket sort.  Every scene's codes, ids and
    nodes equal the oracle's; clustered scenes bring long runs of equal
    codes (ties in input order).  Reference leaves without keep_codes sort
    keys alone (no id payload)."""
    sizes = [100000, 2, 3, 16384, 16385, 64, 40000] + [5] * 40 + [5000, 32768]
    tri, off = W.batched_scenes(seed=31, sizes=sizes)
    for s in (0, 4, 6):
        m = sizes[s]
        tri[int(off[s]):int(off[s + 1])] = W.make_triangles(W.cluster_vertices(31000 + s, m))
    b = ctx.build_batch(tri, off, leaf_order=leaf, keep_codes=keep, sort_strategy=3,
                        sort_digit_bits=10, sort_tile=tile)
    inf = b.info()
    assert inf.sort_strategy == 3 and inf.sort_digit_bits == 10
    assert inf.sort_global_passes == 1
    for s in range(len(sizes)):
        on, ob, osc, oids = oracle.build(_scene(tri, off, s), 30, leaf)
        if keep:
            codes, ids, bounds = b.codes(s)
            assert np.array_equal(codes, osc), "scene %d codes" % s
            assert np.array_equal(ids, oids), "scene %d ids" % s
        assert b.nodes(s).tobytes() == on.tobytes(), "scene %d nodes" % s
    b.rebuild(graph=True)  # the same sort replayed from a graph
    ctx.synchronize()
    for s in range(len(sizes)):
        on = oracle.build(_scene(tri, off, s), 30, leaf)[0]
        assert b.nodes(s).tobytes() == on.tobytes(), "graph replay scene %d" % s
    b.destroy()


def test_batch_segmented_hybrid_one_scene(ctx):
    """A one-scene batch (30-bit keys, no scene bits) of 50,000 clustered
    triangles: the hybrid's one global digit is code bits 20-29, sorted by
    the segmented pass over the scene's four chunks; equal to the oracle."""
    tri, off = W.batched_scenes(seed=32, sizes=[50000])
    tri[:] = W.make_triangles(W.cluster_vertices(32000, 50000))
    b = ctx.build_batch(tri, off, leaf_order=1, keep_codes=True, sort_strategy=3,
                        sort_digit_bits=10, sort_tile=2048)
    assert b.info().sort_global_passes == 1
    on, ob, osc, oids = oracle.build(tri, 30, 1)
    codes, ids, bounds = b.codes(0)
    assert np.array_equal(codes, osc)
    assert np.array_equal(ids, oids)
    assert b.nodes(0).tobytes() == on.tobytes()
    b.destroy()


def test_batch_rejects_bad_arguments(ctx):
    tri, off = W.batched_scenes(seed=3, sizes=[4, 5, 6])
    with pytest.raises(lbvh_amd.LbvhError) as e:
        ctx.build_batch(tri[:10], np.array([0, 4, 5, 10], dtype=np.uint64))  # a 1-triangle scene
    assert e.value.status == 1
    with pytest.raises(lbvh_amd.LbvhError):
        ctx.build_batch(tri, np.array([1, 5, 15], dtype=np.uint64))
    with pytest.raises(ValueError):
        ctx.build_batch(tri, np.array([0, 4, 9], dtype=np.uint64))  # does not end at n
    b = ctx.build_batch(tri, off)
    with pytest.raises(ValueError):
        b.trace(np.zeros((4, 8), np.float32), np.array([0, 4], dtype=np.uint64))
    b.destroy()


def test_batch_of_many_scenes_matches_single_builds(ctx):
    """512 scenes log-uniform in [64, 65536] (about 4.8M triangles): every
    scene's node array equals the single-scene build of that scene."""
    sizes = W.loguniform_sizes(44, 512, 64, 65536)
    tri, off = W.batched_scenes(seed=44, sizes=sizes)
    d = hip.DeviceArray.from_host(tri)
    b = ctx.build_batch(d, off, leaf_order=1)
    ctx.synchronize()
    for s in range(b.n_scenes):
        single = ctx.build_triangles(_scene(tri, off, s), code_bits=30, leaf_order=1)
        assert b.nodes(s).tobytes() == single.nodes().tobytes(), "scene %d" % s
        single.destroy()
    b.destroy()


def test_batch_split_tree_merges_equal_oracle(ctx, mixed, monkeypatch):
    """The mixed batch with every upper-level tree group forced through the
    split merges (LBVH_TREE_GROUP_SLOTS=1): every scene equals the oracle."""
    monkeypatch.setenv("LBVH_TREE_GROUP_SLOTS", "1")
    tri, off = mixed
    b = ctx.build_batch(tri, off, leaf_order=1, keep_codes=True)
    for s in range(b.n_scenes):
        on, ob, osc, oids = oracle.build(_scene(tri, off, s), 30, 1)
        assert b.nodes(s).tobytes() == on.tobytes(), "scene %d nodes" % s
    b.destroy()


@pytest.mark.parametrize("sizes", [[4, 5, 6], [2, 3], [3000, 7000, 2, 50000], [1025, 2]],
                         ids=["3-tiny", "2-tiny", "4-mixed", "2-edge"])
@pytest.mark.parametrize("strategy", [2, 3], ids=["rts", "hybrid"])
def test_batch_of_few_scenes_equals_oracle(ctx, sizes, strategy):
    """Batches of 2-4 scenes have at most 32 key bits (30 + scene bits) in
    64-bit keys: they must still sort as 64-bit keys."""
    tri, off = W.batched_scenes(seed=21, sizes=sizes)
    for leaf in (0, 1):
        b = ctx.build_batch(tri, off, leaf_order=leaf, keep_codes=True, sort_strategy=strategy)
        assert b.info().sort_strategy == strategy
        for s in range(len(sizes)):
            ts = _scene(tri, off, s)
            on, ob, osc, oids = oracle.build(ts, 30, leaf)
            codes, ids, bounds = b.codes(s)
            assert np.array_equal(codes, osc), "scene %d codes" % s
            assert np.array_equal(ids, oids), "scene %d ids" % s
            assert b.nodes(s).tobytes() == on.tobytes(), "scene %d nodes" % s
        b.destroy()
