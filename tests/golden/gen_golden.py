"""Writes the golden fixtures under tests/golden/ from the real tay10r/lbvh.

Run in a container that holds the reference sources (the recipe
oracle/build_oracle.py compiles oracle/ref_driver.cpp against
/root/reference/lbvh.h into oracle/_ref/libreflbvh.so).  The fixtures are
data only: inputs (triangles, rays) and the reference's outputs.

  small.npz        uniform scenes n in {2,3,4,5,7,8,16,17,64,1000,4096}
  adversarial.npz  duplicate codes, tight clusters, a flat axis, 1e6-scale
                   and negative coordinates, repeated blocks
  meshes.npz       the teapot and monkey meshes shipped with the reference
  f64.npz          lbvh::builder<double>: triangle<double> scenes (uniform
                   sizes, duplicates, clusters, a flat axis, 1e6-scale
                   coordinates, the teapot) and plain aabb<double> inputs
                   (wide-range, point and nested boxes); node<double> arrays
                   for both leaf orders
  hashes.json      FNV-1a-64 hashes of inputs and outputs for the full
                   65,536-triangle, 1,048,576-triangle and 16,777,216-triangle
                   workloads

For every fixture case and each (code bits 30/63) x (leaf order reference /
primitive id): centroid bounds, unsorted codes, sorted codes, the node array,
and closest-hit results (prim, t, uv) of a ray set with the reference tie
rule (mode 0) and the lowest-id tie rule (mode 1).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import build_oracle  # noqa: E402
from lbvh_amd import workload as W  # noqa: E402

REF_MODELS = "/root/reference/models"


def load_obj(path):
    """Triangles of an OBJ file (faces fan-triangulated), as vertex triples."""
    verts = []
    tris = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f":
                idx = [int(p.split("/")[0]) for p in parts[1:]]
                idx = [i - 1 if i > 0 else len(verts) + i for i in idx]
                for k in range(1, len(idx) - 1):
                    tris.append(verts[idx[0]] + verts[idx[k]] + verts[idx[k + 1]])
    return np.asarray(tris, dtype=np.float32)


def rays_for(tri, seed, count=256):
    """Half the rays start on the scene box surface in random directions; the
    other half aim from a surrounding sphere at random primitive centres."""
    rng = np.random.default_rng(seed)
    boxes = oracle.triangle_aabb(tri) if len(tri) < 20000 else None
    lo = tri[:, 0:3].min(axis=0) - 1e-3
    hi = tri[:, 0:3].max(axis=0) + 1e-3
    half = count // 2
    r1 = W.box_rays(seed, lo, hi, half)
    if boxes is None:
        boxes = oracle.triangle_aabb(tri[:4096])
    c = ((boxes[:, 0:3] + boxes[:, 3:6]) * 0.5).astype(np.float64)
    centre = (lo.astype(np.float64) + hi) * 0.5
    rad = float(np.max(hi - lo)) * 1.5 + 1e-3
    k = count - half
    pick = rng.integers(0, c.shape[0], size=k)
    u = rng.normal(size=(k, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o = centre + rad * u
    d = c[pick] - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r2 = np.empty((k, 8), dtype=np.float32)
    r2[:, 0:3] = o
    r2[:, 3:6] = d
    r2[:, 6] = 0.0
    r2[:, 7] = np.inf
    return np.concatenate([r1, r2]).astype(np.float32)


def case_record(name, tri, seed):
    rec = {"tri": tri.astype(np.float32)}
    rays = rays_for(tri, seed)
    rec["rays"] = rays
    for bits in (30, 63):
        rec["codes%d" % bits] = oracle.ref_codes(tri, bits)
        for lo in (0, 1):
            nodes, sc, bounds = oracle.ref_build(tri, bits, lo)
            key = "b%d_l%d" % (bits, lo)
            rec[key + "_nodes"] = nodes.view(np.uint32).reshape(-1, 8)
            rec[key + "_sorted"] = sc
            rec["bounds"] = bounds
            if bits == 30:  # traversal depends on the tree; 30-bit trees suffice
                for mode in (0, 1):
                    p, t, uv = oracle.ref_trace(nodes, tri, rays, mode)
                    rec["%s_m%d_prim" % (key, mode)] = p
                    rec["%s_m%d_t" % (key, mode)] = t
                    rec["%s_m%d_uv" % (key, mode)] = uv
    return {name + "/" + k: v for k, v in rec.items()}


def small_cases():
    out = {}
    for n in (2, 3, 4, 5, 7, 8, 16, 17, 64, 1000, 4096):
        v = W.uniform_vertices(100 + n, n, 0.01)
        out.update(case_record("n%d" % n, W.make_triangles(v), 100 + n))
    return out


def adversarial_cases():
    out = {}
    rng = np.random.default_rng(7)
    # every code duplicated: identical triangles
    v = np.tile(np.array([[0.2, 0.3, 0.4, 0.21, 0.3, 0.4, 0.2, 0.31, 0.41]], np.float32), (300, 1))
    out.update(case_record("all_equal", W.make_triangles(v), 11))
    # two tight clusters
    a = rng.uniform(-1e-4, 1e-4, size=(500, 9)) + 0.1
    b = rng.uniform(-1e-4, 1e-4, size=(500, 9)) + 0.9
    out.update(case_record("two_clusters", W.make_triangles(np.concatenate([a, b]).astype(np.float32)), 12))
    # zero extent along z
    v = W.uniform_vertices(13, 1000, 0.01)
    v[:, 2::3] = 0.5
    out.update(case_record("flat_z", W.make_triangles(v), 13))
    # negative, 1e6-scale coordinates
    c = rng.uniform(-1e6, 1e6, size=(2000, 3))
    off = rng.uniform(-1e3, 1e3, size=(2000, 9))
    v = (np.tile(c, (1, 3)) + off).astype(np.float32)
    out.update(case_record("big_negative", W.make_triangles(v), 14))
    # 64 distinct positions repeated 32 times each (long equal-code runs)
    base = W.uniform_vertices(15, 64, 0.01)
    v = np.repeat(base, 32, axis=0)
    out.update(case_record("dup_blocks", W.make_triangles(v), 15))
    return out


def mesh_cases():
    out = {}
    for name in ("teapot", "monkey"):
        path = os.path.join(REF_MODELS, name + ".obj")
        v = load_obj(path)
        out.update(case_record(name, W.make_triangles(v), 21))
    return out


def tri96(v):
    """triangle<double> records {p0, e0 = a - b, e1 = c - a, normal} from
    vertex triples, computed in FP64."""
    v = np.asarray(v, dtype=np.float64).reshape(-1, 9)
    a, b, c = v[:, 0:3], v[:, 3:6], v[:, 6:9]
    e0 = a - b
    e1 = c - a
    nrm = np.stack([e0[:, 1] * e1[:, 2] - e0[:, 2] * e1[:, 1],
                    e0[:, 2] * e1[:, 0] - e0[:, 0] * e1[:, 2],
                    e0[:, 0] * e1[:, 1] - e0[:, 1] * e1[:, 0]], axis=1)
    return np.ascontiguousarray(np.concatenate([a, e0, e1, nrm], axis=1))


def f64_record(name, prims, kind):
    rec = {"prims": prims}
    for lo in (0, 1):
        rec["l%d_nodes" % lo] = oracle.ref_build_f64(prims, kind, lo).view(np.uint64).reshape(-1, 8)
    return {kind + "/" + name + "/" + k: v for k, v in rec.items()}


def f64_cases():
    out = {}
    rng = np.random.default_rng(31)
    for n in (2, 3, 5, 17, 64, 1000, 2048):
        out.update(f64_record("n%d" % n, tri96(rng.random((n, 9))), "triangles"))
    out.update(f64_record("all_equal", tri96(np.tile(rng.random((1, 9)), (300, 1))), "triangles"))
    a = rng.uniform(-1e-9, 1e-9, size=(400, 9)) + 0.1
    b = rng.uniform(-1e-9, 1e-9, size=(400, 9)) + 0.9
    out.update(f64_record("two_clusters", tri96(np.concatenate([a, b])), "triangles"))
    v = rng.random((1000, 9))
    v[:, 2::3] = 0.5
    out.update(f64_record("flat_z", tri96(v), "triangles"))
    c = rng.uniform(-1e6, 1e6, size=(1500, 3))
    v = np.tile(c, (1, 3)) + rng.uniform(-1e3, 1e3, size=(1500, 9))
    out.update(f64_record("big_negative", tri96(v), "triangles"))
    out.update(f64_record("dup_blocks", tri96(np.repeat(rng.random((48, 9)), 24, axis=0)), "triangles"))
    teapot = load_obj(os.path.join(REF_MODELS, "teapot.obj")).astype(np.float64)
    out.update(f64_record("teapot", tri96(teapot), "triangles"))
    # plain boxes through the generic converter path
    lo = rng.uniform(-1e4, 1e4, size=(1500, 3))
    ext = rng.exponential(10.0, size=(1500, 3))
    out.update(f64_record("wide", np.concatenate([lo, lo + ext], axis=1), "aabbs"))
    p = rng.random((700, 3))
    out.update(f64_record("points", np.concatenate([p, p], axis=1), "aabbs"))
    r = np.linspace(1.0, 2.0, 200)[:, None] * np.ones((1, 3))
    out.update(f64_record("nested", np.concatenate([-r, r], axis=1), "aabbs"))
    return out


def config_hashes():
    H = {}
    h = oracle.fnv1a64
    # 65,536 uniform triangles, seed 1, offsets 0.01; 30-bit; 256x256 primary rays
    v, tri = W.scene_uniform(1, 65536, 0.01)
    rays = W.primary_rays(256, 256)
    e = {"tri": h(tri), "rays": h(rays), "codes30": h(oracle.ref_codes(tri, 30))}
    for lo in (0, 1):
        nodes, sc, bounds = oracle.ref_build(tri, 30, lo)
        e["b30_l%d_nodes" % lo] = h(nodes)
        e["b30_l%d_sorted" % lo] = h(sc)
        e["bounds"] = h(bounds)
        for mode in (0, 1):
            p, t, uv = oracle.ref_trace(nodes, tri, rays, mode)
            e["b30_l%d_m%d_hits" % (lo, mode)] = h(np.concatenate([p.view(np.uint32), t.view(np.uint32)]))
            e["b30_l%d_m%d_nhit" % (lo, mode)] = int(np.sum(p != 0xFFFFFFFF))
    nodes, sc, bounds = oracle.ref_build(tri, 63, 1)
    e["b63_l1_nodes"] = h(nodes)
    H["uniform65536_seed1"] = e
    print("uniform65536 done", flush=True)
    # 1,048,576 clustered triangles, seed 2 (Gaussians use libm)
    v, tri = W.scene_clustered(2, 1 << 20)
    e = {"tri": h(tri), "libm_dependent": True}
    for bits in (30, 63):
        for lo in (0, 1):
            nodes, sc, bounds = oracle.ref_build(tri, bits, lo)
            e["b%d_l%d_nodes" % (bits, lo)] = h(nodes)
            e["b%d_l%d_sorted" % (bits, lo)] = h(sc)
    H["clustered1m_seed2"] = e
    print("clustered1m done", flush=True)
    # 16,777,216 uniform triangles, seed 3, offsets 5e-4
    v, tri = W.scene_uniform(3, 1 << 24, 5e-4)
    del v
    e = {"tri": h(tri)}
    for bits in (30, 63):
        for lo in (0, 1):
            nodes, sc, bounds = oracle.ref_build(tri, bits, lo)
            e["b%d_l%d_nodes" % (bits, lo)] = h(nodes)
            e["b%d_l%d_sorted" % (bits, lo)] = h(sc)
            e["bounds"] = h(bounds)
            del nodes, sc
            print("uniform16m", bits, lo, flush=True)
    H["uniform16m_seed3"] = e
    return H


def main():
    build_oracle.build_all()
    if not oracle.ref_available():
        raise SystemExit("reference driver not available: fixtures need the reference sources")
    np.savez_compressed(os.path.join(HERE, "f64.npz"), **f64_cases())
    print("f64 done", flush=True)
    if "--f64-only" in sys.argv:
        return
    np.savez_compressed(os.path.join(HERE, "small.npz"), **small_cases())
    print("small done", flush=True)
    np.savez_compressed(os.path.join(HERE, "adversarial.npz"), **adversarial_cases())
    print("adversarial done", flush=True)
    np.savez_compressed(os.path.join(HERE, "meshes.npz"), **mesh_cases())
    print("meshes done", flush=True)
    if "--no-hashes" not in sys.argv:
        with open(os.path.join(HERE, "hashes.json"), "w") as f:
            json.dump(config_hashes(), f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
