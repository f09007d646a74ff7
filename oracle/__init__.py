"""ORACLE — test infrastructure only.

ctypes access to
  * the clean-room CPU restatement (oracle/_build/liblbvh_oracle.so), and
  * the driver around the real tay10r/lbvh header (oracle/_ref/libreflbvh.so,
    present only where it was built from the reference sources).

Only tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this package, and only as the checker or the timed CPU baseline.  The engine
never calls it.
"""
import ctypes
import os

import numpy as np

from . import build_oracle as _b

F32P = ctypes.POINTER(ctypes.c_float)
F64P = ctypes.POINTER(ctypes.c_double)
U32P = ctypes.POINTER(ctypes.c_uint32)
U64P = ctypes.POINTER(ctypes.c_uint64)
VP = ctypes.c_void_p

LEAF_REFERENCE = 0
LEAF_PRIMITIVE_ID = 1
ISECT_REFERENCE = 0
ISECT_REFERENCE_LOWID = 1
ISECT_WATERTIGHT = 2

NODE_DTYPE = np.dtype([("box", np.float32, (6,)), ("left", np.uint32), ("right", np.uint32)])
assert NODE_DTYPE.itemsize == 32
# lbvh::node<double>: 64 bytes, leaf flag in bit 63.
NODE64_DTYPE = np.dtype([("box", np.float64, (6,)), ("left", np.uint64), ("right", np.uint64)])
assert NODE64_DTYPE.itemsize == 64
LEAF_BIT64 = np.uint64(1 << 63)

_orc = None
_ref = {}


def _p(a, t):
    return a.ctypes.data_as(t)


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(_b.ORACLE_SO):
            raise RuntimeError("oracle library not built (oracle/build_oracle.py)")
        L = ctypes.CDLL(_b.ORACLE_SO)
        L.orc_triangle_aabb.argtypes = [F32P, F32P]
        L.orc_centroid_bounds.argtypes = [F32P, ctypes.c_uint64, F32P]
        L.orc_morton_codes.argtypes = [F32P, ctypes.c_uint64, ctypes.c_int, F32P, VP]
        L.orc_sort_pairs.argtypes = [ctypes.c_int, VP, ctypes.c_uint64, VP, U32P]
        L.orc_karras.argtypes = [ctypes.c_int, VP, ctypes.c_uint64, VP]
        L.orc_karras.restype = ctypes.c_int
        L.orc_fit.argtypes = [VP, ctypes.c_uint64, F32P, U32P]
        L.orc_build.argtypes = [F32P, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, VP, F32P, VP, U32P]
        L.orc_build.restype = ctypes.c_int
        L.orc_trace.argtypes = [VP, ctypes.c_uint64, F32P, F32P, ctypes.c_uint64, ctypes.c_int,
                                ctypes.c_int, U32P, F32P, F32P, U64P]
        L.orc_brute_closest.argtypes = [F32P, ctypes.c_uint64, F32P, ctypes.c_uint64, ctypes.c_int,
                                        ctypes.c_int, U32P, F32P]
        L.orc_t_fp64.argtypes = [F32P, F32P, U32P, ctypes.c_uint64, F64P]
        L.orc_triangle_aabb_f64.argtypes = [F64P, ctypes.c_uint64, F64P]
        L.orc_morton_codes_f64.argtypes = [F64P, ctypes.c_uint64, F64P, U64P]
        L.orc_karras_f64.argtypes = [U64P, ctypes.c_uint64, VP]
        L.orc_karras_f64.restype = ctypes.c_int
        L.orc_fit_f64.argtypes = [VP, ctypes.c_uint64, F64P]
        L.orc_build_f64.argtypes = [F64P, ctypes.c_uint64, ctypes.c_int, VP, U64P, U32P]
        L.orc_build_f64.restype = ctypes.c_int
        L.orc_trace_f64.argtypes = [VP, ctypes.c_uint64, F64P, F64P, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.c_int, U64P, F64P, F64P]
        L.orc_brute_closest_f64.argtypes = [F64P, ctypes.c_uint64, F64P, ctypes.c_uint64,
                                            ctypes.c_int, U64P, F64P]
        L.orc_fnv1a64.argtypes = [VP, ctypes.c_uint64]
        L.orc_fnv1a64.restype = ctypes.c_uint64
        _orc = L
    return _orc


def ref_available(fast=False):
    return os.path.exists(_b.REF_FAST_SO if fast else _b.REF_SO)


def ref(fast=False):
    key = "fast" if fast else "strict"
    if key not in _ref:
        path = _b.REF_FAST_SO if fast else _b.REF_SO
        if not os.path.exists(path):
            raise RuntimeError("reference driver not built: " + path)
        L = ctypes.CDLL(path)
        L.ref_build.argtypes = [F32P, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                VP, VP, F32P, F64P]
        L.ref_build.restype = ctypes.c_int
        L.ref_codes.argtypes = [F32P, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, VP]
        L.ref_codes.restype = ctypes.c_int
        L.ref_trace.argtypes = [VP, ctypes.c_uint64, F32P, F32P, ctypes.c_uint64, ctypes.c_int,
                                ctypes.c_int, U32P, F32P, F32P]
        L.ref_trace.restype = ctypes.c_int
        L.ref_build_f64_triangles.argtypes = [F64P, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, VP]
        L.ref_build_f64_triangles.restype = ctypes.c_int
        L.ref_build_f64_aabbs.argtypes = [F64P, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, VP]
        L.ref_build_f64_aabbs.restype = ctypes.c_int
        L.ref_trace_f64.argtypes = [VP, ctypes.c_uint64, F64P, F64P, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.c_int, U64P, F64P, F64P]
        L.ref_trace_f64.restype = ctypes.c_int
        L.ref_hardware_concurrency.restype = ctypes.c_int
        _ref[key] = L
    return _ref[key]


def _tri(tri):
    return np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 12)


def _code_dtype(bits):
    return np.uint32 if bits == 30 else np.uint64


# ---- restatement -----------------------------------------------------------

def triangle_aabb(tri):
    tri = _tri(tri)
    out = np.empty((tri.shape[0], 6), dtype=np.float32)
    for i in range(tri.shape[0]):
        orc().orc_triangle_aabb(_p(tri[i], F32P), _p(out[i], F32P))
    return out


def centroid_bounds(tri):
    tri = _tri(tri)
    out = np.empty(6, dtype=np.float32)
    orc().orc_centroid_bounds(_p(tri, F32P), tri.shape[0], _p(out, F32P))
    return out


def morton_codes(tri, code_bits, bounds=None):
    tri = _tri(tri)
    if bounds is None:
        bounds = centroid_bounds(tri)
    bounds = np.ascontiguousarray(bounds, dtype=np.float32)
    out = np.empty(tri.shape[0], dtype=_code_dtype(code_bits))
    orc().orc_morton_codes(_p(tri, F32P), tri.shape[0], code_bits, _p(bounds, F32P), out.ctypes.data)
    return out


def sort_pairs(codes, code_bits):
    codes = np.ascontiguousarray(codes, dtype=_code_dtype(code_bits))
    n = codes.shape[0]
    sc = np.empty(n, dtype=codes.dtype)
    ids = np.empty(n, dtype=np.uint32)
    orc().orc_sort_pairs(code_bits, codes.ctypes.data, n, sc.ctypes.data, _p(ids, U32P))
    return sc, ids


def karras(sorted_codes, code_bits):
    sc = np.ascontiguousarray(sorted_codes, dtype=_code_dtype(code_bits))
    n = sc.shape[0]
    nodes = np.zeros(n - 1, dtype=NODE_DTYPE)
    rc = orc().orc_karras(code_bits, sc.ctypes.data, n, nodes.ctypes.data)
    if rc != 0:
        raise ValueError("orc_karras failed")
    return nodes


def fit(nodes, tri, leaf_prim=None):
    nodes = np.array(nodes, dtype=NODE_DTYPE, copy=True)
    tri = _tri(tri)
    lp = None if leaf_prim is None else np.ascontiguousarray(leaf_prim, dtype=np.uint32)
    orc().orc_fit(nodes.ctypes.data, tri.shape[0], _p(tri, F32P), None if lp is None else _p(lp, U32P))
    return nodes


def build(tri, code_bits=30, leaf_order=LEAF_REFERENCE):
    """Returns (nodes, bounds[6], sorted_codes, ids)."""
    tri = _tri(tri)
    n = tri.shape[0]
    nodes = np.zeros(n - 1, dtype=NODE_DTYPE)
    bounds = np.empty(6, dtype=np.float32)
    sc = np.empty(n, dtype=_code_dtype(code_bits))
    ids = np.empty(n, dtype=np.uint32)
    rc = orc().orc_build(_p(tri, F32P), n, code_bits, leaf_order, nodes.ctypes.data,
                         _p(bounds, F32P), sc.ctypes.data, _p(ids, U32P))
    if rc != 0:
        raise ValueError("orc_build failed (n < 2 or bad code width)")
    return nodes, bounds, sc, ids


# ---- builder<double> -------------------------------------------------------

def _f64(a, w):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1, w)


def triangle_aabb_f64(tri96):
    """triangle_aabb_converter<double> over triangle<double> records [n, 12]."""
    tri96 = _f64(tri96, 12)
    out = np.empty((tri96.shape[0], 6), dtype=np.float64)
    orc().orc_triangle_aabb_f64(_p(tri96, F64P), tri96.shape[0], _p(out, F64P))
    return out


def morton_codes_f64(boxes48):
    """Returns (codes u64 in primitive order, centroid bounds[6] f64)."""
    boxes48 = _f64(boxes48, 6)
    codes = np.empty(boxes48.shape[0], dtype=np.uint64)
    bounds = np.empty(6, dtype=np.float64)
    orc().orc_morton_codes_f64(_p(boxes48, F64P), boxes48.shape[0], _p(bounds, F64P),
                               _p(codes, U64P))
    return codes, bounds


def karras_f64(sorted_codes):
    sc = np.ascontiguousarray(sorted_codes, dtype=np.uint64)
    nodes = np.zeros(sc.shape[0] - 1, dtype=NODE64_DTYPE)
    if orc().orc_karras_f64(_p(sc, U64P), sc.shape[0], nodes.ctypes.data) != 0:
        raise ValueError("orc_karras_f64 needs n >= 2")
    return nodes


def fit_f64(nodes, boxes48):
    nodes = np.array(nodes, dtype=NODE64_DTYPE, copy=True)
    boxes48 = _f64(boxes48, 6)
    orc().orc_fit_f64(nodes.ctypes.data, boxes48.shape[0], _p(boxes48, F64P))
    return nodes


def build_f64(boxes48, leaf_order=LEAF_REFERENCE):
    """builder<double> over converter boxes [n, 6].  Returns (nodes, sorted_codes, ids)."""
    boxes48 = _f64(boxes48, 6)
    n = boxes48.shape[0]
    nodes = np.zeros(max(n - 1, 0), dtype=NODE64_DTYPE)
    sc = np.empty(n, dtype=np.uint64)
    ids = np.empty(n, dtype=np.uint32)
    rc = orc().orc_build_f64(_p(boxes48, F64P), n, leaf_order, nodes.ctypes.data, _p(sc, U64P),
                             _p(ids, U32P))
    if rc != 0:
        raise ValueError("orc_build_f64 failed (n < 2)")
    return nodes, sc, ids


def trace_f64(nodes64, tri96, rays64, mode=ISECT_REFERENCE, threads=None):
    """traverser<double> over a node<double> array: triangle<double> records
    [n, 12] and ray<double> records [m, 8].  Returns (prim u64, t f64,
    uv f64[m, 2]); prim is 2^64 - 1 on a miss."""
    tri96 = _f64(tri96, 12)
    rays64 = _f64(rays64, 8)
    nodes64 = np.ascontiguousarray(nodes64, dtype=NODE64_DTYPE)
    m = rays64.shape[0]
    prim = np.empty(m, dtype=np.uint64)
    t = np.empty(m, dtype=np.float64)
    uv = np.empty((m, 2), dtype=np.float64)
    th = threads or min(8, os.cpu_count() or 1)
    orc().orc_trace_f64(nodes64.ctypes.data, tri96.shape[0], _p(tri96, F64P), _p(rays64, F64P), m,
                        mode, th, _p(prim, U64P), _p(t, F64P), _p(uv, F64P))
    return prim, t, uv


def brute_closest_f64(tri96, rays64, threads=None):
    """FP64 closest hit over every triangle (lowest id on equal distance)."""
    tri96 = _f64(tri96, 12)
    rays64 = _f64(rays64, 8)
    m = rays64.shape[0]
    prim = np.empty(m, dtype=np.uint64)
    t = np.empty(m, dtype=np.float64)
    th = threads or min(8, os.cpu_count() or 1)
    orc().orc_brute_closest_f64(_p(tri96, F64P), tri96.shape[0], _p(rays64, F64P), m, th,
                                _p(prim, U64P), _p(t, F64P))
    return prim, t


def trace(nodes, tri, rays, mode=ISECT_REFERENCE, threads=None, counters=False):
    """Returns (prim u32, t f32, uv f32[m,2]) [, counters(popped, boxes, tests)]."""
    tri = _tri(tri)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    m = rays.shape[0]
    prim = np.empty(m, dtype=np.uint32)
    t = np.empty(m, dtype=np.float32)
    uv = np.empty((m, 2), dtype=np.float32)
    cnt = np.zeros(3, dtype=np.uint64)
    th = threads or min(8, os.cpu_count() or 1)
    orc().orc_trace(nodes.ctypes.data, tri.shape[0], _p(tri, F32P), _p(rays, F32P), m, mode, th,
                    _p(prim, U32P), _p(t, F32P), _p(uv, F32P), _p(cnt, U64P))
    if counters:
        return prim, t, uv, cnt
    return prim, t, uv


def brute_closest(tri, rays, mode=ISECT_REFERENCE_LOWID, threads=None):
    tri = _tri(tri)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    m = rays.shape[0]
    prim = np.empty(m, dtype=np.uint32)
    t = np.empty(m, dtype=np.float32)
    th = threads or min(8, os.cpu_count() or 1)
    orc().orc_brute_closest(_p(tri, F32P), tri.shape[0], _p(rays, F32P), m, mode, th,
                            _p(prim, U32P), _p(t, F32P))
    return prim, t


def t_fp64(tri, rays, prim):
    tri = _tri(tri)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    prim = np.ascontiguousarray(prim, dtype=np.uint32)
    out = np.empty(rays.shape[0], dtype=np.float64)
    orc().orc_t_fp64(_p(tri, F32P), _p(rays, F32P), _p(prim, U32P), rays.shape[0], _p(out, F64P))
    return out


def fnv1a64(a):
    """FNV-1a 64-bit hash of an array's raw bytes (as a hex string)."""
    a = np.ascontiguousarray(a)
    return "%016x" % orc().orc_fnv1a64(a.ctypes.data, a.nbytes)


# ---- the real reference (compiled lbvh.h) ---------------------------------

def ref_build(tri, code_bits=30, leaf_order=LEAF_REFERENCE, threads=None, phases=False, fast=False):
    """Returns (nodes, sorted_codes, bounds[6]) [, phase_ms[4]]."""
    tri = _tri(tri)
    n = tri.shape[0]
    nodes = np.zeros(n - 1, dtype=NODE_DTYPE)
    sc = np.empty(n, dtype=_code_dtype(code_bits))
    bounds = np.empty(6, dtype=np.float32)
    ph = np.zeros(4, dtype=np.float64)
    th = threads or ref(fast).ref_hardware_concurrency() or 1
    rc = ref(fast).ref_build(_p(tri, F32P), n, code_bits, leaf_order, th, nodes.ctypes.data,
                             sc.ctypes.data, _p(bounds, F32P), _p(ph, F64P) if phases else None)
    if rc != 0:
        raise ValueError("ref_build failed")
    if phases:
        return nodes, sc, bounds, ph
    return nodes, sc, bounds


def ref_build_f64(prims, kind="triangles", leaf_order=LEAF_REFERENCE, threads=None, fast=False):
    """lbvh::builder<double> as shipped over triangle<double> [n, 12] or
    aabb<double> [n, 6] records.  Returns the node<double> array."""
    w = 12 if kind == "triangles" else 6
    prims = _f64(prims, w)
    n = prims.shape[0]
    nodes = np.zeros(n - 1, dtype=NODE64_DTYPE)
    th = threads or ref(fast).ref_hardware_concurrency() or 1
    L = ref(fast)
    fn = L.ref_build_f64_triangles if kind == "triangles" else L.ref_build_f64_aabbs
    if fn(_p(prims, F64P), n, leaf_order, th, nodes.ctypes.data) != 0:
        raise ValueError("ref_build_f64 failed")
    return nodes


def ref_trace_f64(nodes64, tri96, rays64, mode=0, threads=None):
    """The compiled reference's traverser<double> (mode 0 as shipped, 1 with
    lowest-id ties).  Returns (prim u64, t f64, uv f64[m, 2])."""
    tri96 = _f64(tri96, 12)
    rays64 = _f64(rays64, 8)
    nodes64 = np.ascontiguousarray(nodes64, dtype=NODE64_DTYPE)
    m = rays64.shape[0]
    prim = np.empty(m, dtype=np.uint64)
    t = np.empty(m, dtype=np.float64)
    uv = np.empty((m, 2), dtype=np.float64)
    th = threads or ref().ref_hardware_concurrency() or 1
    rc = ref().ref_trace_f64(nodes64.ctypes.data, tri96.shape[0], _p(tri96, F64P), _p(rays64, F64P),
                             m, mode, th, _p(prim, U64P), _p(t, F64P), _p(uv, F64P))
    if rc != 0:
        raise ValueError("ref_trace_f64 failed")
    return prim, t, uv


def ref_codes(tri, code_bits=30, threads=None):
    tri = _tri(tri)
    out = np.empty(tri.shape[0], dtype=_code_dtype(code_bits))
    th = threads or ref().ref_hardware_concurrency() or 1
    rc = ref().ref_codes(_p(tri, F32P), tri.shape[0], code_bits, th, out.ctypes.data)
    if rc != 0:
        raise ValueError("ref_codes failed")
    return out


def ref_trace(nodes, tri, rays, mode=0, threads=None, fast=False):
    tri = _tri(tri)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    m = rays.shape[0]
    prim = np.empty(m, dtype=np.uint32)
    t = np.empty(m, dtype=np.float32)
    uv = np.empty((m, 2), dtype=np.float32)
    th = threads or ref(fast).ref_hardware_concurrency() or 1
    rc = ref(fast).ref_trace(nodes.ctypes.data, tri.shape[0], _p(tri, F32P), _p(rays, F32P), m,
                             mode, th, _p(prim, U32P), _p(t, F32P), _p(uv, F32P))
    if rc != 0:
        raise ValueError("ref_trace failed")
    return prim, t, uv
