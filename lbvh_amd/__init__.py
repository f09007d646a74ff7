"""lbvh_amd — Python access to the MI355X LBVH engine through its C ABI.

The engine is the native library lbvh_amd/lib/liblbvh_mi355x.so (HIP kernels
for gfx950 behind include/lbvh_mi355x.h).  This module only marshals
arguments: numpy arrays go through the host-pointer entry points (copy in,
copy out); torch CUDA tensors go through the *_device entry points and stay
resident.  There is no CPU fallback: if the library is missing or no gfx950
device is present, calls raise.

Host layouts (the reference's own, lbvh.h):
  triangles float32[n, 12]  lbvh::triangle<float> {p0, e0, e1, normal}
  boxes     float32[n, 6]   lbvh::aabb<float>
  rays      float32[m, 8]   lbvh::ray<float> {pos, dir, tmin, tmax}
  nodes     NODE_DTYPE[n-1] lbvh::node<float> {box[6], left, right}
  boxes64   float64[n, 6]   lbvh::aabb<double> (builder<double> input)
  nodes64   NODE64_DTYPE[n-1] lbvh::node<double> {box[6], u64 left, u64 right}
"""
import ctypes
import os

import numpy as np

from . import build as _build

LEAF_REFERENCE = 0
LEAF_PRIMITIVE_ID = 1
ISECT_REFERENCE = 0
ISECT_REFERENCE_LOWID = 1
ISECT_WATERTIGHT = 2
LEAF_BIT = 0x80000000
MISS = 0xFFFFFFFF

NODE_DTYPE = np.dtype([("box", np.float32, (6,)), ("left", np.uint32), ("right", np.uint32)])
# lbvh::node<double> (builder<double>): leaf flag in bit 63
NODE64_DTYPE = np.dtype([("box", np.float64, (6,)), ("left", np.uint64), ("right", np.uint64)])
LEAF_BIT64 = 1 << 63

STATUS = {0: "LBVH_OK", 1: "LBVH_E_INVALID", 2: "LBVH_E_OOM", 3: "LBVH_E_HIP", 4: "LBVH_E_BUSY",
          5: "LBVH_E_NODEVICE", 6: "LBVH_E_TIMEOUT"}


class LbvhError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


class BuildOpts(ctypes.Structure):
    _fields_ = [("code_bits", ctypes.c_uint32), ("leaf_order", ctypes.c_uint32),
                ("sort_digit_bits", ctypes.c_uint32), ("sort_tile", ctypes.c_uint32),
                ("keep_codes", ctypes.c_uint32), ("sort_strategy", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 2)]


class BvhInfo(ctypes.Structure):
    _fields_ = [("n_prims", ctypes.c_uint64), ("n_nodes", ctypes.c_uint64),
                ("code_bits", ctypes.c_uint32), ("leaf_order", ctypes.c_uint32),
                ("d_nodes", ctypes.c_void_p), ("d_prims", ctypes.c_void_p),
                ("device_bytes", ctypes.c_uint64), ("sort_digit_bits", ctypes.c_uint32),
                ("sort_tile", ctypes.c_uint32), ("sort_strategy", ctypes.c_uint32),
                ("sort_global_digits", ctypes.c_uint32)]


class TuneRecord(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_uint32), ("key_bits", ctypes.c_uint32),
                ("log2_n", ctypes.c_uint32), ("sort_digit_bits", ctypes.c_uint32),
                ("sort_tile", ctypes.c_uint32), ("sort_strategy", ctypes.c_uint32),
                ("picked", ctypes.c_uint32), ("ms", ctypes.c_float),
                ("scalar_bits", ctypes.c_uint32), ("kind", ctypes.c_uint32),
                ("trace_reorder", ctypes.c_uint32), ("trace_waves_per_simd", ctypes.c_uint32),
                ("trace_mode", ctypes.c_uint32)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("n_scenes", ctypes.c_uint32), ("leaf_order", ctypes.c_uint32),
                ("n_prims", ctypes.c_uint64), ("n_nodes", ctypes.c_uint64),
                ("sort_digit_bits", ctypes.c_uint32), ("sort_tile", ctypes.c_uint32),
                ("sort_passes", ctypes.c_uint32), ("sort_strategy", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64), ("d_nodes", ctypes.c_void_p),
                ("d_prims", ctypes.c_void_p), ("sort_global_passes", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("memory_cap_bytes", ctypes.c_uint64), ("lanes", ctypes.c_uint32),
                ("max_inflight", ctypes.c_uint32), ("code_bits", ctypes.c_uint32),
                ("leaf_order", ctypes.c_uint32), ("isect_mode", ctypes.c_uint32),
                ("reserve_prims", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 4)]


class Request(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_uint64), ("tri48", ctypes.c_void_p),
                ("n_tri", ctypes.c_uint64), ("rays32", ctypes.c_void_p),
                ("n_rays", ctypes.c_uint64), ("prim", ctypes.c_void_p), ("t", ctypes.c_void_p),
                ("uv", ctypes.c_void_p), ("flags", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


REQUEST_DEVICE = 1


class RequestInfo(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_uint64), ("status", ctypes.c_int),
                ("coalesced_with", ctypes.c_uint32), ("submit_us", ctypes.c_double),
                ("admit_us", ctypes.c_double), ("ready_us", ctypes.c_double),
                ("done_us", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in (
        "completed", "failed", "launches", "coalesced_requests", "max_coalesced", "rays",
        "evictions", "resident_scenes", "memory_cap_bytes", "memory_used_bytes",
        "memory_peak_bytes")]


VP = ctypes.c_void_p
U64 = ctypes.c_uint64
F32P = ctypes.POINTER(ctypes.c_float)

# Every symbol include/lbvh_mi355x.h declares, with its ctypes signature.
SIGNATURES = {
    "lbvh_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "lbvh_build_opts_init": (None, [ctypes.POINTER(BuildOpts)]),
    "lbvh_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(VP)]),
    "lbvh_ctx_destroy": (None, [VP]),
    "lbvh_ctx_set_stream": (ctypes.c_int, [VP, VP]),
    "lbvh_ctx_synchronize": (ctypes.c_int, [VP]),
    "lbvh_ctx_set_timing": (ctypes.c_int, [VP, ctypes.c_int]),
    "lbvh_ctx_phase_ms": (ctypes.c_int, [VP, F32P, ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_uint32)]),
    "lbvh_ctx_phase_stats": (ctypes.c_int, [VP, ctypes.c_int, F32P, ctypes.c_uint32,
                                            ctypes.POINTER(ctypes.c_uint32),
                                            ctypes.POINTER(ctypes.c_uint32)]),
    "lbvh_ctx_last_error": (ctypes.c_char_p, [VP]),
    "lbvh_ctx_set_autotune": (ctypes.c_int, [VP, ctypes.c_int]),
    "lbvh_ctx_tune_log": (ctypes.c_uint32, [VP, VP, ctypes.c_uint32]),
    "lbvh_ctx_workspace_bytes": (ctypes.c_uint64, [VP]),
    "lbvh_build_triangles": (ctypes.c_int, [VP, VP, U64, ctypes.POINTER(BuildOpts), ctypes.POINTER(VP)]),
    "lbvh_build_triangles_device": (ctypes.c_int, [VP, VP, U64, ctypes.POINTER(BuildOpts),
                                                   ctypes.POINTER(VP)]),
    "lbvh_build_aabbs": (ctypes.c_int, [VP, VP, U64, ctypes.POINTER(BuildOpts), ctypes.POINTER(VP)]),
    "lbvh_build_aabbs_f64": (ctypes.c_int, [VP, VP, U64, ctypes.POINTER(BuildOpts), VP]),
    "lbvh_build_aabbs_f64_device": (ctypes.c_int, [VP, VP, U64, ctypes.POINTER(BuildOpts), VP]),
    "lbvh_trace_closest_f64": (ctypes.c_int, [VP, VP, U64, VP, VP, U64, ctypes.c_int, VP, VP, VP]),
    "lbvh_trace_closest_f64_device": (ctypes.c_int, [VP, VP, U64, VP, VP, U64, ctypes.c_int, VP,
                                                     VP, VP]),
    "lbvh_build_aabbs_device": (ctypes.c_int, [VP, VP, U64, ctypes.POINTER(BuildOpts),
                                               ctypes.POINTER(VP)]),
    "lbvh_build_triangles_noalloc": (ctypes.c_int, [VP, VP, U64, ctypes.POINTER(BuildOpts), VP]),
    "lbvh_refit_triangles": (ctypes.c_int, [VP, VP, VP]),
    "lbvh_refit_triangles_device": (ctypes.c_int, [VP, VP, VP]),
    "lbvh_bvh_get_info": (ctypes.c_int, [VP, ctypes.POINTER(BvhInfo)]),
    "lbvh_bvh_download_nodes": (ctypes.c_int, [VP, VP, VP, U64]),
    "lbvh_bvh_download_codes": (ctypes.c_int, [VP, VP, VP, VP, VP]),
    "lbvh_bvh_destroy": (None, [VP]),
    "lbvh_trace_closest": (ctypes.c_int, [VP, VP, VP, U64, ctypes.c_int, VP, VP, VP]),
    "lbvh_trace_closest_device": (ctypes.c_int, [VP, VP, VP, U64, ctypes.c_int, VP, VP, VP]),
    "lbvh_build_batch": (ctypes.c_int, [VP, VP, VP, ctypes.c_uint32, ctypes.POINTER(BuildOpts),
                                        ctypes.POINTER(VP)]),
    "lbvh_build_batch_device": (ctypes.c_int, [VP, VP, VP, ctypes.c_uint32,
                                               ctypes.POINTER(BuildOpts), ctypes.POINTER(VP)]),
    "lbvh_batch_rebuild": (ctypes.c_int, [VP, VP, VP, ctypes.c_int]),
    "lbvh_batch_refit": (ctypes.c_int, [VP, VP, VP, ctypes.c_int]),
    "lbvh_batch_get_info": (ctypes.c_int, [VP, ctypes.POINTER(BatchInfo)]),
    "lbvh_batch_download_scene": (ctypes.c_int, [VP, VP, ctypes.c_uint32, VP, U64, VP, VP, VP]),
    "lbvh_batch_destroy": (None, [VP]),
    "lbvh_batch_trace_closest": (ctypes.c_int, [VP, VP, VP, VP, ctypes.c_int, VP, VP, VP]),
    "lbvh_batch_trace_closest_device": (ctypes.c_int, [VP, VP, VP, VP, U64, ctypes.c_int, VP,
                                                       VP, VP]),
    "lbvh_server_opts_init": (None, [ctypes.POINTER(ServerOpts)]),
    "lbvh_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ServerOpts),
                                          ctypes.POINTER(VP)]),
    "lbvh_server_destroy": (None, [VP]),
    "lbvh_server_submit": (ctypes.c_int, [VP, ctypes.POINTER(Request), ctypes.POINTER(U64),
                                          ctypes.POINTER(U64)]),
    "lbvh_server_wait": (ctypes.c_int, [VP, U64, ctypes.POINTER(RequestInfo)]),
    "lbvh_server_poll": (ctypes.c_int, [VP, U64, ctypes.POINTER(RequestInfo)]),
    "lbvh_server_release_scene": (ctypes.c_int, [VP, U64]),
    "lbvh_server_get_stats": (ctypes.c_int, [VP, ctypes.POINTER(ServerStats)]),
}

_lib = None


def library_path():
    return _build.ENGINE_SO


def lib():
    """Loads the engine library; raises if it is not built."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError("the HIP engine library is not built (%s); run __graft_entry__.build()"
                               % path)
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _single_runtime():
    """True when the process maps one copy of the HIP runtime."""
    try:
        with open("/proc/self/maps") as f:
            paths = set(line.split()[-1] for line in f if "libamdhip64" in line)
    except OSError:
        return True
    return len(paths) <= 1


def _is_device(x):
    from .hip import DeviceArray
    if isinstance(x, DeviceArray):
        return True
    if _is_torch(x):
        if not x.is_cuda:
            return False
        if not _single_runtime():
            raise RuntimeError("PyTorch and the engine use different HIP runtimes in this "
                               "process; import torch before the engine is first used")
        if not x.is_contiguous():
            raise ValueError("device arrays must be contiguous")
        return True
    return False


def _numel(x):
    return int(x.numel())


def _dev_dtype(x):
    """The numpy dtype of a device array (DeviceArray or torch tensor)."""
    if _is_torch(x):
        try:
            return np.dtype(str(x.dtype).replace("torch.", ""))
        except TypeError:
            return None
    return np.dtype(x.dtype)


def _dev_nbytes(x):
    if _is_torch(x):
        return int(x.numel()) * int(x.element_size())
    return int(x.nbytes)


def _check_dev(x, what, dtypes, count=None, at_least=None, width=1):
    """Rejects a device array whose dtype is not one of `dtypes`, whose
    element count is not a multiple of `width`, or that holds other than
    `count` (or fewer than `at_least`) elements, before any C call: the
    engine reads and writes device memory by the counts it is given."""
    dt = _dev_dtype(x)
    if dt is None or all(dt != np.dtype(d) for d in dtypes):
        raise ValueError("%s: device dtype %s, expected %s"
                         % (what, dt, " or ".join(np.dtype(d).name for d in dtypes)))
    n = _numel(x)
    if n % width:
        raise ValueError("%s: %d elements is not a multiple of %d" % (what, n, width))
    if count is not None and n != count:
        raise ValueError("%s: %d elements, expected %d" % (what, n, count))
    if at_least is not None and n < at_least:
        raise ValueError("%s: %d elements, expected at least %d" % (what, n, at_least))
    return n // width


def _check_dev_out(out, m, with_uv):
    """Device outputs (prim, t, uv) of m rays: u32 / f32 / 2 x f32 each."""
    prim, t, uv = out
    _check_dev(prim, "prim", (np.uint32, np.int32), at_least=m)
    _check_dev(t, "t", (np.float32,), at_least=m)
    if uv is not None:
        _check_dev(uv, "uv", (np.float32,), at_least=2 * m)
    return prim, t, uv


def _dptr(x):
    return ctypes.c_void_p(int(x.data_ptr()))


def _hptr(a):
    return ctypes.c_void_p(a.ctypes.data)


SORT_AUTO, SORT_ONESWEEP, SORT_REDUCE_SCAN, SORT_HYBRID = 0, 1, 2, 3
SORT_NAMES = {1: "onesweep", 2: "reduce-then-scan", 3: "hybrid"}


def _make_opts(code_bits, leaf_order, keep_codes, sort_digit_bits=0, sort_tile=0, sort_strategy=0):
    o = BuildOpts()
    lib().lbvh_build_opts_init(ctypes.byref(o))
    o.code_bits = code_bits
    o.leaf_order = leaf_order
    o.keep_codes = 1 if keep_codes else 0
    o.sort_digit_bits = sort_digit_bits
    o.sort_tile = sort_tile
    o.sort_strategy = sort_strategy
    return o


class Bvh:
    """A device-resident tree (lbvh::bvh<float> equivalent)."""

    def __init__(self, ctx, handle, n, code_bits, leaf_order, keep=None):
        self.ctx = ctx
        self.handle = handle
        self.n = n
        self.code_bits = code_bits
        self.leaf_order = leaf_order
        self._keep = keep  # borrowed device tensors that must outlive the tree

    def info(self):
        inf = BvhInfo()
        self.ctx._check(lib().lbvh_bvh_get_info(self.handle, ctypes.byref(inf)))
        return inf

    def nodes(self):
        out = np.empty(self.n - 1, dtype=NODE_DTYPE)
        self.ctx._check(lib().lbvh_bvh_download_nodes(self.ctx.handle, self.handle, _hptr(out),
                                                      out.nbytes))
        return out

    def codes(self):
        """(sorted codes, sorted ids, centroid bounds[6]); needs keep_codes."""
        codes = np.empty(self.n, dtype=np.uint32 if self.code_bits == 30 else np.uint64)
        ids = np.empty(self.n, dtype=np.uint32)
        bounds = np.empty(6, dtype=np.float32)
        self.ctx._check(lib().lbvh_bvh_download_codes(self.ctx.handle, self.handle, _hptr(codes),
                                                      _hptr(ids), _hptr(bounds)))
        return codes, ids, bounds

    def bounds(self):
        bounds = np.empty(6, dtype=np.float32)
        self.ctx._check(lib().lbvh_bvh_download_codes(self.ctx.handle, self.handle, None, None,
                                                      _hptr(bounds)))
        return bounds

    def destroy(self):
        if self.handle:
            lib().lbvh_bvh_destroy(self.handle)
            self.handle = None
            self._keep = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """Many scenes built in one segmented pass (lbvh_batch).  Scene s is
    primitives [offsets[s], offsets[s + 1]); ids, nodes and hits are
    scene-local."""

    def __init__(self, ctx, handle, offsets, leaf_order, keep=None):
        self.ctx = ctx
        self.handle = handle
        self.offsets = np.asarray(offsets, dtype=np.uint64)
        self.n_scenes = len(self.offsets) - 1
        self.n = int(self.offsets[-1])
        self.leaf_order = leaf_order
        self._keep = keep

    def info(self):
        inf = BatchInfo()
        self.ctx._check(lib().lbvh_batch_get_info(self.handle, ctypes.byref(inf)))
        return inf

    def scene_size(self, s):
        return int(self.offsets[s + 1] - self.offsets[s])

    def nodes(self, s):
        out = np.empty(self.scene_size(s) - 1, dtype=NODE_DTYPE)
        self.ctx._check(lib().lbvh_batch_download_scene(self.ctx.handle, self.handle, s,
                                                        _hptr(out), out.nbytes, None, None, None))
        return out

    def codes(self, s, with_ids=True):
        """(sorted 30-bit codes, scene-local ids or None, centroid bounds[6])."""
        m = self.scene_size(s)
        codes = np.empty(m, dtype=np.uint32)
        ids = np.empty(m, dtype=np.uint32) if with_ids else None
        bounds = np.empty(6, dtype=np.float32)
        self.ctx._check(lib().lbvh_batch_download_scene(
            self.ctx.handle, self.handle, s, None, 0, _hptr(codes),
            _hptr(ids) if ids is not None else None, _hptr(bounds)))
        return codes, ids, bounds

    def rebuild(self, tri=None, graph=False):
        """Queues a rebuild of every scene (from `tri` when given: device memory)."""
        p = None
        if tri is not None:
            if not _is_device(tri):
                raise ValueError("rebuild takes device triangles")
            _check_dev(tri, "triangles", (np.float32,), count=12 * self.n)
            p = _dptr(tri)
            self._keep = tri
        self.ctx._check(lib().lbvh_batch_rebuild(self.ctx.handle, self.handle, p, int(graph)))

    def refit(self, tri=None, graph=False):
        """Queues a refit of every scene on its unchanged topology."""
        p = None
        if tri is not None:
            if not _is_device(tri):
                raise ValueError("refit takes device triangles")
            _check_dev(tri, "triangles", (np.float32,), count=12 * self.n)
            p = _dptr(tri)
            self._keep = tri
        self.ctx._check(lib().lbvh_batch_refit(self.ctx.handle, self.handle, p, int(graph)))

    def trace(self, rays, ray_offsets, mode=ISECT_REFERENCE, out=None, with_uv=True):
        """Closest hits; rays [ray_offsets[s], ray_offsets[s + 1]) go to scene
        s.  numpy rays and offsets -> numpy outputs after the work finished;
        device rays and device offsets -> device outputs, queued."""
        if _is_device(rays):
            if not _is_device(ray_offsets):
                raise ValueError("device rays need device ray offsets")
            m = _check_dev(rays, "rays", (np.float32,), width=8)
            _check_dev(ray_offsets, "ray offsets", (np.uint64, np.int64),
                       count=self.n_scenes + 1)
            if out is None:
                from .hip import DeviceArray
                prim = DeviceArray(m, np.uint32)
                t = DeviceArray(m, np.float32)
                uv = DeviceArray((m, 2), np.float32) if with_uv else None
            else:
                prim, t, uv = _check_dev_out(out, m, with_uv)
            self.ctx._check(lib().lbvh_batch_trace_closest_device(
                self.ctx.handle, self.handle, _dptr(rays), _dptr(ray_offsets), m, mode,
                _dptr(prim), _dptr(t), _dptr(uv) if uv is not None else None))
            return prim, t, uv
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        off = np.ascontiguousarray(ray_offsets, dtype=np.uint64)
        if off.shape[0] != self.n_scenes + 1 or int(off[-1]) != r.shape[0]:
            raise ValueError("ray offsets must have n_scenes + 1 entries ending at the ray count")
        m = r.shape[0]
        prim = np.empty(m, dtype=np.uint32)
        t = np.empty(m, dtype=np.float32)
        uv = np.empty((m, 2), dtype=np.float32) if with_uv else None
        self.ctx._check(lib().lbvh_batch_trace_closest(self.ctx.handle, self.handle, _hptr(r),
                                                       _hptr(off), mode, _hptr(prim), _hptr(t),
                                                       _hptr(uv) if uv is not None else None))
        return prim, t, uv

    def destroy(self):
        if self.handle:
            lib().lbvh_batch_destroy(self.handle)
            self.handle = None
            self._keep = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Server:
    """A request server with continuous batching (lbvh_server).  submit()
    returns at once; wait() / poll() collect a request's (prim, t, uv, info)."""

    def __init__(self, device=0, memory_cap_bytes=None, lanes=None, max_inflight=None,
                 code_bits=None, leaf_order=None, isect_mode=None, reserve_prims=None):
        o = ServerOpts()
        lib().lbvh_server_opts_init(ctypes.byref(o))
        for k, v in (("memory_cap_bytes", memory_cap_bytes), ("lanes", lanes),
                     ("max_inflight", max_inflight), ("code_bits", code_bits),
                     ("leaf_order", leaf_order), ("isect_mode", isect_mode),
                     ("reserve_prims", reserve_prims)):
            if v is not None:
                setattr(o, k, v)
        h = ctypes.c_void_p()
        st = lib().lbvh_server_create(device, ctypes.byref(o), ctypes.byref(h))
        if st != 0:
            raise LbvhError(st, "lbvh_server_create")
        self.handle = h
        self._live = {}  # ticket -> host arrays the request reads and writes

    def submit(self, rays, scene=0, tri=None, with_uv=True, out=None):
        """scene 0 with tri: a new scene; a scene id with tri: refit, then
        trace; a scene id alone: trace.  numpy rays: host request (results
        copied back); device rays (DeviceArray or torch CUDA tensor, with
        device triangles): device request, results in device outputs (`out`
        = (prim, t, uv) or new DeviceArrays).  Returns (ticket, scene id)."""
        req = Request()
        req.scene = scene
        keep = []
        if _is_device(rays):
            m = _check_dev(rays, "rays", (np.float32,), width=8)
            if tri is not None:
                if not _is_device(tri):
                    raise ValueError("a device request needs device triangles")
                _check_dev(tri, "triangles", (np.float32,), width=12)
            if out is None:
                from .hip import DeviceArray
                out = (DeviceArray(m, np.uint32), DeviceArray(m, np.float32),
                       DeviceArray((m, 2), np.float32) if with_uv else None)
            prim, t, uv = _check_dev_out(out, m, True)
            req.flags = REQUEST_DEVICE
            req.rays32 = _dptr(rays).value if m else None
            req.prim = _dptr(prim).value
            req.t = _dptr(t).value
            req.uv = _dptr(uv).value if uv is not None else None
            if tri is not None:
                req.tri48 = _dptr(tri).value
                req.n_tri = _numel(tri) // 12
            keep = [rays, tri]
        else:
            r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
            m = r.shape[0]
            prim = np.empty(m, dtype=np.uint32)
            t = np.empty(m, dtype=np.float32)
            uv = np.empty((m, 2), dtype=np.float32) if with_uv else None
            if tri is not None:
                tri_a = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 12)
                req.tri48 = tri_a.ctypes.data
                req.n_tri = tri_a.shape[0]
                keep.append(tri_a)
            req.rays32 = r.ctypes.data if m else None
            req.prim = prim.ctypes.data
            req.t = t.ctypes.data
            req.uv = uv.ctypes.data if uv is not None else None
            keep.append(r)
        req.n_rays = m
        ticket, sid = U64(), U64()
        st = lib().lbvh_server_submit(self.handle, ctypes.byref(req), ctypes.byref(ticket),
                                      ctypes.byref(sid))
        if st != 0:
            raise LbvhError(st, "lbvh_server_submit")
        self._live[ticket.value] = (keep, prim, t, uv)
        return ticket.value, sid.value

    def _collect(self, fn, ticket):
        info = RequestInfo()
        st = fn(self.handle, ticket, ctypes.byref(info))
        if st == 4 and fn is lib().lbvh_server_poll:
            return None
        _, prim, t, uv = self._live.pop(ticket)
        if st != 0:
            raise LbvhError(st, "request %d" % ticket)
        return prim, t, uv, {"scene": info.scene, "coalesced_with": info.coalesced_with,
                             "submit_us": info.submit_us, "admit_us": info.admit_us,
                             "ready_us": info.ready_us, "done_us": info.done_us}

    def wait(self, ticket):
        return self._collect(lib().lbvh_server_wait, ticket)

    def poll(self, ticket):
        """None while the request runs."""
        return self._collect(lib().lbvh_server_poll, ticket)

    def release_scene(self, scene):
        st = lib().lbvh_server_release_scene(self.handle, scene)
        if st != 0:
            raise LbvhError(st, "lbvh_server_release_scene")

    def stats(self):
        out = ServerStats()
        st = lib().lbvh_server_get_stats(self.handle, ctypes.byref(out))
        if st != 0:
            raise LbvhError(st, "lbvh_server_get_stats")
        return {k: int(getattr(out, k)) for k, _ in ServerStats._fields_}

    def close(self):
        if self.handle:
            lib().lbvh_server_destroy(self.handle)
            self.handle = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """An engine context on one gfx950 device (lbvh_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        st = lib().lbvh_ctx_create(device, ctypes.byref(h))
        if st != 0:
            raise LbvhError(st, lib().lbvh_status_string(st).decode())
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().lbvh_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            msg = lib().lbvh_ctx_last_error(self.handle).decode() if self.handle else ""
            raise LbvhError(st, msg)

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream, a raw hipStream_t int, or None (own stream)."""
        raw = None
        if stream is not None:
            raw = ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream)))
        self._check(lib().lbvh_ctx_set_stream(self.handle, raw))

    def synchronize(self):
        self._check(lib().lbvh_ctx_synchronize(self.handle))

    def set_timing(self, mode=1):
        """0 off, 1 (or True) per phase, 2 one interval per call."""
        self._check(lib().lbvh_ctx_set_timing(self.handle, int(mode)))

    def phase_ms(self):
        buf = (ctypes.c_float * 8)()
        cnt = ctypes.c_uint32()
        self._check(lib().lbvh_ctx_phase_ms(self.handle, buf, 8, ctypes.byref(cnt)))
        return [buf[i] for i in range(cnt.value)]

    def phase_stats(self, kind):
        """(summed phase ms, number of calls) over timed calls of a kind:
        1 build [bounds+encode, sort, tree, fit], 2 trace, 3 refit."""
        buf = (ctypes.c_float * 8)()
        cnt = ctypes.c_uint32()
        calls = ctypes.c_uint32()
        self._check(lib().lbvh_ctx_phase_stats(self.handle, kind, buf, 8, ctypes.byref(cnt),
                                               ctypes.byref(calls)))
        return [buf[i] for i in range(cnt.value)], calls.value

    def set_autotune(self, enable=True):
        """Sort autotuning for builds that leave the sort options at 0."""
        self._check(lib().lbvh_ctx_set_autotune(self.handle, int(bool(enable))))

    def tune_log(self):
        """Every configuration the sort autotuner timed: a list of dicts."""
        n = lib().lbvh_ctx_tune_log(self.handle, None, 0)
        buf = (TuneRecord * max(n, 1))()
        lib().lbvh_ctx_tune_log(self.handle, buf, n)
        return [{"kind": "traversal" if r.kind == 1 else "sort",
                 "batch": bool(r.batch), "key_bits": r.key_bits, "log2_n": r.log2_n,
                 "digit_bits": r.sort_digit_bits, "tile": r.sort_tile,
                 "strategy": SORT_NAMES.get(r.sort_strategy, "-"),
                 "picked": bool(r.picked), "ms": round(r.ms, 4),
                 "scalar_bits": r.scalar_bits, "reorder": bool(r.trace_reorder),
                 "waves_per_simd": r.trace_waves_per_simd, "mode": r.trace_mode}
                for r in buf[:n]]

    def workspace_bytes(self):
        return int(lib().lbvh_ctx_workspace_bytes(self.handle))

    # -- build --------------------------------------------------------------
    def build_triangles(self, tri, code_bits=30, leaf_order=LEAF_REFERENCE, keep_codes=False,
                        sort_digit_bits=0, sort_tile=0, sort_strategy=SORT_AUTO):
        return self._build(tri, 12, "lbvh_build_triangles", code_bits, leaf_order, keep_codes,
                           sort_digit_bits, sort_tile, sort_strategy)

    def build_aabbs(self, boxes, code_bits=30, leaf_order=LEAF_REFERENCE, keep_codes=False,
                    sort_digit_bits=0, sort_tile=0, sort_strategy=SORT_AUTO):
        return self._build(boxes, 6, "lbvh_build_aabbs", code_bits, leaf_order, keep_codes,
                           sort_digit_bits, sort_tile, sort_strategy)

    def _build(self, prims, width, fn, code_bits, leaf_order, keep_codes, sdb, stile, sstrat):
        opts = _make_opts(code_bits, leaf_order, keep_codes, sdb, stile, sstrat)
        out = ctypes.c_void_p()
        if _is_device(prims):
            n = _check_dev(prims, "primitives", (np.float32,), width=width)
            self._check(getattr(lib(), fn + "_device")(self.handle, _dptr(prims), n,
                                                       ctypes.byref(opts), ctypes.byref(out)))
            return Bvh(self, out, n, code_bits, leaf_order, keep=prims)
        a = np.ascontiguousarray(prims, dtype=np.float32).reshape(-1, width)
        n = a.shape[0]
        self._check(getattr(lib(), fn)(self.handle, _hptr(a), n, ctypes.byref(opts),
                                       ctypes.byref(out)))
        return Bvh(self, out, n, code_bits, leaf_order)

    def build_aabbs_f64(self, boxes, leaf_order=LEAF_REFERENCE, out=None, sort_digit_bits=0,
                        sort_tile=0, sort_strategy=SORT_AUTO):
        """lbvh::builder<double> over converter boxes (float64 [n, 6]).
        Host input: returns NODE64_DTYPE[n - 1].  Device input (a
        DeviceArray or a torch float64 tensor): writes the node records into
        `out` (device memory of at least 64 (n - 1) bytes; allocated like
        the input when None) and returns it without synchronizing."""
        opts = _make_opts(63, leaf_order, False, sort_digit_bits, sort_tile, sort_strategy)
        if _is_device(boxes):
            n = _check_dev(boxes, "boxes", (np.float64,), width=6)
            if out is None:
                if _is_torch(boxes):
                    import torch
                    out = torch.empty(max(n - 1, 1) * 8, dtype=torch.int64, device=boxes.device)
                else:
                    from .hip import DeviceArray
                    out = DeviceArray((max(n - 1, 1),), NODE64_DTYPE)
            nbytes = _dev_nbytes(out)
            if nbytes < 64 * max(n - 1, 0):
                raise ValueError("out holds fewer than 64 (n - 1) bytes")
            self._check(lib().lbvh_build_aabbs_f64_device(self.handle, _dptr(boxes), n,
                                                          ctypes.byref(opts), _dptr(out)))
            return out
        a = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        n = a.shape[0]
        nodes = np.zeros(max(n - 1, 1), dtype=NODE64_DTYPE)
        self._check(lib().lbvh_build_aabbs_f64(self.handle, _hptr(a), n, ctypes.byref(opts),
                                               _hptr(nodes)))
        return nodes[:n - 1]

    def trace_f64(self, nodes, tri, rays, mode=ISECT_REFERENCE, out=None, with_uv=True):
        """traverser<double> closest hits over the node<double> array of an
        FP64 build of the triangle<double> records `tri` ([n, 12] float64)
        for ray<double> records `rays` ([m, 8] float64).  Host arrays ->
        numpy (prim u64, t f64, uv f64 [m, 2]) after the work finished;
        device arrays (DeviceArray or torch) -> device outputs, queued.
        prim is 2^64 - 1 on a miss."""
        if _is_device(rays):
            if not (_is_device(nodes) and _is_device(tri)):
                raise ValueError("device rays need device nodes and triangles")
            n = _check_dev(tri, "triangles", (np.float64,), width=12)
            m = _check_dev(rays, "rays", (np.float64,), width=8)
            if _dev_nbytes(nodes) < 64 * max(n - 1, 0):
                raise ValueError("nodes hold fewer than 64 (n - 1) bytes")
            if out is None:
                from .hip import DeviceArray
                prim = DeviceArray(m, np.uint64)
                t = DeviceArray(m, np.float64)
                uv = DeviceArray((m, 2), np.float64) if with_uv else None
            else:
                prim, t, uv = out
                _check_dev(prim, "prim", (np.uint64, np.int64), at_least=m)
                _check_dev(t, "t", (np.float64,), at_least=m)
                if uv is not None:
                    _check_dev(uv, "uv", (np.float64,), at_least=2 * m)
            self._check(lib().lbvh_trace_closest_f64_device(
                self.handle, _dptr(nodes), n, _dptr(tri), _dptr(rays), m, mode, _dptr(prim),
                _dptr(t), _dptr(uv) if uv is not None else None))
            return prim, t, uv
        tr = np.ascontiguousarray(tri, dtype=np.float64).reshape(-1, 12)
        r = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 8)
        nd = np.ascontiguousarray(nodes, dtype=NODE64_DTYPE)
        n, m = tr.shape[0], r.shape[0]
        if nd.shape[0] != n - 1:
            raise ValueError("%d nodes for %d triangles, expected n - 1" % (nd.shape[0], n))
        prim = np.empty(m, dtype=np.uint64)
        t = np.empty(m, dtype=np.float64)
        uv = np.empty((m, 2), dtype=np.float64) if with_uv else None
        self._check(lib().lbvh_trace_closest_f64(self.handle, _hptr(nd), n, _hptr(tr), _hptr(r),
                                                 m, mode, _hptr(prim), _hptr(t),
                                                 _hptr(uv) if uv is not None else None))
        return prim, t, uv

    def build_batch(self, tri, offsets, leaf_order=LEAF_REFERENCE, keep_codes=False,
                    sort_digit_bits=0, sort_tile=0, sort_strategy=SORT_AUTO):
        """Builds every scene of `tri` (scene s = triangles offsets[s] ..
        offsets[s + 1] - 1) in one segmented pass with 30-bit codes."""
        opts = _make_opts(30, leaf_order, keep_codes, sort_digit_bits, sort_tile, sort_strategy)
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        nseg = off.shape[0] - 1
        out = ctypes.c_void_p()
        if _is_device(tri):
            _check_dev(tri, "triangles", (np.float32,), count=12 * int(off[-1]))
            self._check(lib().lbvh_build_batch_device(self.handle, _dptr(tri), _hptr(off), nseg,
                                                      ctypes.byref(opts), ctypes.byref(out)))
            return Batch(self, out, off, leaf_order, keep=tri)
        a = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 12)
        if a.shape[0] != int(off[-1]):
            raise ValueError("offsets must end at the triangle count")
        self._check(lib().lbvh_build_batch(self.handle, _hptr(a), _hptr(off), nseg,
                                           ctypes.byref(opts), ctypes.byref(out)))
        return Batch(self, out, off, leaf_order)

    def refit_triangles(self, bvh, tri):
        if _is_device(tri):
            _check_dev(tri, "triangles", (np.float32,), count=12 * bvh.n)
            self._check(lib().lbvh_refit_triangles_device(self.handle, bvh.handle, _dptr(tri)))
            bvh._keep = tri
            return
        a = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 12)
        if a.shape[0] != bvh.n:
            raise ValueError("refit needs the same primitive count")
        self._check(lib().lbvh_refit_triangles(self.handle, bvh.handle, _hptr(a)))

    # -- traversal ---------------------------------------------------------
    def trace(self, bvh, rays, mode=ISECT_REFERENCE, out=None, with_uv=True):
        """Closest hit per ray.  numpy rays -> numpy (prim, t, uv) after the
        work has finished; device rays (DeviceArray or a torch CUDA tensor) ->
        device outputs, queued on the context stream (synchronize before
        reading)."""
        if _is_device(rays):
            m = _check_dev(rays, "rays", (np.float32,), width=8)
            if out is None:
                from .hip import DeviceArray
                prim = DeviceArray(m, np.uint32)
                t = DeviceArray(m, np.float32)
                uv = DeviceArray((m, 2), np.float32) if with_uv else None
            else:
                prim, t, uv = _check_dev_out(out, m, with_uv)
            self._check(lib().lbvh_trace_closest_device(self.handle, bvh.handle, _dptr(rays), m,
                                                        mode, _dptr(prim), _dptr(t),
                                                        _dptr(uv) if uv is not None else None))
            return prim, t, uv
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        m = r.shape[0]
        prim = np.empty(m, dtype=np.uint32)
        t = np.empty(m, dtype=np.float32)
        uv = np.empty((m, 2), dtype=np.float32) if with_uv else None
        self._check(lib().lbvh_trace_closest(self.handle, bvh.handle, _hptr(r), m, mode,
                                             _hptr(prim), _hptr(t),
                                             _hptr(uv) if uv is not None else None))
        return prim, t, uv
