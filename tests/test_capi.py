"""CPU tests of the C-ABI library: it loads, exports every function that
include/*.h declares, and its host-only entry points behave (no GPU work)."""
import ctypes
import os
import re
import subprocess

import pytest

import lbvh_amd
from lbvh_amd import build as B
from conftest import ROOT

HEADER_DIR = os.path.join(ROOT, "include")


def declared_functions():
    names = []
    for fn in sorted(os.listdir(HEADER_DIR)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(HEADER_DIR, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(lbvh_\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


needs_lib = pytest.mark.skipif(not os.path.exists(B.ENGINE_SO), reason="engine library not built")


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("lbvh_ctx_create", "lbvh_build_triangles", "lbvh_build_triangles_device",
                     "lbvh_build_aabbs", "lbvh_refit_triangles", "lbvh_bvh_download_nodes",
                     "lbvh_trace_closest", "lbvh_trace_closest_device", "lbvh_bvh_destroy",
                     "lbvh_status_string"):
        assert required in names


@needs_lib
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(B.ENGINE_SO)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", B.ENGINE_SO], capture_output=True,
                         text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    assert set(declared_functions()) <= exported


@needs_lib
def test_python_binding_covers_header():
    assert set(declared_functions()) == set(lbvh_amd.SIGNATURES)


@needs_lib
def test_host_only_entry_points():
    lib = lbvh_amd.lib()
    assert lib.lbvh_status_string(0) == b"ok"
    assert lib.lbvh_status_string(6) == b"bounded wait expired"
    o = lbvh_amd.BuildOpts()
    o.code_bits = 0
    lib.lbvh_build_opts_init(ctypes.byref(o))
    assert o.code_bits == 30 and o.leaf_order == 0 and o.keep_codes == 0
    # null handles are rejected without touching a device
    assert lib.lbvh_ctx_synchronize(None) == 1
    assert lib.lbvh_ctx_create(0, None) == 1
    assert lib.lbvh_ctx_set_autotune(None, 1) == 1
    assert lib.lbvh_ctx_tune_log(None, None, 0) == 0
    assert lib.lbvh_batch_get_info(None, None) == 1
    assert lib.lbvh_server_submit(None, None, None, None) == 1
    assert lib.lbvh_server_get_stats(None, None) == 1
    assert lib.lbvh_trace_closest_f64(None, None, 2, None, None, 0, 0, None, None, None) == 1
    assert lib.lbvh_trace_closest_f64_device(None, None, 2, None, None, 0, 0, None, None,
                                             None) == 1
    # server defaults: 288 GiB cap, 4 lanes, 64 in flight, 63-bit codes,
    # primitive-id leaves, reference intersector
    so = lbvh_amd.ServerOpts()
    lib.lbvh_server_opts_init(ctypes.byref(so))
    assert so.memory_cap_bytes == 288 << 30 and so.lanes == 4 and so.max_inflight == 64
    assert so.code_bits == 63 and so.leaf_order == 1 and so.isect_mode == 0
    # bad server options are rejected before any device is touched
    so.lanes = 0
    out = ctypes.c_void_p()
    assert lib.lbvh_server_create(0, ctypes.byref(so), ctypes.byref(out)) == 1
    so.lanes = 4
    so.code_bits = 31
    assert lib.lbvh_server_create(0, ctypes.byref(so), ctypes.byref(out)) == 1


@needs_lib
def test_struct_layouts_match_the_header():
    """The ctypes mirrors have the C structs' sizes (x86-64 ABI)."""
    assert ctypes.sizeof(lbvh_amd.BuildOpts) == 32
    assert ctypes.sizeof(lbvh_amd.BvhInfo) == 64
    assert ctypes.sizeof(lbvh_amd.BatchInfo) == 72
    assert ctypes.sizeof(lbvh_amd.TuneRecord) == 52
    assert ctypes.sizeof(lbvh_amd.ServerOpts) == 48
    assert ctypes.sizeof(lbvh_amd.Request) == 72
    assert ctypes.sizeof(lbvh_amd.RequestInfo) == 48
    assert ctypes.sizeof(lbvh_amd.ServerStats) == 88


def test_engine_fails_loudly_without_library(monkeypatch):
    monkeypatch.setattr(lbvh_amd, "_lib", None)
    monkeypatch.setattr(lbvh_amd, "library_path", lambda: "/nonexistent/liblbvh_mi355x.so")
    with pytest.raises(RuntimeError):
        lbvh_amd.lib()


def test_product_does_not_use_the_oracle():
    """The engine package never imports or loads the oracle."""
    pkg = os.path.join(ROOT, "lbvh_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hip", ".hpp", ".c", ".h")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "import oracle" not in text and "from oracle" not in text, fn
                assert "lbvh_oracle" not in text and "libreflbvh" not in text, fn
