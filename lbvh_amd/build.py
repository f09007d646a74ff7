"""Build recipe for the engine's native libraries (always rebuilds).

  lbvh_amd/lib/liblbvh_mi355x.so : the C-ABI engine (HIP kernels for gfx950 +
                                   host runtime), declared in include/lbvh_mi355x.h
  lbvh_amd/lib/libworkload.so    : synthetic scene / ray generators
  lbvh_amd/lib/dropin_minimal    : compile-conformance program for the C++
                                   drop-in header include/lbvh_mi355x.hpp

Floating-point flags are part of the contract: results must match the
reference CPU path bit for bit, so device and host code are compiled with
contraction off, IEEE-correct division and square root and no denormal
flushing.  Packed-FP32 instructions are disabled on the device.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")

ENGINE_SO = os.path.join(LIB, "liblbvh_mi355x.so")
WORKLOAD_SO = os.path.join(LIB, "libworkload.so")
DROPIN_BIN = os.path.join(LIB, "dropin_minimal")

HIP_SOURCES = ["capi.cpp", "batch_api.cpp", "serve.cpp", "capi_f64.cpp", "capi_noalloc.cpp", "build_kernels.hip", "build_f64.hip", "radix_sort.hip", "trace.hip", "trace_f64.hip", "batch.hip"]

ARCH = os.environ.get("LBVH_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    "--offload-arch=" + ARCH,
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    print("[build] " + " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_workload():
    os.makedirs(LIB, exist_ok=True)
    tmp = WORKLOAD_SO + ".tmp"
    _run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
          os.path.join(CSRC, "workload.c"), "-o", tmp, "-lm"])
    os.replace(tmp, WORKLOAD_SO)


def build_engine():
    os.makedirs(LIB, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(LIB, "_obj_" + src.replace(".", "_") + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        _run(["hipcc"] + lang + HIP_FLAGS + ["-I" + INCLUDE, "-I" + CSRC, "-c", path, "-o", obj])
        objs.append(obj)
    tmp = ENGINE_SO + ".tmp"
    _run(["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp, "-lpthread"])
    os.replace(tmp, ENGINE_SO)
    for o in objs:
        os.remove(o)


def build_dropin():
    src = os.path.join(ROOT, "tests", "cpp", "dropin_minimal.cpp")
    if not os.path.exists(src):
        return
    tmp = DROPIN_BIN + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I" + INCLUDE, src, "-o", tmp,
          "-L" + LIB, "-llbvh_mi355x", "-Wl,-rpath,$ORIGIN", "-pthread"])
    os.replace(tmp, DROPIN_BIN)


def build_all():
    build_workload()
    build_engine()
    build_dropin()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "workload": build_workload, "engine": build_engine,
         "dropin": build_dropin}[w]()
