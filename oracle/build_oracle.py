"""ORACLE build recipe — test infrastructure only.

Builds
  * oracle/_build/liblbvh_oracle.so : the clean-room C restatement
    (lbvh_oracle.c), strict IEEE flags;
  * oracle/_ref/libreflbvh.so       : ref_driver.cpp compiled against the real
    tay10r/lbvh header where it lies (/root/reference/lbvh.h), strict flags
    (-O2 -ffp-contract=off), the parity build;
  * oracle/_ref/libreflbvh_fast.so  : the same driver with the reference's
    release flags (Makefile LBVH_RELEASE: -O3 -ffast-math; -march is pinned to
    x86-64-v3 instead of native so the library runs on any current x86 host),
    used only for CPU-baseline timing of the build, never for traversal parity.

The reference build happens only where /root/reference exists; elsewhere the
previously built files are used as they are.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = "/root/reference"
BUILD_DIR = os.path.join(HERE, "_build")
REF_OUT = os.path.join(HERE, "_ref")

ORACLE_SO = os.path.join(BUILD_DIR, "liblbvh_oracle.so")
REF_SO = os.path.join(REF_OUT, "libreflbvh.so")
REF_FAST_SO = os.path.join(REF_OUT, "libreflbvh_fast.so")

STRICT = ["-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-pthread"]


def _run(cmd):
    print("[oracle] " + " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_oracle():
    os.makedirs(BUILD_DIR, exist_ok=True)
    tmp = ORACLE_SO + ".tmp"
    _run(["gcc", "-std=c11"] + STRICT + [os.path.join(HERE, "lbvh_oracle.c"), "-o", tmp, "-lm"])
    os.replace(tmp, ORACLE_SO)
    return ORACLE_SO


def build_ref():
    if not os.path.exists(os.path.join(REF_DIR, "lbvh.h")):
        print("[oracle] reference sources absent; keeping prebuilt oracle/_ref", flush=True)
        return None
    os.makedirs(REF_OUT, exist_ok=True)
    src = os.path.join(HERE, "ref_driver.cpp")
    tmp = REF_SO + ".tmp"
    _run(["g++", "-std=c++17", "-I" + REF_DIR] + STRICT + [src, "-o", tmp])
    os.replace(tmp, REF_SO)
    tmp = REF_FAST_SO + ".tmp"
    _run(["g++", "-std=c++17", "-I" + REF_DIR, "-O3", "-ffast-math", "-march=x86-64-v3",
          "-fPIC", "-shared", "-pthread", src, "-o", tmp])
    os.replace(tmp, REF_FAST_SO)
    return REF_SO


def build_all():
    build_oracle()
    build_ref()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
