"""Build libwcfe.so (the gfx950 HIP library behind include/wcfe.h) in-tree.

    python -m wholecellfe_amd.build            # compile every source
    python -m wholecellfe_amd.build --if-stale # only when a source is newer

hipcc cross-compiles gfx950 code objects without a GPU present.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwcfe.so")
REPO = os.path.dirname(HERE)
COMPAT_SRC = os.path.join(REPO, "tests", "compat_check.cpp")
COMPAT_BIN = os.path.join(BUILD, "compat_check")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["k_assemble.hip", "k_lu.hip", "k_csr.hip", "k_reference.hip", "k_boundary.hip", "wcfe_vtk.hip",
           "wcfe_api.hip", "wcfe_noalloc.hip"]
HEADERS = ["wcfe_device.hpp", "wcfe_internal.hpp"]

FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _deps():
    inc = os.path.join(os.path.dirname(HERE), "include", "wcfe.h")
    return [os.path.join(CSRC, h) for h in HEADERS] + [inc]


def stale():
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    files = [os.path.join(CSRC, s) for s in SOURCES] + _deps() + [os.path.abspath(__file__)]
    return any(os.path.getmtime(f) > t for f in files)


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, r.stdout, r.stderr))
    return obj


def build(force=True, verbose=False):
    if not force and not stale():
        return LIB
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(6, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    # host-only check program for include/wcfe_compat.hpp (links libwcfe.so)
    if os.path.exists(COMPAT_SRC):
        cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(REPO, "include"), COMPAT_SRC, "-o", COMPAT_BIN,
               "-L" + HERE, "-lwcfe", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compat_check build failed:\n%s\n%s" % (r.stdout, r.stderr))
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--if-stale", action="store_true")
    a = ap.parse_args()
    build(force=not a.if_stale, verbose=True)
    sys.exit(0)
