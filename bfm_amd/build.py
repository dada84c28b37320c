"""Builds the native libraries of bfm_amd in-tree.

    lib/libbfmhip.so   HIP kernels for gfx950 + the bfmx C ABI (include/bfmx.h)
    lib/libbfm.so.1    the libbfm C API (include/bfm/*.h), C99 host code that
                       calls libbfmhip.so for the planar-elasticity path
                       (lib/libbfm.so links to it)

Usage: python -m bfm_amd.build [--jobs N]
"""

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
HIP_DIR = os.path.join(HERE, "csrc", "hip")
HOST_DIR = os.path.join(HERE, "csrc", "host")
LIB_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "lib", "obj")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-I", INC, "-I", HIP_DIR,
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-fvisibility=hidden",
]
C_FLAGS = ["-O2", "-std=gnu99", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I", INC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r


def _sources(d, ext):
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(ext)) if os.path.isdir(d) else []


def public_api():
    """Function names declared by include/bfm/*.h."""
    import re
    names = set()
    hdir = os.path.join(INC, "bfm")
    for h in os.listdir(hdir):
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(hdir, h)).read(), flags=re.S)
        names |= set(re.findall(r"\b(bfm_\w+)\s*\(", text))
    return names


def build(jobs=8, verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    hip_src = _sources(HIP_DIR, ".hip")
    objs = []

    def compile_hip(src):
        obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_hip, hip_src))
    hip_lib = os.path.join(LIB_DIR, "libbfmhip.so")
    tmp = hip_lib + ".tmp"
    vscript = os.path.join(OBJ_DIR, "bfmhip.map")
    with open(vscript, "w") as f:
        f.write("{ global: bfmx_*; local: *; };\n")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-Wl,-soname,libbfmhip.so", f"-Wl,--version-script={vscript}"])
    os.replace(tmp, hip_lib)

    host_src = _sources(HOST_DIR, ".c")
    if host_src:
        hobjs = []
        for src in host_src:
            obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
            _run([CC, *C_FLAGS, "-c", src, "-o", obj])
            hobjs.append(obj)
        bfm_lib = os.path.join(LIB_DIR, "libbfm.so")
        tmp = bfm_lib + ".tmp"
        # export exactly the functions the public headers declare
        vscript = os.path.join(OBJ_DIR, "bfm.map")
        with open(vscript, "w") as f:
            f.write("{ global: " + " ".join(f"{n};" for n in sorted(public_api())) + " local: *; };\n")
        _run([CC, "-shared", "-o", tmp, *hobjs, "-L", LIB_DIR, "-lbfmhip", "-Wl,-rpath,$ORIGIN",
              "-Wl,-soname,libbfm.so.1", f"-Wl,--version-script={vscript}", "-lm", "-lpthread"])
        # libbfm.so.1 (the soname clients record, as the reference's SOVERSION 1)
        # and libbfm.so, the link-time name, pointing at it
        os.replace(tmp, bfm_lib + ".1")
        if os.path.lexists(bfm_lib):
            os.remove(bfm_lib)
        os.symlink("libbfm.so.1", bfm_lib)
    shutil.rmtree(OBJ_DIR, ignore_errors=True)
    if verbose:
        print("built", os.listdir(LIB_DIR))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    build(a.jobs, verbose=True)
