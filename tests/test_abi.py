"""The C-ABI boundary, checked without a GPU.

* libbfmhip.so loads and exports exactly the entry points include/bfmx.h
  declares; without a GPU it reports zero devices and refuses to create a
  system (no CPU fallback).
* libbfm.so exports every function the 15 public headers declare (the
  reference's 67, libbfm/CMakeLists.txt:27-28 + headers) and nothing named
  bfm_* beyond them.
* Struct layouts: sizes and field offsets from the C compiler on include/
  equal the ctypes declarations used by the Python bindings and, where the
  reference tree is present, the reference headers.
"""

import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

from bfm_amd import _ffi as F
from bfm_amd import device as D

INC = os.path.join(ROOT, "include")
REF_INC = "/root/reference/libbfm/src"
LIB_DIR = os.path.join(ROOT, "bfm_amd", "lib")


def declared(header_text, prefix):
    text = re.sub(r"/\*.*?\*/", "", header_text, flags=re.S)
    return set(re.findall(rf"\b({prefix}\w+)\s*\(", text))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_bfmx_exports_match_header():
    hdr = open(os.path.join(INC, "bfmx.h")).read()
    decl = declared(hdr, "bfmx_")
    assert decl == set(D.EXPORTS)
    assert exported(os.path.join(LIB_DIR, "libbfmhip.so")) == decl


def test_bfmx_without_gpu_refuses():
    n = D.device_count()
    if n > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(D.BfmxError) as e:
        D.System(4, np.zeros((4, 2)), np.array([[0, 1, 2, 3]]))
    assert e.value.code == -6


EXTENSION_HEADERS = {"gpu.h"}   # this implementation's additions to the API


def public_functions(include_extensions=False):
    names = set()
    for h in sorted(os.listdir(os.path.join(INC, "bfm"))):
        if h in EXTENSION_HEADERS and not include_extensions:
            continue
        names |= declared(open(os.path.join(INC, "bfm", h)).read(), "bfm_")
    return names


def test_header_function_set_matches_bindings():
    names = public_functions()
    assert len(names) == 67
    assert names == set(F.SIGNATURES)


def test_reference_headers_declare_the_same_functions():
    if not os.path.isdir(REF_INC):
        pytest.skip("reference tree not present")
    ref = set()
    for h in os.listdir(os.path.join(REF_INC, "bfm")):
        if h.endswith(".h"):
            ref |= declared(open(os.path.join(REF_INC, "bfm", h)).read(), "bfm_")
    assert ref == public_functions()


def test_libbfm_exports():
    path = os.path.join(LIB_DIR, "libbfm.so")
    if not os.path.exists(path):
        pytest.fail("libbfm.so is not built")
    ex = {s for s in exported(path) if s.startswith("bfm_")}
    assert ex == public_functions(include_extensions=True)


def layout(inc):
    tmp = os.path.join(ROOT, "tests", "_abi_build")
    os.makedirs(tmp, exist_ok=True)
    exe = os.path.join(tmp, "abi_" + str(abs(hash(inc))))
    subprocess.run(["gcc", "-std=gnu99", "-I", inc, os.path.join(ROOT, "tests", "abi_layout.c"), "-o", exe],
                   check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    shutil.rmtree(tmp, ignore_errors=True)
    return dict((k, int(v)) for k, v in (line.split() for line in out.splitlines()))


def test_struct_layout_matches_ctypes():
    lay = layout(INC)
    classes = {name: getattr(F, name) for name in dir(F) if name.startswith("bfm_") and name.endswith("_t")}
    checked = 0
    for key, val in lay.items():
        struct, field = key.split(".")
        if struct == "enum":
            continue
        cls = classes[struct]
        if field == "sizeof":
            assert C.sizeof(cls) == val, key
        else:
            off = getattr(cls, field).offset if field not in ("full", "band") else cls.u.offset
            assert off == val, key
        checked += 1
    assert checked > 90


def test_struct_layout_matches_reference_headers():
    if not os.path.isdir(REF_INC):
        pytest.skip("reference tree not present")
    assert layout(INC) == layout(REF_INC)
