"""The ctypes mirrors of include/bfmx.h's structs (bfm_amd/device.py) have the
C layout gcc gives them: every field offset and the size."""

import ctypes as C
import os
import shutil
import subprocess

from bfm_amd import device as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def c_layout(struct, fields, tmp):
    src = os.path.join(tmp, "lay.c")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bfmx.h"', "int main(void) {",
             f'printf("sizeof %zu\\n", sizeof({struct}));']
    lines += [f'printf("{f} %zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines += ["return 0;", "}"]
    open(src, "w").write("\n".join(lines))
    exe = os.path.join(tmp, "lay")
    subprocess.run(["gcc", "-std=gnu99", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return dict((k, int(v)) for k, v in (line.split() for line in out.splitlines()))


def test_bfmx_struct_layouts(tmp_path):
    for struct, cls in (("bfmx_instance_t", D.Instance), ("bfmx_stats_t", D.Stats)):
        fields = [f for f, _ in cls._fields_]
        lay = c_layout(struct, fields, str(tmp_path))
        assert lay.pop("sizeof") == C.sizeof(cls), struct
        for f in fields:
            assert lay[f] == getattr(cls, f).offset, (struct, f)
