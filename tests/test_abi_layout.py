"""The Python host's ctypes structures match include/piks3d.h field by
field: a C program compiled against the header prints every field's offset
and every structure's size, which must equal ctypes' (a field the header
does not have fails to compile)."""
import ctypes
import os
import shutil
import subprocess

import pytest

import piks3d_amd as pk

HERE = os.path.dirname(os.path.abspath(__file__))
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

STRUCTS = [("piks_params", pk.PiksParams), ("piks_info", pk.PiksInfo),
           ("piks_serve_params", pk.ServeParams), ("piks_serve_result", pk.ServeResult)]


def test_ctypes_structures_match_the_header(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "piks3d.h"', "int main(void) {"]
    want = []
    for cname, cls in STRUCTS:
        lines.append('  printf("%%zu\\n", sizeof(%s));' % cname)
        want.append(ctypes.sizeof(cls))
        for field, _ in cls._fields_:
            lines.append('  printf("%%zu\\n", offsetof(%s, %s));' % (cname, field))
            want.append(getattr(cls, field).offset)
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I" + INCLUDE, "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == want
