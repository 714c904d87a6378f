"""The ctypes mirrors of the C-ABI structs (p2r_amd/_capi.py) against the headers, on the CPU.

A small C program built by gcc from include/p2r.h and include/p2r_gen.h prints the size of every struct the
bindings pass by pointer or by value and the offset of each field; the ctypes classes must agree field for field.
A field added to a header but not to its mirror (or a changed type) shifts a size or an offset and fails here,
before a call would read past the caller's struct.
"""
import ctypes as C
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "p2r-tests_amd"))

from p2r_amd import _capi  # noqa: E402  (ctypes declarations only; loads no library)

STRUCTS = ["p2r_shape", "p2r_params", "p2r_tuning", "p2r_tune_entry", "p2r_tune_report", "p2r_gen_config"]


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("abi")
    lines = ['#include <stdio.h>', '#include "p2r_gen.h"', "int main(void) {"]
    for s in STRUCTS:
        lines.append(f'  printf("{s} %zu\\n", sizeof({s}));')
        for f, _ in getattr(_capi, s)._fields_:
            lines.append(f'  printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines += ["  return 0;", "}"]
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    return {k: int(v) for k, v in (line.split() for line in out if line)}


@pytest.mark.parametrize("name", STRUCTS)
def test_ctypes_struct_matches_header(c_layout, name):
    cls = getattr(_capi, name)
    assert C.sizeof(cls) == c_layout[name], f"sizeof({name})"
    for f, _ in cls._fields_:
        assert getattr(cls, f).offset == c_layout[f"{name}.{f}"], f"offsetof({name}, {f})"
