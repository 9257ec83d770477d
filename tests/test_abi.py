"""The C ABI boundary without a GPU: libprt.so exports every function that
include/prt.h declares, and the ctypes mirrors of its structs (the Python
binding) have the C layout (sizes and field offsets from a program compiled
against the header with gcc)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from prender_amd import scene as S
from prender_amd import prt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "prt.h")

STRUCTS = {
    "prt_material": S.PrtMaterial, "prt_entity": S.PrtEntity, "prt_camera": S.PrtCamera,
    "prt_sampling": S.PrtSampling, "prt_scene_desc": S.PrtSceneDesc, "prt_ray": S.PrtRay, "prt_hit": S.PrtHit,
    "prt_stats": S.PrtStats, "prt_render_opts": S.PrtRenderOpts, "prt_batch_opts": S.PrtBatchOpts,
    "prt_dump_opts": S.PrtDumpOpts,
}


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_list():
    assert declared_functions() == sorted(prt.EXPORTED)


def test_library_exports_every_declared_function():
    if not os.path.exists(prt.LIB_PATH):
        pytest.skip("libprt.so is not built (python -m prender_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", prt.LIB_PATH], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_library_loads_and_reports_its_abi_version():
    if not os.path.exists(prt.LIB_PATH):
        pytest.skip("libprt.so is not built (python -m prender_amd.build)")
    L = prt.lib()
    m = re.search(r"#define PRT_ABI_VERSION (\d+)", open(HEADER).read())
    assert L.prt_abi_version() == int(m.group(1))
    assert L.prt_last_error() is not None


def test_struct_layouts_match_c(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    got = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    c_layout = {}
    for l in got:
        if l:
            a, b, v = l.split()
            c_layout[(a, b)] = int(v)
    for cname, py in STRUCTS.items():
        assert c_layout[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert c_layout[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
