"""The C ABI library loads without a GPU, exports every function include/crisoptix.h declares,
and the ctypes mirrors have the C layouts."""
import ctypes as C
import os
import re
import subprocess

import pytest

from crisoptix_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "crisoptix.h")
LIB = os.path.join(ROOT, "crisoptix_amd", "libcrisoptix_amd.so")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(cx_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("cx_ctx_create", "cx_scene_build", "cx_trace_closest", "cx_entries_build", "cx_entries_query",
              "cx_renderer_create", "cx_renderer_frame", "cx_pass_launch"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.cx_abi_version() == abi.ABI_VERSION


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "sz.c"
    fields = {
        "cx_launch_params": abi.LaunchParams, "cx_camera": abi.CameraData, "cx_material": abi.Material,
        "cx_scene_desc": abi.SceneDesc, "cx_scene_info": abi.SceneInfo, "cx_frame_stats": abi.FrameStats,
        "cx_state_view": abi.StateView, "cx_texture": abi.Texture, "cx_material_tex": abi.MaterialTex,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in fields.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n") if l)
    for cname, py in fields.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)
