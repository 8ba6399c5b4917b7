"""Builds tests/cpp/_build/render_frames, the C++ program over include/crisoptix.hpp that the
tests run, with g++ against the in-tree libcrisoptix_amd.so (found at run time through an
$ORIGIN-relative rpath, so the tree can move)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "render_frames")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def command(out: str = OUT) -> list[str]:
    return ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
            "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROCM, "include"),
            os.path.join(HERE, "render_frames.cpp"), "-o", out,
            "-L", os.path.join(ROOT, "crisoptix_amd"), "-lcrisoptix_amd",
            "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
            "-Wl,-rpath,$ORIGIN/../../../crisoptix_amd", "-Wl,-rpath," + os.path.join(ROCM, "lib")]


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    r = subprocess.run(command(tmp), capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for tests/cpp/render_frames.cpp:\n" + r.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
