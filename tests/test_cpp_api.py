"""The C++ host API (include/crisoptix.hpp): tests/cpp/render_frames.cpp builds a scene from
meshes, a ShaderBindingTable of materials and emitters, renders frames with
Renderer::LaunchFrame and, on a second renderer, through a Pipeline of raygen passes in the
order of src/Renderer.cpp:259-371.

CPU: the header compiles on its own, and the program links against the library and reports the
ABI version it was built for; the instanced case's model-space scene turns back to the world
scene bit for bit.  GPU: every LaunchFrame frame matches the oracle within 1e-4
relative per channel (as the frame-level parity tests), every Pipeline frame equals its
LaunchFrame frame bit for bit, and the camera basis the C++ Camera computes is the one the
Python Camera computes (within one float32 ulp: the C++ class keeps the field of view and
aspect ratio as float, as the reference's does)."""
import importlib.util
import os
import subprocess

import numpy as np
import pytest

from crisoptix_amd import abi, scenes
from parity_util import rel_mismatch, tent_grid_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("cx_cpp_build", os.path.join(ROOT, "tests", "cpp", "build.py"))
cpp_build = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(cpp_build)

LIB = os.path.join(ROOT, "crisoptix_amd", "libcrisoptix_amd.so")


def test_instanced_scene_files_turn_back_exactly():
    for sc in (scenes.cornell_box(1, 8), tent_grid_scene()):
        for v in (sc.positions, sc.normals, sc.emitter_faces):
            model_space(v)


def test_header_compiles_alone(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text('#include "crisoptix.hpp"\nint main() { crisoptix::ShaderBindingTable s; '
                   's.AddCallableEntry(crisoptix::Material{}); s.CreateSBT(); return (int)s.Rows().size() - 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I",
                    os.path.join(ROOT, "include"), str(src)], check=True)


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_program_links_and_reports_abi(tmp_path):
    exe = str(tmp_path / "render_frames")
    subprocess.run(cpp_build.command(exe), check=True)
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.dirname(LIB))   # the rpath is relative to the tree
    out = subprocess.run([exe, "--abi"], capture_output=True, text=True, check=True, env=env).stdout.split()
    assert out == ["abi", str(abi.ABI_VERSION), str(abi.ABI_VERSION)]


def mesh_runs(sc):
    """Meshes of a flattened scene: maximal runs of triangles with one material and visibility,
    each over its own contiguous vertex range (as the scene builders lay them out)."""
    T = sc.num_triangles
    mask = sc.tri_mask if sc.tri_mask is not None else np.full(T, abi.CX_VIS_ALL, np.uint8)
    key = sc.tri_material.astype(np.int64) * 4 + mask
    starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    ends = np.r_[starts[1:], T]
    runs, next_v = [], 0
    for a, b in zip(starts, ends):
        lo, hi = int(sc.indices[a:b].min()), int(sc.indices[a:b].max())
        assert lo == next_v, "mesh vertex ranges must tile the vertex array in triangle order"
        runs.append((lo, hi - lo + 1, int(a), int(b - a), int(sc.tri_material[a]), int(mask[a])))
        next_v = hi + 1
    assert next_v == len(sc.positions)
    return runs


def model_space(v):
    """Points or directions [..., 3] in the model space of the instanced case: turned by -90
    degrees about y, (x, y, z) -> (-z, y, x), so that the instance transform (x, y, z) -> (z, y, -x)
    turns them back.  Checked to come back bit for bit through the C++ Mat4's float arithmetic
    (((m0 x + m4 y) + m8 z) + m12 per row)."""
    v = np.asarray(v, np.float32)
    f = np.stack([-v[..., 2], v[..., 1], v[..., 0]], -1).astype(np.float32)
    one, zero = np.float32(1), np.float32(0)
    back = np.stack([((zero * f[..., 0] + zero * f[..., 1]) + one * f[..., 2]) + zero,
                     ((zero * f[..., 0] + one * f[..., 1]) + zero * f[..., 2]) + zero,
                     ((-one * f[..., 0] + zero * f[..., 1]) + zero * f[..., 2]) + zero], -1).astype(np.float32)
    assert np.array_equal(back.view(np.uint32), v.view(np.uint32)), "the turn does not come back exactly"
    return f


def write_scene(path, sc, cam, instanced=False):
    runs = mesh_runs(sc)
    has_uv = sc.uvs is not None
    assert not sc.textures
    pos, nrm, faces = sc.positions, sc.normals, sc.emitter_faces
    if instanced:
        pos, nrm, faces = model_space(pos), model_space(nrm), model_space(faces)
    with open(path, "wb") as f:
        f.write(b"CXS1")
        f.write(np.asarray([len(sc.positions), sc.num_triangles, len(sc.materials), len(sc.emitter_color),
                            len(sc.emitter_faces), len(runs), int(has_uv)], "<u4").tobytes())
        for a, dt in ((pos, "<f4"), (nrm, "<f4")):
            f.write(np.ascontiguousarray(a, dt).tobytes())
        if has_uv:
            f.write(np.ascontiguousarray(sc.uvs, "<f4").tobytes())
        f.write(np.ascontiguousarray(sc.indices, "<u4").tobytes())
        f.write(np.ascontiguousarray(sc.materials, "<f4").tobytes())
        f.write(np.ascontiguousarray(faces, "<f4").tobytes())
        f.write(np.ascontiguousarray(sc.emitter_face_offset, "<u4").tobytes())
        f.write(np.ascontiguousarray(sc.emitter_color, "<f4").tobytes())
        f.write(np.asarray(runs, "<u4").tobytes())
        f.write(np.asarray([*cam.pos, *cam.view_dir, *cam.up, cam.fov], "<f4").tobytes())


def read_frames(path, W, H):
    raw = open(path, "rb").read()
    size, n = np.frombuffer(raw[:4], "<f4")[0], int(np.frombuffer(raw[4:8], "<u4")[0])
    rec = 48 + 4 + 2 * W * H * 12
    assert len(raw) == 8 + n * rec
    frames = []
    for i in range(n):
        b = raw[8 + i * rec: 8 + (i + 1) * rec]
        cam = np.frombuffer(b[:48], "<f4").reshape(4, 3)
        fi = int(np.frombuffer(b[48:52], "<u4")[0])
        imgs = np.frombuffer(b[52:], "<f4").reshape(2, H, W, 3)
        frames.append((cam, fi, imgs[0], imgs[1]))
    return float(size), frames


def overrides(p, base):
    out = []
    for name, ctype in abi.LaunchParams._fields_:
        if name in ("random", "width", "height", "frameIdx", "spatialRoundIdx", "pad0", "trackPrefixStats"):
            continue
        a, b = getattr(p, name), getattr(base, name)
        if a != b:
            out.append(f"{name}={a:.9g}" if isinstance(a, float) else f"{name}={a}")
    return out


CASES = {
    "pt_cornell": dict(scene="cornell", mode="path_tracer", W=32, H=24, frames=2),
    "cond_cornell": dict(scene="cornell", mode="conditional_restir", W=40, H=32, frames=3,
                         over=dict(gatherRadius=0.08)),
    # traced lights (meshes of visibility CX_VIS_SURFACE) and a turning camera
    "cond_tent_orbit": dict(scene="tent16", mode="conditional_restir", W=48, H=32, frames=3, orbit=True,
                            over=dict(gatherRadius=0.6)),
    # the same scene built from a Model placed by a ModelInstance and emitters from meshes with
    # the instance's transform
    "cond_tent_instanced": dict(scene="tent16", mode="conditional_restir", W=48, H=32, frames=2, orbit=True,
                                over=dict(gatherRadius=0.6), instanced=True),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_cpp_frames_match_oracle(case, tmp_path, orc, cx_runtime):
    c = CASES[case]
    print(f"case {case}", flush=True)
    W, H = c["W"], c["H"]
    sc = scenes.cornell_box(1, 8) if c["scene"] == "cornell" else tent_grid_scene()
    cam = scenes.cornell_camera(W, H) if c["scene"] == "cornell" else scenes.orbit_camera(sc, W, H)
    oscene = orc.OracleScene(sc)
    p = abi.workload_params(c["mode"], W, H, oscene.size)
    for k, v in c.get("over", {}).items():
        setattr(p, k, v)
    scene_file, out_file = tmp_path / "scene.bin", tmp_path / "frames.bin"
    write_scene(scene_file, sc, cam, instanced=c.get("instanced", False))
    args = [cpp_build.OUT, str(scene_file), str(out_file), str(W), str(H), str(c["frames"]), "4321",
            f"orbit={int(bool(c.get('orbit')))}", f"instanced={int(bool(c.get('instanced')))}",
            *overrides(p, abi.default_params(W, H, oscene.size))]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "expected error" in r.stdout
    size, frames = read_frames(out_file, W, H)
    assert size == oscene.size
    oren = orc.OracleRenderer(oscene, W, H, p, seed=4321)
    for f, (ccam, fi, img_launch, img_pipe) in enumerate(frames):
        if c.get("orbit") and f > 0:
            cam.rotate_around_origin([0, 1, 0], 2 * np.pi / 64)
        py = cam.data()
        for row, k in enumerate(("pos", "U", "V", "W")):
            want = np.asarray(py[k], np.float32)
            assert np.all(np.abs(ccam[row] - want) <= np.spacing(np.abs(want)) + 1e-12), (k, ccam[row], want)
        cd = abi.CameraData.from_dict(dict(zip(("pos", "U", "V", "W"), ccam)))
        ref = oren.frame(cd, f == 0 or bool(c.get("orbit"))).copy()
        bad = rel_mismatch(img_launch, ref).any(-1)
        print(f"{case} frame {f}: frameIdx {fi}, mismatching pixels {int(bad.sum())}/{W * H}, "
              f"mean {ref.mean():.5f} vs {img_launch.mean():.5f}")
        assert not bad.any()
        assert np.array_equal(img_pipe.view(np.uint32), img_launch.view(np.uint32))
        assert fi == 0
