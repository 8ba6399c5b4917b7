"""OBJ / MTL scene loading (Model::LoadMaterials / LoadMesh, src/Model.cpp:9-374; implicit
emitters of Scene::BuildEmitterTable, src/Scene.cpp:176-204), and the loaded scene through the
frame loop: on the CPU with the oracle, and bit-exact against it on the GPU."""
import numpy as np
import pytest

from crisoptix_amd import abi, assets, scenes

OBJ = """
mtllib m.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
vn 0 0 1
usemtl red
f 1//1 2//1 3//1 4//1
usemtl lamp
f -5 -4 -1
f 1 3 5
usemtl undefined_one
f 2 3 5
"""
MTL = """
newmtl red
Kd 0.8 0.1 0.1
Ks 0.04 0.04 0.04
Pr 0.35
newmtl lamp   # an emitter
Kd 0.5 0.5 0.5
Ke 10 10 10
map_Kd lamp.png
"""


def test_parse_obj_and_mtl():
    P, N, UV, T, tm, mats = assets.parse_obj(OBJ, lambda name: MTL)
    assert T.shape == (5, 3)                     # quad -> 2, then 3 triangles
    assert tm == ["red", "red", "lamp", "lamp", "undefined_one"]
    assert np.allclose(N[:6], [0, 0, 1])         # given normals
    assert np.allclose(P[6:9], [[0, 0, 0], [1, 0, 0], [0, 0, 1]])   # negative indices
    n = N[9]                                     # flat normal of f 1 3 5
    e = np.cross(P[10] - P[9], P[11] - P[9])
    assert np.allclose(n, e / np.linalg.norm(e))
    assert mats["red"]["Pr"] == 0.35 and mats["lamp"]["Ke"] == (10.0, 10.0, 10.0)
    assert mats["lamp"]["map_Kd"] == "lamp.png"
    assert np.all(UV == 0)                       # no vt: uv (0, 0)


def test_uvs_and_flip():
    obj = "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nvt 0.25 0.5\nvt 0.75 0.125\nvt 2 -1\nusemtl m\nf 1/3 2/1 3/2\n"
    mtl = "newmtl m\nKd 1 1 1\nKe 1 1 1\nmap_Kd -s 2 2 1 t.png\nmap_Ks t.png\nmap_Pr r.png\n"
    P, N, UV, T, tm, mats = assets.parse_obj(obj, lambda n: mtl)
    assert UV.tolist() == [[2, -1], [0.25, 0.5], [0.75, 0.125]]
    loaded = []
    img = {"t.png": np.full((2, 3, 4), 7, np.uint8), "r.png": np.full((1, 1, 4), 9, np.uint8)}

    def tl(name):
        loaded.append(name)
        return img[name]

    sc = assets.scene_from_obj(obj, lambda n: mtl, texture_loader=tl, flip_uv=True, spec_tex_usage=1)
    assert loaded == ["t.png", "r.png"]          # one load per file
    assert sc.material_tex.tolist() == [[0, 0, 1, 1]]
    assert np.allclose(sc.uvs, [[2, 2], [0.25, 0.5], [0.75, 0.875]])   # v -> 1 - v


def test_load_obj_reads_textures_from_files(tmp_path):
    from crisoptix_amd import image
    obj, mtl, tex = assets.cornell_box_obj(textured=True)
    (tmp_path / "box.obj").write_text(obj)
    (tmp_path / "cornell.mtl").write_text(mtl)
    for name, im in tex.items():
        image.write_png(str(tmp_path / name), im)
    sc = assets.load_obj(str(tmp_path / "box.obj"), trace_emitters=False)
    assert len(sc.textures) == 2                  # checker.dds was read as checker.png
    assert np.array_equal(sc.textures[0], tex["checker.png"]) and np.array_equal(sc.textures[1], tex["rough.png"])
    assert sc.uvs.shape == (sc.positions.shape[0], 2) and sc.uvs.max() == 3.0


def test_scene_from_obj_materials_and_emitters():
    sc = assets.scene_from_obj(OBJ, lambda name: MTL)
    assert sc.materials.shape == (3, 10)
    assert np.allclose(sc.materials[0], [0.8, 0.1, 0.1, 0.04, 0.04, 0.04, 0, 0, 0, 0.35])
    assert np.allclose(sc.materials[2], [0.6, 0.6, 0.6, 0, 0, 0, 0, 0, 0, 1.0])   # assimp's default
    assert sc.emitter_face_offset.tolist() == [0, 2]
    assert np.allclose(sc.emitter_color, [[1, 1, 1]])    # implicit emitters have colour 1
    assert sc.num_triangles == 5
    sc2 = assets.scene_from_obj(OBJ, lambda name: MTL, trace_emitters=False)
    assert sc2.num_triangles == 3 and sc2.emitter_face_offset.tolist() == [0, 2]


def cornell_obj_scene(trace_emitters=False, textured=False):
    if textured:
        obj, mtl, tex = assets.cornell_box_obj(textured=True)
        return assets.scene_from_obj(obj, lambda name: mtl, trace_emitters=trace_emitters, name="cornell_obj_tex",
                                     texture_loader=lambda n: tex[n.replace(".dds", ".png")])
    obj, mtl = assets.cornell_box_obj()
    return assets.scene_from_obj(obj, lambda name: mtl, trace_emitters=trace_emitters, name="cornell_obj")


def test_cornell_obj_orientation():
    sc = cornell_obj_scene(trace_emitters=True)
    P, N, tri = sc.positions, sc.normals, sc.indices
    centroid = P[tri].mean(1)
    nt = N[tri[:, 0]]
    room = np.asarray([2.78, 2.78, 2.78])
    # walls, floor, ceiling and light (the first 12 triangles) face into the room; each block's
    # 10 triangles (top and four sides) face away from that block's centre
    assert np.all((nt[:12] * (room - centroid[:12])).sum(1) > 0)
    for b in range(2):
        sl = slice(12 + 10 * b, 22 + 10 * b)
        centre = centroid[sl].mean(0) - np.asarray([0.0, 0.1, 0.0])
        assert np.all((nt[sl] * (centroid[sl] - centre)).sum(1) > 0)
    assert sc.emitter_face_offset[-1] == 2


def obj_camera(W, H):
    return scenes.Camera(pos=np.asarray([2.78, 2.73, -8.0], np.float32), view_dir=np.asarray([0, 0, 1], np.float32),
                         up=np.asarray([0, 1, 0], np.float32), aspect=W / H, fov=np.radians(39.0))


def test_oracle_renders_obj_scene(orc):
    sc = cornell_obj_scene()
    os_ = orc.OracleScene(sc)
    W = H = 32
    cd = abi.CameraData.from_dict(obj_camera(W, H).data())
    p = abi.workload_params("path_tracer", W, H, os_.size)
    p.maxPathLen = 4
    r = orc.OracleRenderer(os_, W, H, p, seed=1)
    img = np.mean([r.frame(cd, i == 0).copy() for i in range(16)], axis=0)
    # implicit emitters have unit colour (Scene.cpp:194), so the box is dim
    assert np.isfinite(img).all() and img.mean() > 1e-3
    # the red wall is on the image's left (camera at -z looking +z with +x to the left), the
    # green one on its right
    assert img[:, :4, 0].mean() > 2 * img[:, :4, 1].mean()
    assert img[:, -4:, 1].mean() > 1.5 * img[:, -4:, 0].mean()


def test_textured_obj_changes_the_image(orc):
    W = H = 24
    cd = abi.CameraData.from_dict(obj_camera(W, H).data())
    imgs = []
    for textured in (False, True):
        os_ = orc.OracleScene(cornell_obj_scene(textured=textured))
        p = abi.workload_params("path_tracer", W, H, os_.size)
        p.maxPathLen = 3
        r = orc.OracleRenderer(os_, W, H, p, seed=2)
        imgs.append(np.mean([r.frame(cd, i == 0).copy() for i in range(8)], axis=0))
    # the checker floor (two blues on a white-Kd material) fills the lower rows
    lo0, lo1 = imgs[0][-6:], imgs[1][-6:]
    assert np.isfinite(imgs[1]).all()
    assert lo1[..., 2].mean() / lo1[..., 0].mean() > 1.5 * lo0[..., 2].mean() / lo0[..., 0].mean()


@pytest.mark.gpu
@pytest.mark.parametrize("textured", [False, True])
@pytest.mark.parametrize("mode", ["path_tracer", "conditional_restir"])
def test_gpu_obj_scene_bit_exact(mode, textured, orc, cx_runtime, cx_ctx):
    print(f"obj scene {mode} textured={textured}", flush=True)
    sc = cornell_obj_scene(textured=textured)
    W, H = 48, 40
    cd = abi.CameraData.from_dict(obj_camera(W, H).data())
    gs = cx_runtime.Scene(cx_ctx, sc)
    p = abi.workload_params(mode, W, H, gs.size)
    p.maxPathLen = 4
    if mode == "conditional_restir":
        p.gatherRadius = 0.08 * gs.size
    ren = cx_runtime.Renderer(W, H, cx_ctx, obj_camera(W, H), gs, params=p, seed=4)
    oren = orc.OracleRenderer(orc.OracleScene(sc), W, H, p, seed=4)
    for f in range(3):
        g = ren.launch_frame(cd, f == 0).numpy()
        o = oren.frame(cd, f == 0)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), f
    assert g.mean() > 0
