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
    P, N, T, tm, mats = assets.parse_obj(OBJ, lambda name: MTL)
    assert T.shape == (5, 3)                     # quad -> 2, then 3 triangles
    assert tm == ["red", "red", "lamp", "lamp", "undefined_one"]
    assert np.allclose(N[:6], [0, 0, 1])         # given normals
    assert np.allclose(P[6:9], [[0, 0, 0], [1, 0, 0], [0, 0, 1]])   # negative indices
    n = N[9]                                     # flat normal of f 1 3 5
    e = np.cross(P[10] - P[9], P[11] - P[9])
    assert np.allclose(n, e / np.linalg.norm(e))
    assert mats["red"]["Pr"] == 0.35 and mats["lamp"]["Ke"] == (10.0, 10.0, 10.0)
    assert mats["lamp"]["textures"] == ["lamp.png"]


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


def cornell_obj_scene(trace_emitters=False):
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


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["path_tracer", "conditional_restir"])
def test_gpu_obj_scene_bit_exact(mode, orc, cx_runtime, cx_ctx):
    sc = cornell_obj_scene()
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
