"""The oracle's FP64 path-replay adjoint against double-precision central differences of its
own replayed estimate (the evaluation table perturbed, paths and decisions fixed).  This pins
the BSDF derivatives (brdf.cu:99-190 with F0 = albedo), the emitter-colour derivatives and the
product rule over replayed paths, for ReSTIR PT (temporal and spatial reuse, shifted samples
re-traced from their records) and the path tracer, before the GPU adjoint is checked against it."""
import numpy as np
import pytest

from crisoptix_amd import abi, scenes

W, H = 48, 32


def adjoint_scene():
    """The reduced object grid with its lights as traced panels.  (A tent of lights has vertices
    on one light face sampling points of the same face, where the cosine is zero in exact
    arithmetic: its sign differs between the float frame and the double replay, and so does
    such a sample's tiny contribution.)"""
    return scenes.object_grid(grid=3, seed=5, ground_quads=20, ico_subdiv=2, num_lights=4, lights="panels")


def setup(orc, mode, textured=False):
    sc = adjoint_scene()
    if textured:
        sc = scenes.with_textures(sc)
    os_ = orc.OracleScene(sc)
    name = "restir_pt_cornell" if mode == "restir_pt" else "path_tracer"
    p = abi.workload_params(name, W, H, os_.size)
    p.maxPathLen = 4
    cd = abi.CameraData.from_dict(scenes.orbit_camera(sc, W, H).data())
    r = orc.OracleRenderer(os_, W, H, p, seed=5)
    for f in range(4):
        out = r.frame(cd, f == 0).copy()
    return sc, os_, r, out


@pytest.mark.parametrize("mode,textured", [("restir_pt", False), ("path_tracer", False), ("restir_pt", True)])
def test_fp64_adjoint_matches_central_differences(orc, mode, textured):
    """Textured: texture-driven diffuse colours and roughnesses have zero gradient (and zero
    central difference); OccRoughMetal materials carry the (1 - metal) factor on the albedo."""
    sc, os_, r, out = setup(orc, mode, textured)
    g = np.random.default_rng(0).standard_normal((H, W, 3)).astype(np.float32)
    gA, gR, gE, rep, st = r.adjoint(g)
    assert st[0] > 100 and st[2] == 0
    # the replayed estimate is the frame's output (double vs float arithmetic)
    assert np.allclose(rep, out, rtol=1e-4, atol=1e-7)
    mats = sc.materials.astype(np.float32)
    emit = sc.emitter_color.astype(np.float32)
    if textured:
        mt = sc.material_tex
        assert np.all(gA[mt[:, 0] >= 0] == 0)
        assert np.all(gR[(mt[:, 2] >= 0) | (mt[:, 3] == abi.CX_SPEC_TEX_OCC_ROUGH_METAL)] == 0)
        metal = (mt[:, 3] == abi.CX_SPEC_TEX_OCC_ROUGH_METAL) & (mt[:, 0] < 0)
        assert np.abs(gA[metal]).max() > 0 and np.abs(gR[(mt[:, 2] < 0) & ~metal]).max() > 0

    def J(m, e):
        os_.set_materials(m, 2)
        os_.set_emitter_colors(e, 2)
        rp = r.adjoint(None)[3]
        os_.set_materials(mats, 2)
        os_.set_emitter_colors(emit, 2)
        return float(np.sum(rp * g))

    fd, ad = [], []
    for m in range(mats.shape[0]):
        for col, grad in [(0, gA[m, 0]), (1, gA[m, 1]), (2, gA[m, 2]), (9, gR[m])]:
            a, b = mats.copy(), mats.copy()
            h = np.float32(1e-4 if col == 9 else 1e-3)   # albedo enters linearly; roughness does not
            a[m, col] += h
            b[m, col] -= h
            h2 = float(a[m, col]) - float(b[m, col])
            fd.append((J(a, emit) - J(b, emit)) / h2)
            ad.append(grad)
    for e in range(emit.shape[0]):
        for c in range(3):
            a, b = emit.copy(), emit.copy()
            a[e, c] += np.float32(1e-2)
            b[e, c] -= np.float32(1e-2)
            h2 = float(a[e, c]) - float(b[e, c])
            fd.append((J(mats, a) - J(mats, b)) / h2)
            ad.append(gE[e, c])
    fd, ad = np.asarray(fd), np.asarray(ad)
    scale = np.abs(ad).max()
    assert scale > 0
    err = np.abs(fd - ad) / np.maximum(np.abs(ad), 1e-2 * scale)
    assert err.max() < 1e-3, (err.max(), int(err.argmax()))
