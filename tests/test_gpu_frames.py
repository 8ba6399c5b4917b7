"""Frame-level parity: the HIP frame loop against the oracle's restatement of the reference
shader logic (shader/*.cu, include/graph/restir/*.h) on identical seeded inputs.

Per-pixel radiance must agree within 1e-4 relative after every frame; reservoir records,
RNG states and neighbour sets are compared bit for bit.  Both execution modes of the HIP
library (wavefront state machines and one megakernel per pass) are checked; the wavefront mode
with its default (two rounds per machine, then the in-place tail kernel), with every round as
traversal + step (rounds0) and with the tail after one round (rounds1)."""
import numpy as np
import pytest

from crisoptix_amd import abi, scenes
from parity_util import rel_mismatch, small_grid_scene, tent_grid_scene

pytestmark = pytest.mark.gpu

CASES = {
    "pt_cornell": dict(scene="cornell", mode="path_tracer", W=64, H=48, frames=3),
    "restir_pt_cornell": dict(scene="cornell", mode="restir_pt_cornell", W=64, H=48, frames=6),
    "restir_pt_temporal": dict(scene="cornell", mode="restir_pt_cornell", W=40, H=40, frames=5,
                               over=dict(prefixEnableSpatial=0)),
    "cond_cornell": dict(scene="cornell", mode="conditional_restir", W=48, H=40, frames=4,
                         over=dict(gatherRadius=0.08)),
    "cond_cornell_gatherN2": dict(scene="cornell", mode="conditional_restir", W=32, H=24, frames=3,
                                  over=dict(gatherRadius=0.08, gatherN=2, suffixSpatialCount=2, windowRadius=6)),
    "cond_grid_orbit": dict(scene="grid3", mode="conditional_restir", W=48, H=32, frames=4, orbit=True,
                            over=dict(gatherRadius=0.3)),
    # the same with the BVH built on the GPU (PLOC): same images as with the host-built tree
    "cond_grid_orbit_devbvh": dict(scene="grid3", mode="conditional_restir", W=48, H=32, frames=4, orbit=True,
                                   over=dict(gatherRadius=0.3), builder="device"),
    # textured materials (diffuse, roughness and OccRoughMetal specular textures, wrapped uvs)
    "pt_grid_tex": dict(scene="grid3_tex", mode="path_tracer", W=48, H=32, frames=3),
    "restir_pt_grid_tex": dict(scene="grid3_tex", mode="restir_pt_cornell", W=48, H=32, frames=4),
    "cond_grid_tex_orbit": dict(scene="grid3_tex", mode="conditional_restir", W=48, H=32, frames=4, orbit=True,
                                over=dict(gatherRadius=0.3)),
    # the Conditional ReSTIR workload reduced (tent of 16 traced lights, k = 8): the final gather
    # sees every neighbour count from 1 to 8, and at least 250 pixels gather all 8 neighbours
    "cond_tent_orbit": dict(scene="tent16", mode="conditional_restir", W=48, H=32, frames=4, orbit=True,
                            over=dict(gatherRadius=0.6), full_k=250, every_count=True),
    # the ends of the neighbour count the reference accepts (gatherM in [1, 32], Renderer.cpp:462):
    # no neighbours at all, one, and 31 (the wide query kernel), at least 200 pixels gathering 31
    "cond_cornell_gatherM1": dict(scene="cornell", mode="conditional_restir", W=40, H=32, frames=3,
                                  over=dict(gatherRadius=0.08, gatherM=1)),
    "cond_tent_gatherM2": dict(scene="tent16", mode="conditional_restir", W=40, H=32, frames=3, orbit=True,
                               over=dict(gatherRadius=0.6, gatherM=2), full_k=300),
    "cond_tent_gatherM32": dict(scene="tent16", mode="conditional_restir", W=40, H=32, frames=3, orbit=True,
                                over=dict(gatherRadius=1.2, gatherM=32), full_k=200),
    # the closed Cornell box with its light traced as geometry (C9')
    "cond_cornell_traced_light": dict(scene="cornell_tl", mode="conditional_restir", W=40, H=32, frames=3,
                                      over=dict(gatherRadius=0.08)),
}


def _scene(name):
    if name == "cornell":
        return scenes.cornell_box(1, 8)
    if name == "cornell_tl":
        return scenes.cornell_box(1, 8, trace_emitters=True)
    if name == "tent16":
        return tent_grid_scene()
    sc = small_grid_scene()
    return scenes.with_textures(sc) if name.endswith("_tex") else sc


def _setup(case, orc, rt, ctx, exec_mode="wavefront"):
    c = CASES[case]
    sc = _scene(c["scene"])
    W, H = c["W"], c["H"]
    gscene = rt.Scene(ctx, sc, builder=c.get("builder", "host"))
    oscene = orc.OracleScene(sc)
    assert abs(gscene.size - oscene.size) == 0.0
    p = abi.workload_params(c["mode"], W, H, gscene.size)
    for k, v in c.get("over", {}).items():
        setattr(p, k, v)
    cam = scenes.cornell_camera(W, H) if c["scene"].startswith("cornell") else scenes.orbit_camera(sc, W, H)
    ren = rt.Renderer(W, H, ctx, cam, gscene, params=p, seed=1234)
    if exec_mode.startswith("rounds"):   # wavefront rounds (0: all), then the in-place tail kernel
        ren.set_exec_mode("wavefront")
        ren.set_wave_rounds(int(exec_mode[6:]))
    else:
        ren.set_exec_mode(exec_mode)
    oren = orc.OracleRenderer(oscene, W, H, p, seed=1234)
    return c, cam, ren, oren


def _compare_prefix(g, o, which_valid):
    valid = (o["flags"] >> 16) & 1
    assert np.array_equal(g["flags"], o["flags"])
    v = valid.astype(bool) & which_valid
    assert np.array_equal(g["rngState"][v], o["rngState"][v])
    for name in ("f", "postReconF", "p", "wSum", "confidence", "pathLen", "reconOutDir"):
        gv = np.asarray(g[name])[v]
        ov = np.asarray(o[name])[v]
        assert not rel_mismatch(gv, ov).any(), name
    assert np.array_equal(g["primary_prim"][v], o["primaryInt"]["prim"][v])
    assert np.array_equal(g["last_prim"][v], o["lastInt"]["prim"][v])


@pytest.mark.parametrize("exec_mode", ["wavefront", "megakernel", "rounds0", "rounds1"])
@pytest.mark.parametrize("case", list(CASES))
def test_frames_match_oracle(case, exec_mode, orc, cx_runtime, cx_ctx):
    print(f"case {case} exec {exec_mode}", flush=True)
    c, cam, ren, oren = _setup(case, orc, cx_runtime, cx_ctx, exec_mode)
    P = c["W"] * c["H"]
    worst = 0
    for f in range(c["frames"]):
        if c.get("orbit") and f > 0:
            cam.rotate_around_origin([0, 1, 0], 2 * np.pi / 64)
        cd = abi.CameraData.from_dict(cam.data())
        changed = f == 0 or bool(c.get("orbit"))
        out_o = oren.frame(cd, changed).copy()
        out_g = ren.launch_frame(cd, changed).numpy()
        bad = rel_mismatch(out_g, out_o).any(axis=-1)
        exact = np.array_equal(out_g.view(np.uint32), out_o.view(np.uint32))
        print(f"{case} frame {f}: mismatching pixels {bad.sum()}/{P}, bit-exact {exact}, "
              f"mean {out_o.mean():.5f} vs {out_g.mean():.5f}")
        worst = max(worst, int(bad.sum()))
        if bad.any():
            idx = np.argwhere(bad)[:5]
            for y, x in idx:
                print("   px", x, y, out_o[y, x], out_g[y, x])
    assert worst == 0
    assert ren.stats().queueOverflow == 0
    front, back, _, _ = oren.buffers()
    st = ren.state()
    assert (st.frontBufferIdx, st.backBufferIdx) == (front, back)
    if c["mode"] != "path_tracer":
        go = ren.gbuffer()
        oo = oren.gbuffer()
        assert np.array_equal(go["primaryIntValid"], oo[:, 2])
        assert np.array_equal(go["canonFlags"], oo[:, 3])
        assert np.array_equal(go["rngState"], oo[:, 0].astype(np.uint64) | (oo[:, 1].astype(np.uint64) << np.uint64(32)))
        for which in (front, back):
            _compare_prefix(ren.prefix(which), {n: oren.prefix(which)[n] for n in orc.PREFIX_DT.names},
                            np.ones(P, bool))
    if c["mode"] == "conditional_restir":
        gs = ren.suffix(front)
        os_ = oren.suffix(front)
        assert np.array_equal(gs["flags"], os_["flags"])
        v = ((os_["flags"] >> 16) & 1).astype(bool)
        for name in ("f", "postReconF", "p", "wSum", "confidence"):
            assert not rel_mismatch(np.asarray(gs[name])[v], np.asarray(os_[name])[v]).any(), name
        gi, gd, gc = ren.neighbors()
        oi, od, oc = oren.neighbors()
        assert np.array_equal(gc, oc)
        for p in range(P):
            n = oc[p]
            assert np.array_equal(gi[p, :n], oi[p, :n])
            assert np.array_equal(gd[p, :n].view(np.uint32), od[p, :n].view(np.uint32))
        k = int(c.get("over", {}).get("gatherM", 9)) - 1
        hist = np.bincount(oc, minlength=k + 1)
        print(f"{case}: mean neighbours {oc.mean():.2f}, histogram {hist.tolist()}, "
              f"valid suffixes {v.mean():.2f}")
        if c.get("every_count"):
            assert (hist[1:] > 0).all(), hist
        if "full_k" in c:
            assert hist[k] >= c["full_k"], hist
