"""Pins of the oracle's FP64 truth renderer (orc_truth_render), the expectation of the
reference's PathTracer estimator (path_gen.h:389-472):

* a closed form: one bounce with NEE probability 1 on a diffuse+GGX plane under a tiny emitter
  equals f(l) L / d^2 (NEE branch, emitter model of include/model/Emitter.h:34-54 without an
  area or cosine factor, the GGX value of brdf.cu:115-190 with F0 = albedo);
* the float oracle's path tracer, accumulated over many frames, agrees with the truth within
  the statistical error (a dropped or extra factor in either estimator shows up as a bias).
"""
import math

import numpy as np
import pytest

from crisoptix_amd import abi, scenes


def plane_scene(albedo, rough, light_pos, radiance, half=50.0):
    v = np.asarray([[-half, 0, -half], [half, 0, -half], [half, 0, half], [-half, 0, half]], np.float32)
    n = np.tile(np.asarray([[0, 1, 0]], np.float32), (4, 1))
    idx = np.asarray([[0, 2, 1], [0, 3, 2]], np.uint32)
    mats = np.asarray([[*albedo, 0.04, 0.04, 0.04, 0, 0, 0, rough]], np.float32)
    s = 1e-4
    lp = np.asarray(light_pos, np.float32)
    face = np.asarray([[lp + [-s, 0, -s], lp + [s, 0, -s], lp + [0, 0, s]]], np.float32)
    return scenes.Scene(positions=v, normals=n, indices=idx, tri_material=np.zeros(2, np.uint32), materials=mats,
                        emitter_faces=face, emitter_face_offset=np.asarray([0, 1], np.uint32),
                        emitter_color=np.asarray([[radiance] * 3], np.float32), name="plane")


def ggx_closed_form(albedo, rough, n, v, l):
    """brdf.cu:115-190 at one direction pair, F0 = albedo (metal = 0 mix), times cos."""
    albedo = np.asarray(albedo, np.float64)
    h = (l + v) / np.linalg.norm(l + v)
    nl, nv, nh, lh = n @ l, n @ v, n @ h, l @ h
    a2 = rough * rough
    D = a2 / (math.pi * ((nh * a2 - nh) * nh + 1) ** 2)
    V = 0.5 / (nl * math.sqrt(nv * nv * (1 - a2) + a2) + nv * math.sqrt(nl * nl * (1 - a2) + a2))
    F = albedo + (1 - albedo) * (1 - lh) ** 5
    return (albedo / math.pi + D * V * F) * nl


def test_single_bounce_closed_form(orc):
    albedo, rough, L = (0.7, 0.4, 0.2), 0.6, 10.0
    light = np.asarray([1.5, 3.0, -0.5])
    sc = plane_scene(albedo, rough, light, L)
    os_ = orc.OracleScene(sc)
    cam_pos = np.asarray([-1.0, 2.0, 0.5])
    target = np.zeros(3)
    fwd = (target - cam_pos) / np.linalg.norm(target - cam_pos)
    cam = abi.CameraData.from_dict({"pos": cam_pos, "U": [1e-7, 0, 0], "V": [0, 0, 1e-7], "W": fwd})
    mean, var = orc.truth_render(os_, cam, 1, 1, 1, 1.0, 64, seed=3)
    x = np.zeros(3)
    n = np.asarray([0.0, 1.0, 0.0])
    v = -fwd
    d = light - x
    dist = np.linalg.norm(d)
    expect = ggx_closed_form(albedo, rough, n, v, d / dist) * L / dist ** 2
    assert np.allclose(mean[0, 0], expect, rtol=2e-5), (mean[0, 0], expect)
    assert var.max() < 1e-9 * expect.max() ** 2


def test_float_path_tracer_agrees_with_truth(orc):
    W = H = 8
    sc = scenes.cornell_box(1, 4)
    os_ = orc.OracleScene(sc)
    camd = scenes.cornell_camera(W, H).data()
    cam = abi.CameraData.from_dict(camd)
    truth, tvar = orc.truth_render(os_, cam, W, H, 4, 0.3, 8192, seed=11)
    p = abi.workload_params("path_tracer", W, H, os_.size)
    p.maxPathLen = 4
    p.enableAccum = 1
    frames = 4096
    ren = orc.OracleRenderer(os_, W, H, p, seed=5)
    for f in range(frames):
        img = ren.frame(cam, f == 0)
    img = img.astype(np.float64)
    # Per-sample variance from the truth run; variance of the frame average is that / frames.
    sample_var = tvar * 8192
    se = math.sqrt(tvar.sum() + (sample_var / frames).sum()) / truth.size
    diff = img.mean() - truth.mean()
    print(f"image means: float PT {img.mean():.5f}, FP64 truth {truth.mean():.5f}, diff {diff:.2e}, se {se:.2e}, "
          f"rel-MSE {orc.rel_mse(img, truth):.2e}")
    assert abs(diff) < 5 * se
    # Per pixel: standardised differences are consistent with noise.
    z = (img - truth) / np.sqrt(tvar + sample_var / frames + 1e-30)
    assert np.mean(z ** 2) < 2.0


def _canonical_chain(orc, sc, cam, W, H, L, frames, seed):
    """Frames of Conditional ReSTIR with every reuse off and gatherM = 1: the final gather's output is
    then one canonical path's f/p per pixel and frame (final_gather.cu:76-91)."""
    os_ = orc.OracleScene(sc)
    p = abi.workload_params("conditional_restir", W, H, os_.size)
    p.maxPathLen = L
    p.gatherM = 1
    p.prefixEnableTemporal = p.prefixEnableSpatial = 0
    p.suffixEnableTemporal = p.suffixEnableSpatial = 0
    ren = orc.OracleRenderer(os_, W, H, p, seed=seed)
    imgs = np.stack([ren.frame(cam, f == 0).astype(np.float64).copy() for f in range(frames)])
    return os_, imgs


@pytest.mark.parametrize("name", ["cornell", "tent"])
def test_conditional_restir_canonical_chain_is_unbiased(orc, name):
    """Pins TracePrefix, TraceSuffix and the canonical combination (path_gen.h:74-387,
    final_gather.cu:76-91) against the FP64 truth.  The prefix stops at its last vertex without an
    NEE draw; the suffix starts there and holds at most maxPathLen - prefixLen vertices, the last
    with NEE forced and p not multiplied by neeProb (an NEE probability of 1 there), so its paths
    carry light from vertices 1 .. maxPathLen - 1: the expectation is the path tracer's at
    maxPathLen - 1.  No material of these scenes emits, so the length-1 emission term of
    path_gen.h:204 never applies.  The test statistic is each 4x4 block's mean against the truth
    (5 standard errors from the frames' own variance); a dropped 1/neeProb or 1/(1 - neeProb), an
    extra vertex or a missing NEE term each shift them by far more."""
    if name == "cornell":
        W, H, L = 8, 8, 4
        sc = scenes.cornell_box(1, 4)
        camd = scenes.cornell_camera(W, H).data()
        frames = 4096
    else:
        W, H, L = 8, 8, 6
        sc = scenes.object_grid(grid=3, seed=5, ground_quads=20, ico_subdiv=2, num_lights=16)
        camd = scenes.orbit_camera(sc, W, H).data()
        frames = 8192
    cam = abi.CameraData.from_dict(camd)
    os_, imgs = _canonical_chain(orc, sc, cam, W, H, L, frames, seed=21)
    truth, tvar = orc.truth_render(os_, cam, W, H, L - 1, 0.3, 16384, seed=13)
    mean = imgs.mean(0)
    var_mean = imgs.var(0, ddof=1) / frames
    bm = mean.reshape(H // 4, 4, W // 4, 4, 3).mean((1, 3))
    bt = truth.reshape(H // 4, 4, W // 4, 4, 3).mean((1, 3))
    bse = np.sqrt((var_mean + tvar).reshape(H // 4, 4, W // 4, 4, 3).sum((1, 3))) / 16
    z = (bm - bt) / (bse + 1e-30)
    print(f"{name}: image mean {mean.mean():.5f} vs truth(L-1) {truth.mean():.5f}; block z {np.round(z.ravel(), 2)}")
    assert np.all(np.abs(z) < 5.0)
    # the truth one vertex longer is far off: the comparison can tell the path lengths apart
    longer, lvar = orc.truth_render(os_, cam, W, H, L, 0.3, 16384, seed=14)
    zl = (mean.mean() - longer.mean()) / (np.sqrt((var_mean + lvar).sum()) / mean.size)
    print(f"{name}: against the truth at maxPathLen {L}: z {zl:.1f}")
    assert abs(zl) > 5.0
