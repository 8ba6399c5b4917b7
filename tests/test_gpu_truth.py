"""Converged HIP images against the FP64 truth (tests/golden/truth_*.npz, written by
tests/golden/make_truth.py from the oracle's double-precision PathTracer estimator).

rel-MSE = mean over pixels and channels of (I - I_ref)^2 / (I_ref^2 + 1e-2).

* PathTracer, 4096 accumulated frames: an unbiased estimate of the truth's expectation, so the
  rel-MSE must be what the per-sample variance predicts (no bias term), and the image mean must
  agree within 5 standard errors.
* ReSTIR PT and Conditional ReSTIR: the reference's estimators deviate from the PathTracer
  expectation in known ways (emission scaling, self re-trace, reconnection index, prefix
  Jacobian at the primary hit), so their rel-MSE is measured and reported, with a loose bound.
"""
import math
import os

import numpy as np
import pytest

from crisoptix_amd import abi, scenes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FRAMES = 4096


def rel_mse(img, ref):
    return float(np.mean((np.asarray(img, np.float64) - ref) ** 2 / (ref ** 2 + 1e-2)))


def load_truth(name):
    with np.load(os.path.join(HERE, "golden", f"truth_{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def accumulate(rt, ctx, mode, frames, exec_mode="wavefront"):
    t = load_truth("cornell_64x64")
    W = H = 64
    sc = scenes.cornell_box(1, 8)
    gs = rt.Scene(ctx, sc)
    p = abi.workload_params(mode, W, H, gs.size)
    p.maxPathLen = 4
    p.enableAccum = 1
    cam = scenes.cornell_camera(W, H)
    cd = abi.CameraData.from_dict(cam.data())
    assert np.array_equal(np.concatenate([np.asarray(getattr(cd, k), np.float32) for k in ("pos", "U", "V", "W")]),
                          t["camera"])
    ren = rt.Renderer(W, H, ctx, cam, gs, params=p, seed=99)
    ren.set_exec_mode(exec_mode)
    for f in range(frames):
        ren.launch_frame(cd, f == 0)
    return ren.output.numpy().astype(np.float64), t


def test_path_tracer_converges_to_truth(cx_runtime, cx_ctx):
    img, t = accumulate(cx_runtime, cx_ctx, "path_tracer", FRAMES)
    truth, tvar, spp = t["mean"], t["var"], int(t["spp"])
    sample_var = tvar * spp
    expect = float(np.mean((sample_var / FRAMES + tvar) / (truth ** 2 + 1e-2)))
    got = rel_mse(img, truth)
    se = math.sqrt((sample_var / FRAMES + tvar).sum()) / truth.size
    print(f"PathTracer {FRAMES} frames: rel-MSE {got:.3e} (noise-only expectation {expect:.3e}), "
          f"mean {img.mean():.5f} vs truth {truth.mean():.5f} (se {se:.1e})")
    assert abs(img.mean() - truth.mean()) < 5 * se
    assert got < 1.5 * expect


@pytest.mark.parametrize("mode", ["restir_pt_cornell", "conditional_restir"])
def test_restir_rel_mse_reported(mode, cx_runtime, cx_ctx):
    img, t = accumulate(cx_runtime, cx_ctx, mode, FRAMES)
    got = rel_mse(img, t["mean"])
    print(f"{mode} {FRAMES} frames: rel-MSE vs FP64 truth {got:.3e}, mean {img.mean():.5f} vs {t['mean'].mean():.5f}")
    assert np.isfinite(img).all()
    assert got < 0.5
