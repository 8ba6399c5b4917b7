"""Edge cases of the frame loop against the oracle: tiny and odd resolutions (1 x 1, 7 x 5,
33 x 17), the shortest paths (max path length 1 to 3), and a scene whose geometry is empty (only
explicit emitters: every ray misses, no prefix is valid and the entry tree is empty).  Radiance
must agree per pixel within 1e-4 relative after every frame, in both execution modes."""
import dataclasses

import numpy as np
import pytest

from crisoptix_amd import abi, scenes
from parity_util import rel_mismatch

pytestmark = pytest.mark.gpu

MODES = ("path_tracer", "restir_pt_cornell", "conditional_restir")


def _run(orc, rt, ctx, sc, mode, W, H, frames=3, exec_mode="wavefront", over=None):
    gscene = rt.Scene(ctx, sc)
    oscene = orc.OracleScene(sc)
    p = abi.workload_params(mode, W, H, gscene.size)
    for k, v in (over or {}).items():
        setattr(p, k, v)
    cam = scenes.cornell_camera(W, H)
    ren = rt.Renderer(W, H, ctx, cam, gscene, params=p, seed=77)
    ren.set_exec_mode(exec_mode)
    oren = orc.OracleRenderer(oscene, W, H, p, seed=77)
    for f in range(frames):
        cd = abi.CameraData.from_dict(cam.data())
        out_o = oren.frame(cd, f == 0).copy()
        out_g = ren.launch_frame(cd, f == 0).numpy()
        bad = rel_mismatch(out_g, out_o).any(axis=-1)
        print(f"{mode} {W}x{H} {exec_mode} frame {f}: mismatching pixels {int(bad.sum())}/{W * H}, "
              f"mean {out_o.mean():.5f} vs {out_g.mean():.5f}", flush=True)
        assert not bad.any()
    assert ren.stats().queueOverflow == 0
    return out_g


@pytest.mark.parametrize("exec_mode", ["wavefront", "megakernel"])
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("size", [(1, 1), (7, 5), (33, 17)])
def test_odd_resolutions(size, mode, exec_mode, orc, cx_runtime, cx_ctx):
    over = dict(gatherRadius=0.08) if mode == "conditional_restir" else None
    _run(orc, cx_runtime, cx_ctx, scenes.cornell_box(1, 8), mode, *size, exec_mode=exec_mode, over=over)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("max_len", [1, 2, 3])
def test_short_paths(max_len, mode, orc, cx_runtime, cx_ctx):
    over = dict(maxPathLen=max_len)
    if mode == "conditional_restir":   # the prefix may not be longer than the path
        over["gatherRadius"] = 0.08
        over["prefixLen"] = min(2, max_len)
    _run(orc, cx_runtime, cx_ctx, scenes.cornell_box(1, 8), mode, 24, 16, over=over)


@pytest.mark.parametrize("exec_mode", ["wavefront", "megakernel"])
@pytest.mark.parametrize("mode", MODES)
def test_empty_geometry(mode, exec_mode, orc, cx_runtime, cx_ctx):
    sc = scenes.cornell_box(1, 8)
    sc = dataclasses.replace(sc, indices=np.zeros((0, 3), np.uint32), tri_material=np.zeros(0, np.uint32))
    out = _run(orc, cx_runtime, cx_ctx, sc, mode, 16, 12, exec_mode=exec_mode,
               over=dict(gatherRadius=0.08) if mode == "conditional_restir" else None)
    assert np.isfinite(out).all()


def test_prefix_longer_than_path_is_rejected(cx_runtime, cx_ctx):
    """Conditional ReSTIR with prefixLen > maxPathLen is refused by cx_renderer_set_params
    (CX_ERR_INVALID_ARG) instead of running with an empty suffix range."""
    sc = scenes.cornell_box(1, 8)
    gscene = cx_runtime.Scene(cx_ctx, sc)
    p = abi.workload_params("conditional_restir", 8, 8, gscene.size)
    p.maxPathLen, p.prefixLen = 1, 2
    with pytest.raises(cx_runtime.CxError, match="prefixLen"):
        cx_runtime.Renderer(8, 8, cx_ctx, scenes.cornell_camera(8, 8), gscene, params=p, seed=1)
