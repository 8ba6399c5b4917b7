"""Multi-view batch (cx_batch_frame): views of two scenes at uneven resolutions rendered on
the context's stream pool give, frame after frame, the same images bit for bit as the plain
per-view frames on the default stream; and every launch setting the autotuner chooses from
(BVH width, queue grid, wavefront rounds before the tail kernel, execution mode) renders the
same image."""
import math

import numpy as np
import pytest

from crisoptix_amd import abi, scenes

pytestmark = pytest.mark.gpu

SPECS = [(0, 64, 48), (1, 96, 56), (0, 40, 32), (1, 128, 72)]


def views(rt, ctx, gscenes, hosts):
    cams, rens = [], []
    for i, (si, w, h) in enumerate(SPECS):
        cam = scenes.orbit_camera(hosts[si], w, h)
        cam.rotate_around_origin([0, 1, 0], 0.7 * i)
        p = abi.workload_params("conditional_restir", w, h, gscenes[si].size)
        p.maxPathLen = 6
        rens.append(rt.Renderer(w, h, ctx, cam, gscenes[si], params=p, seed=50 + i))
        cams.append(cam)
    return cams, rens


def advance(cams):
    out = []
    for c in cams:
        c.rotate_around_origin([0, 1, 0], 2 * math.pi / 64)
        out.append(abi.CameraData.from_dict(c.data()))
    return out


def test_batch_equals_plain(cx_runtime, cx_ctx):
    hosts = [scenes.object_grid(2, seed=41, ground_quads=20, ico_subdiv=2),
             scenes.object_grid(3, seed=42, ground_quads=30, ico_subdiv=2)]
    gs = [cx_runtime.Scene(cx_ctx, h) for h in hosts]
    cp, rp = views(cx_runtime, cx_ctx, gs, hosts)
    cb, rb = views(cx_runtime, cx_ctx, gs, hosts)
    batch = cx_runtime.Batch(cx_ctx, rb)
    for f in range(4):
        for r, cd in zip(rp, advance(cp)):
            r.launch_frame(cd, True)
        batch.frame(advance(cb), [True] * len(rb))
        batch.sync()
        for i, (a, b) in enumerate(zip(rp, rb)):
            assert np.array_equal(a.output.numpy().view(np.uint32), b.output.numpy().view(np.uint32)), (f, i)
    assert any(r.output.numpy().mean() > 0 for r in rp)


def test_tuning_settings_render_the_same_image(cx_runtime, cx_ctx):
    host = scenes.object_grid(3, seed=43, ground_quads=30, ico_subdiv=2)
    W, H = 80, 48
    ref = None
    for width in (2, 4, 8):
        g = cx_runtime.Scene(cx_ctx, host, width=width)
        for mode, grid, rounds in (("wavefront", 1024, 0), ("wavefront", 8192, 1), ("wavefront", 4096, 3),
                                   ("megakernel", None, None)):
            cam = scenes.orbit_camera(host, W, H)
            p = abi.workload_params("conditional_restir", W, H, g.size)
            p.maxPathLen = 6
            r = cx_runtime.Renderer(W, H, cx_ctx, cam, g, params=p, seed=9)
            r.set_exec_mode(mode)
            if grid:
                r.set_queue_grid(grid)
            if rounds is not None:
                r.set_wave_rounds(rounds)
            for f in range(3):
                img = r.launch_frame(advance([cam])[0], True).numpy()
            if ref is None:
                ref = img
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (width, mode, grid, rounds)
