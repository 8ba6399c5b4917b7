"""The multi-view batch and the inverse-rendering loop at their workload sizes.

Batch: 16 views with seeded resolutions between 320 x 240 and 1920 x 1080 over the three object
grids of 64 k, 404 k and 2.27 M triangles (crisoptix_amd/batch.py), Conditional ReSTIR at max
path length 6 with 8 gathered suffixes; the batched enqueue (every view on its own stream, each
view's frame replayed from its captured hipGraph) gives every view's image bit for bit as the
plain per-view frames, frame after frame, and no wavefront queue overflows.

Inverse rendering: 8 views at 640 x 480 of the 407 k-triangle grid with 256 object materials +
ground (crisoptix_amd/inverse.py optimize_scene), Adam at lr 1e-2 on albedo and roughness from
a seeded U[0.3, 0.7] start.  The step recorded once as a graph and replayed follows the eagerly
enqueued step (parameters within 1e-5 after 4 steps: the adjoint's float64 atomics add in a
different order), and the relative parameter error on the materials the views observe falls at
every checkpoint of the graph-replayed loop."""
import math

import numpy as np
import pytest

from crisoptix_amd import abi, batch, inverse, scenes

pytestmark = pytest.mark.gpu


def test_batch_16_views_equal_plain_frames(cx_runtime, cx_ctx):
    hosts = batch.build_scenes()
    print("batch scenes:", [h.num_triangles for h in hosts], flush=True)
    specs = batch.view_specs(16)
    gscenes = [cx_runtime.Scene(cx_ctx, h) for h in hosts]
    plain = batch.Views(cx_ctx, hosts, gscenes, specs)
    batched = batch.Views(cx_ctx, hosts, gscenes, specs, graph=True)
    b = cx_runtime.Batch(cx_ctx, batched.rens)
    for f in range(3):
        for ren, cd in zip(plain.rens, plain.advance()):
            ren.launch_frame(cd, True)
        b.frame(batched.advance(), [True] * len(specs))
        b.sync()
        for i, (rp, rb) in enumerate(zip(plain.rens, batched.rens)):
            a, c = rp.output.numpy(), rb.output.numpy()
            assert np.array_equal(a.view(np.uint32), c.view(np.uint32)), (f, i, specs[i])
        print(f"batch frame {f}: 16 views equal, mean radiance "
              f"{np.mean([r.output.numpy().mean() for r in plain.rens]):.4f}", flush=True)
    assert all(r.stats().queueOverflow == 0 for r in plain.rens + batched.rens)
    assert sum(s["width"] * s["height"] for s in specs) > 5_000_000
    assert {s["scene"] for s in specs} == {0, 1, 2}


W, H, VIEWS = 640, 480, 8


def _views(rt, ctx, sc, gs):
    out = []
    for v in range(VIEWS):
        cam = scenes.orbit_camera(sc, W, H)
        cam.rotate_around_origin([0, 1, 0], 2 * math.pi * v / VIEWS)
        ren = rt.Renderer(W, H, ctx, cam, gs, params=inverse._params(gs.size, W, H), seed=100 + v)
        out.append((ren, abi.CameraData.from_dict(cam.data())))
    return out


def test_inverse_rendering_8_views_graph_follows_eager_and_converges(cx_runtime, cx_ctx):
    import torch
    sc = inverse.optimize_scene()
    truth = sc.materials.astype(np.float32)
    M = truth.shape[0]
    print(f"inverse scene: {sc.num_triangles} triangles, {M} materials", flush=True)
    assert sc.num_triangles > 400_000 and M == 257
    gs = cx_runtime.Scene(cx_ctx, sc)
    views = _views(cx_runtime, cx_ctx, sc, gs)
    targets = [inverse._accumulate(ren, cd, 32) for ren, cd in views]
    clips = [inverse._clip_level(t) for t in targets]
    tt = [torch.from_numpy(t) for t in targets]
    seen = np.zeros(M, np.int64)
    for ren, _ in views:
        prim = ren.prefix(ren.state().frontBufferIdx)["primary_prim"]
        seen += np.bincount(sc.tri_material[prim[prim != 0xFFFFFFFF]], minlength=M)
    observed = seen >= 200
    rng = np.random.default_rng(12)
    al0 = rng.uniform(0.3, 0.7, (M, 3)).astype(np.float32)
    ro0 = rng.uniform(0.3, 0.7, M).astype(np.float32)

    eager = inverse.DeviceStep(gs, views, tt, clips, truth, al0, ro0, lr=1e-2, seed=4)
    for _ in range(4):
        eager.step()
    eager.sync()
    ea, er = eager.albedo.cpu().numpy(), eager.rough.cpu().numpy()
    graph = inverse.DeviceStep(gs, views, tt, clips, truth, al0, ro0, lr=1e-2, seed=4)
    graph.step()
    graph.capture()
    for _ in range(3):
        graph.replay()
    graph.sync()
    assert not np.allclose(ea, al0)
    np.testing.assert_allclose(graph.albedo.cpu().numpy(), ea, rtol=0, atol=1e-5)
    np.testing.assert_allclose(graph.rough.cpu().numpy(), er, rtol=0, atol=1e-5)

    def rel_err():
        al, ro = graph.albedo.cpu().numpy()[observed], graph.rough.cpu().numpy()[observed]
        return (float(np.linalg.norm(al - truth[observed, 0:3]) / np.linalg.norm(truth[observed, 0:3])),
                float(np.linalg.norm(ro - truth[observed, 9]) / np.linalg.norm(truth[observed, 9])))

    errs = [rel_err()]
    for _ in range(CHECKPOINTS):
        for _ in range(STEPS_PER_CHECKPOINT):
            graph.replay()
        graph.sync()
        errs.append(rel_err())
    print(f"{int(observed.sum())} observed materials; rel error (albedo, roughness) every "
          f"{STEPS_PER_CHECKPOINT} steps: {[(round(a, 4), round(r, 4)) for a, r in errs]}", flush=True)
    for (a0, r0), (a1, r1) in zip(errs, errs[1:]):
        assert a1 < a0
        assert r1 < r0


CHECKPOINTS, STEPS_PER_CHECKPOINT = 5, 20
