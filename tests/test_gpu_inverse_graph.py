"""The device-resident inverse-rendering step (crisoptix_amd.inverse.DeviceStep), run eagerly and
recorded once with torch.cuda.graph: frames draw their random numbers on the device
(cx_renderer_set_device_random) and materials come from device tensors
(cx_scene_set_materials_device), so each replay must equal the eager step that starts from the
same state.  With learning rate 0 the parameters stay fixed and the steps differ only by their
random numbers: every replay's residual image equals the eager one bit for bit, and its
gradients up to the summation order of the adjoint's double-precision atomics.  Also: materials
written from device memory render exactly like the host upload, and device-random frames depend
only on their seed."""
import numpy as np
import pytest

from crisoptix_amd import abi, inverse, scenes

pytestmark = pytest.mark.gpu

W, H, VIEWS = 64, 48, 2


def _setup(rt, ctx):
    import math
    sc = scenes.inverse_scene()
    gs = rt.Scene(ctx, sc)
    views, targets, clips = [], [], []
    for v in range(VIEWS):
        cam = scenes.orbit_camera(sc, W, H)
        cam.rotate_around_origin([0, 1, 0], 2 * math.pi * v / VIEWS)
        ren = rt.Renderer(W, H, ctx, cam, gs, params=inverse._params(gs.size, W, H), seed=10 + v)
        cd = abi.CameraData.from_dict(cam.data())
        img = ren.launch_frame(cd, True).numpy().copy()
        views.append((ren, cd))
        targets.append(img)
        clips.append(inverse._clip_level(img))
    rng = np.random.default_rng(5)
    M = sc.materials.shape[0]
    return sc, gs, views, targets, clips, rng.uniform(0.3, 0.7, (M, 3)), rng.uniform(0.3, 0.7, M)


def _record(ds):
    ds.sync()
    return ([i.cpu().numpy().copy() for i in ds.imgs], ds.gA.cpu().numpy().copy(), ds.gR.cpu().numpy().copy())


def test_graph_replays_equal_eager_steps(cx_runtime, cx_ctx):
    import torch
    sc, gs, views, targets, clips, al, ro = _setup(cx_runtime, cx_ctx)
    tt = [torch.from_numpy(t) for t in targets]
    eager = inverse.DeviceStep(gs, views, tt, clips, sc.materials, al, ro, lr=0.0, seed=3)
    ref = []
    for _ in range(4):
        eager.step()
        ref.append(_record(eager))
    graph = inverse.DeviceStep(gs, views, tt, clips, sc.materials, al, ro, lr=0.0, seed=3)
    graph.step()
    got = [_record(graph)]
    graph.capture()
    for _ in range(3):
        graph.replay()
        got.append(_record(graph))
    for k, ((ri, ra, rr), (gi, ga, gr)) in enumerate(zip(ref, got)):
        for a, b in zip(ri, gi):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
        assert np.abs(ra).max() > 0
        np.testing.assert_allclose(ga, ra, rtol=1e-9, atol=1e-15 * np.abs(ra).max())
        np.testing.assert_allclose(gr, rr, rtol=1e-9, atol=1e-15 * np.abs(rr).max())
    # consecutive steps draw different random numbers
    assert not np.array_equal(ref[0][0][0], ref[1][0][0])


def test_graph_steps_follow_eager_parameters(cx_runtime, cx_ctx):
    import torch
    sc, gs, views, targets, clips, al, ro = _setup(cx_runtime, cx_ctx)
    tt = [torch.from_numpy(t) for t in targets]
    eager = inverse.DeviceStep(gs, views, tt, clips, sc.materials, al, ro, lr=1e-2, seed=4)
    for _ in range(4):
        eager.step()
    graph = inverse.DeviceStep(gs, views, tt, clips, sc.materials, al, ro, lr=1e-2, seed=4)
    graph.step()
    graph.capture()
    for _ in range(3):
        graph.replay()
    eager.sync()
    graph.sync()
    ea, er = eager.albedo.cpu().numpy(), eager.rough.cpu().numpy()
    ga, gr = graph.albedo.cpu().numpy(), graph.rough.cpu().numpy()
    assert not np.allclose(ea, al)   # the parameters moved
    np.testing.assert_allclose(ga, ea, rtol=0, atol=1e-5)
    np.testing.assert_allclose(gr, er, rtol=0, atol=1e-5)


def test_device_materials_equal_host_upload(cx_runtime, cx_ctx):
    """cx_scene_set_materials_device writes the same tables as cx_scene_set_materials: frames
    rendered after either upload of perturbed materials are bit-identical."""
    import torch
    sc = scenes.inverse_scene()
    mats = scenes.perturbed_materials(sc.materials, seed=8)
    out = []
    for device in (False, True):
        gs = cx_runtime.Scene(cx_ctx, sc)
        if device:
            rows = torch.from_numpy(np.ascontiguousarray(mats, np.float32)).cuda()
            gs.set_materials_device(rows.data_ptr(), rows.shape[0], 3, None)
            torch.cuda.synchronize()
        else:
            gs.set_materials(mats, 3)
        cam = scenes.orbit_camera(sc, W, H)
        ren = cx_runtime.Renderer(W, H, cx_ctx, cam, gs, params=inverse._params(gs.size, W, H), seed=21)
        cd = abi.CameraData.from_dict(cam.data())
        out.append([ren.launch_frame(cd, f == 0).numpy().copy() for f in range(2)])
    for a, b in zip(*out):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_device_random_frames_reproducible(cx_runtime, cx_ctx):
    """Device-random frames depend only on the seed: two renderers seeded alike render the same
    frames, and consecutive frames differ."""
    sc = scenes.inverse_scene()
    gs = cx_runtime.Scene(cx_ctx, sc)
    frames = []
    for _ in range(2):
        cam = scenes.orbit_camera(sc, W, H)
        ren = cx_runtime.Renderer(W, H, cx_ctx, cam, gs, params=inverse._params(gs.size, W, H), seed=1)
        ren.set_device_random(True, 99)
        cd = abi.CameraData.from_dict(cam.data())
        frames.append([ren.launch_frame(cd, f == 0).numpy().copy() for f in range(3)])
    for a, b in zip(*frames):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert not np.array_equal(frames[0][1], frames[0][2])
