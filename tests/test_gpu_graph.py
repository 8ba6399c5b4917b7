"""Graph mode (cx_renderer_set_graph_mode): each frame is one replay of a hipGraph captured
from the eager frame, per front-buffer parity, with the camera, random number and frame index
read from device memory.  Every replay is checked bit for bit against an eager renderer of the
same seed, with that renderer's eager frames and host synchronisations between the replays;
parameter changes, a reset, the megakernel mode and a multi-view batch recapture or replay as
the rules say.  The frame's entry sort and tree build are memset-free kernels, so a capture
records kernel and copy nodes only."""
import math

import numpy as np
import pytest

from crisoptix_amd import abi, scenes
from parity_util import small_grid_scene

pytestmark = pytest.mark.gpu


def _pair(rt, ctx, mode, W, H, exec_mode="wavefront"):
    sc = scenes.cornell_box(1, 8) if mode != "conditional_restir" else small_grid_scene()
    gs = rt.Scene(ctx, sc)
    p = abi.workload_params(mode, W, H, gs.size)
    if mode == "conditional_restir":
        p.gatherRadius = 0.3
    make_cam = (lambda: scenes.cornell_camera(W, H)) if mode != "conditional_restir" else \
        (lambda: scenes.orbit_camera(sc, W, H))
    ce, cg = make_cam(), make_cam()
    e = rt.Renderer(W, H, ctx, ce, gs, params=p, seed=77)
    g = rt.Renderer(W, H, ctx, cg, gs, params=p, seed=77)
    e.set_exec_mode(exec_mode)
    g.set_exec_mode(exec_mode)
    g.set_graph_mode(True)
    return ce, cg, e, g, p


def _step(ce, cg, e, g, orbit, f, tag):
    if orbit:
        ce.rotate_around_origin([0, 1, 0], 2 * math.pi / 64)
        cg.rotate_around_origin([0, 1, 0], 2 * math.pi / 64)
    changed = f == 0 or orbit
    oe = e.launch_frame(abi.CameraData.from_dict(ce.data()), changed).numpy()
    og = g.launch_frame(abi.CameraData.from_dict(cg.data()), changed).numpy()
    assert np.array_equal(oe.view(np.uint32), og.view(np.uint32)), (tag, f)
    return oe


@pytest.mark.parametrize("mode,exec_mode", [("path_tracer", "wavefront"), ("restir_pt_cornell", "wavefront"),
                                            ("conditional_restir", "wavefront"),
                                            ("conditional_restir", "megakernel")])
def test_graph_replays_equal_eager_frames(mode, exec_mode, cx_runtime, cx_ctx):
    print(f"graph {mode} {exec_mode}", flush=True)
    W, H = 48, 32
    ce, cg, e, g, p = _pair(cx_runtime, cx_ctx, mode, W, H, exec_mode)
    orbit = mode == "conditional_restir"
    for f in range(6):
        out = _step(ce, cg, e, g, orbit, f, "first")
    assert g.graph_captures() == 2          # one capture per front-buffer parity
    assert np.isfinite(out).all() and out.mean() > 0
    # a parameter change recaptures both parities
    p2 = abi.LaunchParams.from_buffer_copy(p)
    p2.maxPathLen = max(2, int(p.maxPathLen) - 1)
    e.set_params(p2)
    g.set_params(p2)
    for f in range(6, 10):
        _step(ce, cg, e, g, orbit, f, "params")
    assert g.graph_captures() == 4
    # a reset restarts the parity sequence; the round bounds shrink, so both recapture again
    e.reset()
    g.reset()
    for f in range(10, 14):
        _step(ce, cg, e, g, orbit, f, "reset")
    assert g.graph_captures() == 6


def test_graph_batch_equals_eager(cx_runtime, cx_ctx):
    hosts = [scenes.object_grid(2, seed=41, ground_quads=20, ico_subdiv=2),
             scenes.object_grid(3, seed=42, ground_quads=30, ico_subdiv=2)]
    gs = [cx_runtime.Scene(cx_ctx, h) for h in hosts]
    specs = [(0, 64, 48), (1, 96, 56), (0, 40, 32), (1, 128, 72)]

    def views(graph):
        cams, rens = [], []
        for i, (si, w, h) in enumerate(specs):
            cam = scenes.orbit_camera(hosts[si], w, h)
            cam.rotate_around_origin([0, 1, 0], 0.7 * i)
            p = abi.workload_params("conditional_restir", w, h, gs[si].size)
            p.maxPathLen = 6
            r = cx_runtime.Renderer(w, h, cx_ctx, cam, gs[si], params=p, seed=50 + i)
            r.set_graph_mode(graph)
            rens.append(r)
            cams.append(cam)
        return cams, rens

    def advance(cams):
        out = []
        for c in cams:
            c.rotate_around_origin([0, 1, 0], 2 * math.pi / 64)
            out.append(abi.CameraData.from_dict(c.data()))
        return out

    cp, rp = views(False)
    cb, rb = views(True)
    batch = cx_runtime.Batch(cx_ctx, rb)
    for f in range(5):
        for r, cd in zip(rp, advance(cp)):
            r.launch_frame(cd, True)
        batch.frame(advance(cb), [True] * len(rb))
        batch.sync()
        for i, (a, b) in enumerate(zip(rp, rb)):
            assert np.array_equal(a.output.numpy().view(np.uint32), b.output.numpy().view(np.uint32)), (f, i)
    assert all(r.graph_captures() == 2 for r in rb)


@pytest.mark.parametrize("graph", [False, True])
def test_final_phase_overlap_is_invisible(graph, cx_runtime, cx_ctx, orc):
    """The entry store and tree build on the side stream (the default) against the sequential
    order and the oracle: images, neighbour lists and the prefix statistics (trackPrefixStats)
    agree bit for bit, frame after frame."""
    print(f"overlap graph={graph}", flush=True)
    sc = scenes.cornell_box(1, 8)          # a closed scene: many valid entries and neighbours
    W, H = 48, 32
    gs = cx_runtime.Scene(cx_ctx, sc)
    p = abi.workload_params("conditional_restir", W, H, gs.size)
    p.gatherRadius = 0.08
    p.trackPrefixStats = 1
    cams = [scenes.cornell_camera(W, H) for _ in range(3)]
    on = cx_runtime.Renderer(W, H, cx_ctx, cams[0], gs, params=p, seed=5)
    off = cx_runtime.Renderer(W, H, cx_ctx, cams[1], gs, params=p, seed=5)
    off.set_overlap(False)
    on.set_graph_mode(graph)
    oren = orc.OracleRenderer(orc.OracleScene(sc), W, H, p, seed=5)
    for f in range(5):
        cds = [abi.CameraData.from_dict(c.data()) for c in cams]
        a = on.launch_frame(cds[0], f == 0).numpy()
        b = off.launch_frame(cds[1], f == 0).numpy()
        o = oren.frame(cds[2], f == 0)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f
        assert np.array_equal(a.view(np.uint32), o.view(np.uint32)), f
        ia, da, ca = on.neighbors()
        ib, db, cb = off.neighbors()
        io, do, co = oren.neighbors()
        assert np.array_equal(ca, cb) and np.array_equal(ca, co), f
        for q in np.flatnonzero(ca):
            n = ca[q]
            assert np.array_equal(ia[q, :n], io[q, :n]) and np.array_equal(ib[q, :n], io[q, :n])
            assert np.array_equal(da[q, :n].view(np.uint32), do[q, :n].view(np.uint32))
        sa, sb = on.stats(), off.stats()
        assert (sa.totalNeighCount, sa.minNeighCount, sa.maxNeighCount) == \
               (sb.totalNeighCount, sb.minNeighCount, sb.maxNeighCount), f
    assert ca.sum() > 50


def test_caller_graph_records_conditional_restir_frame(cx_runtime, cx_ctx):
    """A Conditional ReSTIR frame recorded into the caller's graph (torch.cuda.graph on the
    caller's stream): inside a caller's capture the frame keeps to that stream (no side-stream
    fork) and allocates nothing, so each replay must equal an eager renderer's next frame bit for
    bit (device-random frames, accumulation off, static camera), with eager frames and host
    synchronisations between the replays.  Graph mode inside a caller's capture, and a frame
    whose buffers would first be allocated there, are refused with an error before anything is
    enqueued."""
    import torch
    W, H = 32, 24
    sc = scenes.cornell_box(1, 8)
    gs = cx_runtime.Scene(cx_ctx, sc)
    p = abi.workload_params("conditional_restir", W, H, gs.size)
    p.gatherRadius = 0.08
    cam = scenes.cornell_camera(W, H)
    cd = abi.CameraData.from_dict(cam.data())
    eager = cx_runtime.Renderer(W, H, cx_ctx, cam, gs, params=p, seed=5)
    rec = cx_runtime.Renderer(W, H, cx_ctx, cam, gs, params=p, seed=5)
    for r in (eager, rec):
        r.set_device_random(True, 11)
    ref = [eager.launch_frame(cd, f == 0).numpy().copy() for f in range(5)]
    assert np.isfinite(ref[-1]).all() and ref[-1].mean() > 0
    stream = torch.cuda.Stream()
    out = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    dst = cx_runtime.DevicePtr(out.data_ptr())
    with torch.cuda.stream(stream):
        rec.launch_frame(cd, True, output=dst, stream=stream.cuda_stream)   # eager: allocates every buffer
    stream.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref[0].view(np.uint32))
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        rec.launch_frame(cd, False, output=dst, stream=stream.cuda_stream)
    side = torch.ones(1024, device="cuda")
    for f in range(1, 5):
        graph.replay()
        stream.synchronize()
        got = out.cpu().numpy()
        print(f"caller graph replay {f}: mean {got.mean():.5f}", flush=True)
        assert np.array_equal(got.view(np.uint32), ref[f].view(np.uint32)), f
        side.mul_(1.0001)                     # eager work and a synchronisation between replays
        torch.cuda.synchronize()
    # refused inside a capture: graph mode, and a renderer that never ran a frame
    fresh = cx_runtime.Renderer(W, H, cx_ctx, cam, gs, params=p, seed=5)
    rec.set_graph_mode(True)
    for r, what in ((rec, "graph mode"), (fresh, "eagerly first")):
        g2 = torch.cuda.CUDAGraph()
        with pytest.raises(cx_runtime.CxError, match=what):
            with torch.cuda.graph(g2, stream=stream):
                r.launch_frame(cd, False, output=dst, stream=stream.cuda_stream)
