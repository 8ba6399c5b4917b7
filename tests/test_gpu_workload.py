"""The 1080p Conditional ReSTIR workload itself (the benchmark's scene and view: 16 x 16 object
grid under the tent of 16 traced lights, 1.06 M triangles, k = 8 gathered suffixes), where the
oracle is too slow to run: two implementations of the frame must agree bit for bit.  The
wavefront mode at its defaults (queue traversal with per-wavefront ranges, tail kernels, the
final gather's pixel list and record-major neighbour copy), the wavefront mode with every round
as traversal + step, and the one-kernel-per-pass megakernel mode render the same two orbiting
frames; the frames gather from all 8 neighbours, no queue overflows, and the neighbour
statistics (total, fewest, most per gathering pixel; summed per wavefront in the wavefront
mode, per pixel in the megakernel) agree."""
import math

import numpy as np
import pytest

from crisoptix_amd import abi, scenes

pytestmark = pytest.mark.gpu

W, H = 1920, 1080


def test_workload_frames_agree_across_execution_modes(cx_runtime, cx_ctx):
    sc = scenes.object_grid(16, 2)
    gs = cx_runtime.Scene(cx_ctx, sc)
    print(f"workload scene: {sc.num_triangles} triangles", flush=True)
    images, nstats = {}, {}
    for mode, rounds in (("wavefront", None), ("wavefront", 0), ("megakernel", None)):
        cam = scenes.orbit_camera(sc, W, H)
        p = abi.workload_params("conditional_restir", W, H, gs.size)
        p.trackPrefixStats = 1
        ren = cx_runtime.Renderer(W, H, cx_ctx, cam, gs, params=p, seed=77)
        ren.set_exec_mode(mode)
        if rounds is not None:
            ren.set_wave_rounds(rounds)
        frames = []
        for f in range(2):
            if f:
                cam.rotate_around_origin([0, 1, 0], 2 * math.pi / 64)
            frames.append(ren.launch_frame(abi.CameraData.from_dict(cam.data()), True).numpy().copy())
        st = ren.stats()
        assert st.queueOverflow == 0
        nstats[(mode, rounds)] = (st.totalNeighCount, st.minNeighCount, st.maxNeighCount)
        if mode == "wavefront" and rounds is None:
            _, _, counts = ren.neighbors()
            assert (counts == 8).sum() > 100_000, "gathering pixels should find all 8 neighbours"
        images[(mode, rounds)] = frames
        print(f"{mode} rounds={rounds}: mean radiance {np.mean(frames[-1]):.5f}", flush=True)
        del ren
    assert len(set(nstats.values())) == 1, nstats
    assert nstats[("wavefront", None)][2] == 8
    ref = images[("wavefront", None)]
    for key, frames in images.items():
        for f, (a, b) in enumerate(zip(frames, ref)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (key, f)
