"""Autotuner of the frame loop's launch settings.

For one scene and one view it times the frame under every combination of
  * BVH width (children per node: 2, 4, 8; cx_scene_desc.bvhWidth),
  * queue grid (blocks of the grid-stride traversal and step kernels; cx_renderer_set_queue_grid),
  * execution mode (wavefront queues or one kernel per pass; cx_renderer_set_exec_mode),
  * wavefront rounds per machine before the in-place tail kernel (cx_renderer_set_wave_rounds),
and keeps the fastest.  All settings give bit-identical images (the tests check every one), so
the choice is a pure performance decision.  Results are cached in a JSON file keyed by the
scene's name and triangle count, the resolution and the renderer type.

    python -m crisoptix_amd.tune --grid 10 --width 960 --height 540
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import abi, runtime, scenes

WIDTHS = (2, 4, 8)
QUEUE_GRIDS = (2048, 4096, 8192, 16384)
EXEC_MODES = ("wavefront", "megakernel")
WAVE_ROUNDS = (0, 1, 2, 3)


def _key(scene, w, h, params) -> str:
    return f"{scene.name}:{scene.num_triangles}:{w}x{h}:type{int(params.rendererType)}:len{int(params.maxPathLen)}"


def time_frames(ren, cam, frames: int, warmup: int = 2) -> float:
    """Median GPU ms per frame of an orbiting view."""
    for _ in range(warmup):
        cam.rotate_around_origin([0, 1, 0], 2 * math.pi / 64)
        ren.launch_frame(abi.CameraData.from_dict(cam.data()), True)
    runtime.synchronize()
    t0, t1 = runtime.Event(), runtime.Event()
    ms = []
    for _ in range(frames):
        cam.rotate_around_origin([0, 1, 0], 2 * math.pi / 64)
        cd = abi.CameraData.from_dict(cam.data())
        t0.record()
        ren.launch_frame(cd, True)
        t1.record()
        t1.synchronize()
        ms.append(t0.elapsed_ms(t1))
    return float(np.median(ms))


def tune(ctx, scene, w: int, h: int, params, camera_factory, frames: int = 5, gscenes: dict | None = None,
         cache_path: str | None = None, log=None) -> dict:
    """Returns {"bvh_width", "queue_grid", "wave_rounds", "exec", "ms", "table"} for the view;
    `gscenes` maps a BVH width to an already built runtime.Scene of `scene` (built here when
    missing).  The queue grid is swept at the default wave rounds and the wave rounds at the best
    queue grid of each width, so the table holds 4 + 4 wavefront rows and one megakernel row per
    width."""
    key = _key(scene, w, h, params)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            return cache[key]
    gscenes = gscenes if gscenes is not None else {}
    table = []
    best = None
    for width in WIDTHS:
        if width not in gscenes:
            gscenes[width] = runtime.Scene(ctx, scene, width=width)
        gs = gscenes[width]
        if gs.info.bvhWidth != width:
            continue   # the wide tree's stack bound did not fit; that width is the binary one
        def run(mode, grid=None, rounds=None):
            nonlocal best
            cam = camera_factory()
            ren = runtime.Renderer(w, h, ctx, cam, gs, params=params, seed=1)
            ren.set_exec_mode(mode)
            if grid:
                ren.set_queue_grid(grid)
            if rounds is not None:
                ren.set_wave_rounds(rounds)
            ms = time_frames(ren, cam, frames)
            row = {"bvh_width": width, "exec": mode, "queue_grid": grid, "wave_rounds": rounds, "ms": ms}
            table.append(row)
            if log:
                log(f"  {key}: {row}")
            if best is None or ms < best["ms"]:
                best = dict(row)
            del ren
            return ms

        grid_ms = {g: run("wavefront", g, None) for g in QUEUE_GRIDS}
        g_best = min(grid_ms, key=grid_ms.get)
        for rounds in WAVE_ROUNDS:
            run("wavefront", g_best, rounds)
        run("megakernel")
    best["table"] = table
    if cache_path:
        cache[key] = best
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, cache_path)
    return best


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=10)
    ap.add_argument("--width", type=int, default=960)
    ap.add_argument("--height", type=int, default=540)
    ap.add_argument("--frames", type=int, default=5)
    ap.add_argument("--cache")
    a = ap.parse_args(argv)
    ctx = runtime.Context(0)
    sc = scenes.object_grid(a.grid, seed=20 + a.grid, ground_quads=10 * a.grid)
    p = abi.workload_params("conditional_restir", a.width, a.height, sc.scene_size())
    p.maxPathLen = 6
    t0 = time.perf_counter()
    best = tune(ctx, sc, a.width, a.height, p, lambda: scenes.orbit_camera(sc, a.width, a.height), a.frames,
                cache_path=a.cache, log=lambda m: print(m, file=sys.stderr, flush=True))
    best["seconds"] = time.perf_counter() - t0
    print(json.dumps(best), flush=True)


if __name__ == "__main__":
    main()
