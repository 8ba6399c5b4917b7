"""Multi-view batch: 16 views of uneven resolution over 3 scenes, plain per-view frames against
one batched enqueue (cx_batch_frame), with the autotuner's settings.

Workload: three object-grid scenes built like the 1080p benchmark scene with grids 4, 10 and 24
(icospheres of subdivision 4 and boxes on a 10·grid x 10·grid-quad ground; about 64 k, 404 k
and 2.3 M triangles), 16 views with seeded resolutions between 320 x 240 and 1920 x 1080 (one
seeded parameter t in [0, 1] per view: (320 + 1600 t) x (240 + 840 t), rounded to multiples of
8), each view on a seeded scene and orbit angle; Conditional ReSTIR, max path length 6, eight
gathered suffixes.

plain:   for each view, cx_renderer_frame on the default stream, one after another;
batched: cx_batch_frame, every view on its own stream of the context's pool, so the small views'
         launch-bound passes overlap with the others' work; with --graph each view's frame is
         one replay of its captured hipGraph (cx_renderer_set_graph_mode).
The HIP runtime spreads a process's streams over GPU_MAX_HW_QUEUES hardware queues (HIP's default
is 4); the 16 per-view streams overlap best on 16 queues, so the command line sets
GPU_MAX_HW_QUEUES to --hw-queues (default 16; 4 gives HIP's default) before the first HIP call,
over any value in the environment.  Both legs run in the one process, so both see the same queues.
Both legs run without the final-phase side stream (cx_renderer_set_overlap): over these 16 views it
gained nothing in either leg at 4 queues, and at 16 queues it slowed the plain leg by a tenth.
Both run the same renderers' code on the same seeds: the per-view images are compared bit for
bit every frame (equal images, so equal rel-MSE against any reference).

    python -m crisoptix_amd.batch [--frames 8] [--tune] [--graph] [--hw-queues 16]
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

GRIDS = (4, 10, 24)


def build_scenes():
    return [scenes.object_grid(g, seed=20 + g, ground_quads=10 * g) for g in GRIDS]


def view_specs(n: int = 16, seed: int = 31):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        t = float(rng.random())
        w = int(round((320 + 1600 * t) / 8)) * 8
        h = int(round((240 + 840 * t) / 8)) * 8
        out.append({"scene": int(rng.integers(0, len(GRIDS))), "width": w, "height": h,
                    "angle": float(rng.uniform(0, 2 * math.pi))})
    return out


def make_params(w, h, size):
    p = abi.workload_params("conditional_restir", w, h, size)
    p.maxPathLen = 6
    return p


class Views:
    def __init__(self, ctx, hosts, gscenes, specs, settings=None, graph=False, overlap=True):
        self.cams, self.rens, self.specs = [], [], specs
        for i, s in enumerate(specs):
            sc, gs = hosts[s["scene"]], gscenes[s["scene"]]
            cam = scenes.orbit_camera(sc, s["width"], s["height"])
            cam.rotate_around_origin([0, 1, 0], s["angle"])
            ren = runtime.Renderer(s["width"], s["height"], ctx, cam, gs,
                                   params=make_params(s["width"], s["height"], gs.size), seed=1000 + i)
            if settings:
                st = settings[i]
                ren.set_exec_mode(st["exec"])
                if st.get("queue_grid"):
                    ren.set_queue_grid(st["queue_grid"])
                if st.get("wave_rounds") is not None:
                    ren.set_wave_rounds(st["wave_rounds"])
            ren.set_graph_mode(graph)
            ren.set_overlap(overlap)
            self.cams.append(cam)
            self.rens.append(ren)

    def advance(self):
        out = []
        for cam in self.cams:
            cam.rotate_around_origin([0, 1, 0], 2 * math.pi / 64)
            out.append(abi.CameraData.from_dict(cam.data()))
        return out


def run(a) -> dict:
    ctx = runtime.Context(0)
    t0 = time.perf_counter()
    hosts = build_scenes()
    specs = view_specs(a.views)
    settings = None
    tuning = None
    widths = {i: 4 for i in range(len(GRIDS))}
    if a.tune:
        from . import tune
        tuning = {}
        for si, sc in enumerate(hosts):
            # tune each scene at the median resolution of its views
            vs = [s for s in specs if s["scene"] == si] or specs[:1]
            m = sorted(vs, key=lambda s: s["width"])[len(vs) // 2]
            w, h = m["width"], m["height"]
            best = tune.tune(ctx, sc, w, h, make_params(w, h, sc.scene_size()),
                             lambda sc=sc, w=w, h=h: scenes.orbit_camera(sc, w, h), frames=a.tune_frames,
                             cache_path=a.tune_cache,
                             log=lambda msg: print(msg, file=sys.stderr, flush=True))
            tuning[GRIDS[si]] = {k: v for k, v in best.items() if k != "table"}
            widths[si] = best["bvh_width"]
        settings = [tuning[GRIDS[s["scene"]]] for s in specs]
    gscenes = [runtime.Scene(ctx, sc, width=widths[i]) for i, sc in enumerate(hosts)]
    t_build = time.perf_counter() - t0
    plain = Views(ctx, hosts, gscenes, specs, settings, overlap=False)
    batched = Views(ctx, hosts, gscenes, specs, settings, graph=a.graph, overlap=False)
    batch = runtime.Batch(ctx, batched.rens)
    P = sum(s["width"] * s["height"] for s in specs)
    mismatches = 0
    tp, tb = [], []
    for f in range(a.warmup + a.frames):
        cams = plain.advance()
        runtime.synchronize()
        t1 = time.perf_counter()
        for ren, cd in zip(plain.rens, cams):
            ren.launch_frame(cd, True)
        runtime.synchronize()
        t2 = time.perf_counter()
        cams_b = batched.advance()
        runtime.synchronize()
        t3 = time.perf_counter()
        batch.frame(cams_b, [True] * len(cams_b))
        batch.sync()
        t4 = time.perf_counter()
        if f >= a.warmup:
            tp.append(t2 - t1)
            tb.append(t4 - t3)
        for rp, rb in zip(plain.rens, batched.rens):
            if not np.array_equal(rp.output.numpy().view(np.uint32), rb.output.numpy().view(np.uint32)):
                mismatches += 1
    ms_plain = 1000 * float(np.median(tp))
    ms_batch = 1000 * float(np.median(tb))
    return {
        "workload": f"batch_{a.views}_views_objgrid{'-'.join(map(str, GRIDS))}_conditional_restir_len6",
        "triangles": [sc.num_triangles for sc in hosts],
        "views": [{"scene_grid": GRIDS[s["scene"]], "resolution": [s["width"], s["height"]]} for s in specs],
        "pixels_per_batch_frame": P,
        "frames": a.frames,
        "ms_per_batch_frame_plain": ms_plain,
        "ms_per_batch_frame_batched": ms_batch,
        "speedup": ms_plain / ms_batch,
        "mpix_per_s_plain": P / ms_plain / 1e3,
        "mpix_per_s_batched": P / ms_batch / 1e3,
        "image_mismatches": mismatches,
        "tuned": tuning,
        "graph_mode": bool(a.graph),
        "final_phase_overlap": False,
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "4 (HIP default)"),
        "build_seconds": t_build,
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--views", type=int, default=16)
    ap.add_argument("--frames", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tune", action="store_true")
    ap.add_argument("--graph", action="store_true", help="batched views replay captured frame graphs")
    ap.add_argument("--tune-frames", type=int, default=4)
    ap.add_argument("--tune-cache")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (1-32, over the environment's value); "
                         "takes effect only before the process's first HIP call")
    a = ap.parse_args(argv)
    if not 1 <= a.hw_queues <= 32:
        ap.error("--hw-queues must be between 1 and 32")
    os.environ["GPU_MAX_HW_QUEUES"] = str(a.hw_queues)
    print(json.dumps(run(a)), flush=True)


if __name__ == "__main__":
    main()
