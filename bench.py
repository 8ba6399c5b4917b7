"""Benchmark of the Conditional ReSTIR frame loop on one MI355X.

Workload: 1920x1080 Conditional ReSTIR (prefix/suffix reservoirs, entry-tree neighbour search,
final gather with 8 suffixes per pixel) on the ~1.06 M-triangle object grid (16 x 16 seeded
icospheres of subdivision 4 and boxes, scale U[0.3, 1.5], random rotation, on a 200 x 200-quad
ground plane, GGX roughness U[0.05, 1], albedo U[0.1, 0.9]) lit by 16 emitter triangles of
radiance lognormal(2, 0.5) that form a tent over the grid and are traced as geometry (closest-hit
rays only), max path length 8, camera inside the tent orbiting by 2*pi/64 per frame.  Synthetic,
seeded scene data; every frame runs the full pass sequence of Renderer::LaunchFrame.

One step = one frame.  ms_per_step = the median over the timed frames (frames W+1 .. W+K) of
HIP-event time around one frame; value = counted rays (closest-hit + shadow) of the timed frames
per second of their summed time.  roofline: the dominant kernel is the ray traversal
(k_q_trace), timed by HIP events around each of its launches on the frame's stream; its
algorithmic bytes are 48 B per closest-hit ray (32 B ray + 16 B hit) and 36 B per shadow ray
(32 B ray + 4 B flag).  The line reports the workload's reuse density (valid suffixes, prefix
entries, neighbours per gathering pixel) and exits with status 3 if a wavefront queue ever
overflowed.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/frame + Mrays/s at 1080p Conditional ReSTIR; rel-MSE vs FP64 truth"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes, from the reference's record sizes: per pixel of the final gather without
# its neighbours (primary interaction 40, RNG 32, neighbour list write 64 + read 128, output
# 24) and of the whole frame without them (6,704 - 2,432), plus 2 x 152 B per neighbour found
# (its suffix reservoir read by the validity and the shift loops), counted from the neighbour
# statistics of each frame; per ray 48 B (closest hit) or 36 B (shadow).
FINAL_GATHER_STATE_BYTES_PER_PX = 288
FRAME_STATE_BYTES_PER_PX = 4272
NEIGHBOR_BYTES = 2 * 152
CLOSEST_RAY_BYTES = 48
SHADOW_RAY_BYTES = 36
WORKLOAD_SEED = 2


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--grid", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-scale", type=int, default=4,
                    help="CPU baseline renders (W/scale) x (H/scale) pixels of the same workload")
    ap.add_argument("--cpu-frames", type=int, default=6, help="measured CPU frames (after 2 warm-up frames)")
    ap.add_argument("--exec", choices=["wavefront", "megakernel"], default="wavefront",
                    help="execution mode of the HIP passes (identical results)")
    ap.add_argument("--relmse-frames", type=int, default=4096,
                    help="accumulated frames of the rel-MSE measurement (0 = skip)")
    ap.add_argument("--bvh-width", type=int, choices=[2, 4, 8], default=4, help="children per BVH node")
    ap.add_argument("--bvh-max-leaf", type=int, default=0,
                    help="most triangles per BVH leaf (0: the library's default, 2)")
    ap.add_argument("--bvh-builder", choices=["host", "device"], default="host",
                    help="binned SAH on the host or PLOC on the GPU (same hits, different trees)")
    ap.add_argument("--wave-rounds", type=int, default=None,
                    help="trace/step rounds of every machine before the in-place tail kernel (0: all rounds; "
                         "default: 3 for suffix walks, 2 for the other machines)")
    ap.add_argument("--queue-grid", type=int, default=0,
                    help="most blocks of the grid-stride queue kernels (0: the library's default, 8192)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="run the entry build and neighbour query in sequence instead of beside the gather walks")
    return ap.parse_args()


def traffic_from_profiles(kernel, width):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary of this
    workload (FETCH_SIZE and WRITE_SIZE, each scaled by a copy kernel of known size), if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_1080p_bvh{width}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split()[-1])
    return None


def rel_mse_vs_truth(runtime, ctx, gscene, scene, frames, exec_mode):
    """rel-MSE = mean over pixels and channels of (I - I_ref)^2 / (I_ref^2 + 1e-2) of HIP frames
    of a static view against committed FP64 truths of the PathTracer estimator (written by
    tests/golden/make_truth.py): the path tracer at maxPathLen 8 against the truth at 8, and
    Conditional ReSTIR at maxPathLen 8 against the truth at 7, the path space it covers (its
    suffix stops one vertex early).  The frames are kept as 16 batch means: the rel-MSE of their
    plain mean (finite pixels; non-finite ones counted) and of their per-pixel median, each beside
    the rel-MSE of an all-black image and with the exposure normalised to a unit-mean truth
    (crisoptix_amd/metrics.py), since the reference's ReSTIR estimators are heavy-tailed."""
    from crisoptix_amd import abi, metrics, scenes
    out = {"frames": frames}
    for mode, name in (("conditional_restir", "objgrid16_96x54_L7"), ("path_tracer", "objgrid16_96x54")):
        path = os.path.join(ROOT, "tests", "golden", f"truth_{name}.npz")
        with np.load(path, allow_pickle=False) as z:
            truth, tvar, tcam, spp = z["mean"], z["var"], z["camera"], int(z["spp"])
        h, w = truth.shape[:2]
        cam = scenes.orbit_camera(scene, w, h)
        cd = abi.CameraData.from_dict(cam.data())
        if not np.array_equal(np.concatenate([np.asarray(getattr(cd, k), np.float32) for k in ("pos", "U", "V", "W")]),
                              tcam):
            raise RuntimeError(f"truth fixture {name} was rendered from another view")
        p = abi.workload_params(mode, w, h, gscene.size)
        ren = runtime.Renderer(w, h, ctx, cam, gscene, params=p, seed=3)
        ren.set_exec_mode(exec_mode)
        rep = metrics.batch_report(metrics.frame_batches(ren, cd, frames, 16), truth, tvar)
        rep.update({"truth": name, "truth_spp": spp, "view": f"{w}x{h} static, orbit camera frame 0"})
        out[mode] = {k: (round(v, 7) if isinstance(v, float) else v) for k, v in rep.items()}
    return out


def cpu_baseline(scene, params_for, W, H, scale, frames=6, warmup=2):
    """The oracle (host C++17 restatement of the shader logic, OpenMP over pixels) on the same
    workload at (W/scale) x (H/scale): `warmup` frames discarded, then the median frame time of
    `frames` frames; Mrays/s = rays of the measured frames / their summed time."""
    from crisoptix_amd import abi, scenes
    from oracle import oracle as O

    w, h = max(1, W // scale), max(1, H // scale)
    os_ = O.OracleScene(scene)
    p = params_for(w, h, os_.size)
    cam = scenes.orbit_camera(scene, w, h)
    ren = O.OracleRenderer(os_, w, h, p, seed=1)
    rays = 0
    times = []
    for f in range(warmup + frames):
        cam.rotate_around_origin([0, 1, 0], 2 * math.pi / 64)
        t0 = time.perf_counter()
        ren.frame(abi.CameraData.from_dict(cam.data()), True)
        dt = time.perf_counter() - t0
        if f >= warmup:
            times.append(dt)
            rays += int(ren.rays().sum())
    med = float(np.median(times))
    return {
        "value": rays / sum(times) / 1e6,
        "unit": "Mrays/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": f"{frames} frames after {warmup} warm-up frames of the same workload at {w}x{h}, oracle host "
                  f"C++ with OpenMP; median {med * 1000:.0f} ms/frame, {rays / frames / (w * h):.2f} rays/px",
        "ms_per_frame_at_sample": med * 1000,
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (requested --gpus {a.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    from crisoptix_amd import abi, runtime, scenes

    W, H = a.width, a.height
    t0 = time.perf_counter()
    scene = scenes.object_grid(a.grid, WORKLOAD_SEED)
    ctx = runtime.Context(0)
    gscene = runtime.Scene(ctx, scene, width=a.bvh_width, builder=a.bvh_builder, max_leaf=a.bvh_max_leaf)
    log(f"scene: {scene.num_triangles} triangles, BVH ({a.bvh_builder}) {gscene.info.numNodes} binary nodes, depth "
        f"{gscene.info.maxDepth}, width {gscene.info.bvhWidth} ({gscene.info.numWideNodes} wide nodes, stack bound "
        f"{gscene.info.stackBound}), BVH build {gscene.info.buildMs:.1f} ms, scene ready in "
        f"{time.perf_counter() - t0:.1f}s")

    def params_for(w, h, size):
        return abi.workload_params("conditional_restir", w, h, size)

    params = params_for(W, H, gscene.size)
    params.trackPrefixStats = 1   # the neighbour statistics of PrefixAccelStruct::Stats, every frame
    cam = scenes.orbit_camera(scene, W, H)
    ren = runtime.Renderer(W, H, ctx, cam, gscene, params=params, seed=1)
    ren.enable_timing(True)
    ren.set_exec_mode(a.exec)
    ren.set_overlap(not a.no_overlap)
    if a.wave_rounds is not None:
        ren.set_wave_rounds(a.wave_rounds)
    if a.queue_grid:
        ren.set_queue_grid(a.queue_grid)

    def frame():
        cam.rotate_around_origin([0, 1, 0], 2 * math.pi / 64)
        ren.launch_frame(abi.CameraData.from_dict(cam.data()), True)

    for _ in range(a.warmup):
        frame()
    runtime.synchronize()
    log(f"warmup done ({a.warmup} frames)")
    start, end = runtime.Event(), runtime.Event()
    pass_ms = np.zeros(abi.CX_PASS_COUNT)
    pass_rays = np.zeros((abi.CX_PASS_COUNT, 2))
    total_rays = 0
    trace_ms = 0.0
    trace_launches = 0
    tail_rays = np.zeros(2)
    entries = 0
    neigh_total = 0
    neigh_max = 0
    overflow = 0
    frame_ms = []
    for i in range(a.steps):
        start.record()
        frame()
        end.record()
        end.synchronize()
        frame_ms.append(start.elapsed_ms(end))
        st = ren.stats()
        pass_ms += np.asarray(st.passMs)
        pass_rays[:, 0] += np.asarray(st.passClosestRays)
        pass_rays[:, 1] += np.asarray(st.passShadowRays)
        total_rays += st.closestRays + st.shadowRays
        entries += st.numEntries
        neigh_total += st.totalNeighCount
        neigh_max = max(neigh_max, st.maxNeighCount)
        overflow |= st.queueOverflow
        trace_ms += st.traceMs
        trace_launches += st.traceLaunches
        tail_rays += (st.tailClosestRays, st.tailShadowRays)
    gpu_ms = float(np.sum(frame_ms))
    ms_per_frame = float(np.median(frame_ms))
    mrays = total_rays / (gpu_ms / 1000.0) / 1e6
    P = W * H
    # reuse density of the last timed frame: valid suffixes (front buffer) and the neighbour
    # count of every pixel (0 for pixels that did not gather)
    sv = ren.state()
    suffix_flags = runtime.download(sv.suffixRes[sv.frontBufferIdx], P * 16, np.uint32)[2::4]
    valid_suffix = float(np.mean((suffix_flags >> 16) & 1))
    ncount = runtime.download(sv.neighCount, P * 4, np.uint32)
    k = int(params.gatherM) - 1
    hist = np.bincount(ncount, minlength=k + 1)
    gathering = int((ncount > 0).sum())
    neigh_per_frame = neigh_total / a.steps
    fg = abi.PASS_NAMES.index("final_gather")
    fg_ms = pass_ms[fg] / a.steps
    fg_bytes = FINAL_GATHER_STATE_BYTES_PER_PX * P + NEIGHBOR_BYTES * neigh_per_frame + (
        pass_rays[fg, 0] * CLOSEST_RAY_BYTES + pass_rays[fg, 1] * SHADOW_RAY_BYTES) / a.steps
    fg_gbs = fg_bytes / (fg_ms / 1000.0) / 1e9
    ray_bytes = (pass_rays[:, 0].sum() * CLOSEST_RAY_BYTES + pass_rays[:, 1].sum() * SHADOW_RAY_BYTES) / a.steps
    if trace_launches:
        # Dominant kernel of the wavefront mode: the traversal launches, algorithmic bytes = the
        # rays they answer (all rays but those the tail kernels trace in place) x (32 B ray +
        # 16 B hit, or 32 B ray + 4 B flag).
        roof_kernel = "k_q_trace"
        launches_per_frame = trace_launches / a.steps
        avg_launch_ms = trace_ms / trace_launches
        launch_ray_bytes = ray_bytes - (tail_rays[0] * CLOSEST_RAY_BYTES + tail_rays[1] * SHADOW_RAY_BYTES) / a.steps
        bytes_per_launch = launch_ray_bytes / launches_per_frame
    else:
        roof_kernel = "k_final_gather"
        avg_launch_ms = fg_ms
        bytes_per_launch = fg_bytes
    achieved = bytes_per_launch / (avg_launch_ms / 1000.0) / 1e9
    frame_bytes = FRAME_STATE_BYTES_PER_PX * P + NEIGHBOR_BYTES * neigh_per_frame + ray_bytes
    out = runtime.DeviceArray.numpy(ren.output)
    log("per-pass ms: " + ", ".join(f"{n} {m / a.steps:.3f}" for n, m in zip(abi.PASS_NAMES, pass_ms)))
    log("per-pass rays (closest, shadow): " + ", ".join(
        f"{n} {int(c / a.steps)}/{int(s / a.steps)}" for n, (c, s) in zip(abi.PASS_NAMES, pass_rays)))
    log(f"frame ms median {ms_per_frame:.3f} (min {min(frame_ms):.3f}, max {max(frame_ms):.3f}); rays/frame "
        f"{total_rays / a.steps:.0f} ({total_rays / a.steps / P:.2f}/px), valid suffixes {valid_suffix:.3f}, "
        f"entries/frame {entries / a.steps:.0f}, neighbours/frame {neigh_per_frame:.0f} (histogram {hist.tolist()}), "
        f"mean radiance {float(out.mean()):.5f}, finite {bool(np.isfinite(out).all())}, queue overflow {overflow}")

    relmse = None
    if a.relmse_frames > 0 and a.grid == 16:
        log(f"rel-MSE vs FP64 truth ({a.relmse_frames} accumulated frames) ...")
        t1 = time.perf_counter()
        relmse = rel_mse_vs_truth(runtime, ctx, gscene, scene, a.relmse_frames, a.exec)
        log(f"rel-MSE: {relmse} ({time.perf_counter() - t1:.1f}s)")

    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(scene, params_for, W, H, a.cpu_sample_scale, a.cpu_frames)
        log(f"cpu baseline: {cpu['value']:.2f} Mrays/s on {cpu['cores']} threads ({cpu['sample']})")

    line = {
        "metric": METRIC,
        "value": round(mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_frame, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene, orbiting camera)",
        "config": {
            "workload": f"conditional_restir_{W}x{H}_objgrid{a.grid}_tent",
            "exec_mode": a.exec,
            "final_phase_overlap": not a.no_overlap,
            "queue_grid": a.queue_grid or 8192,
            "wave_rounds": a.wave_rounds if a.wave_rounds is not None else "3 for suffix walks, 2 for the others",
            "resolution": [W, H],
            "triangles": scene.num_triangles,
            "max_path_len": int(params.maxPathLen),
            "gather_m": int(params.gatherM),
            "ms_per_frame_median": round(ms_per_frame, 4),
            "ms_per_frame_mean": round(gpu_ms / a.steps, 4),
            "rays_per_frame": int(total_rays / a.steps),
            "rays_per_pixel": round(total_rays / a.steps / P, 3),
            "valid_suffix_fraction": round(valid_suffix, 4),
            "entries_per_frame": int(entries / a.steps),
            "neighbors_per_frame": int(neigh_per_frame),
            "neighbors_per_gathering_pixel": round(float(ncount[ncount > 0].mean()) if gathering else 0.0, 3),
            "neighbors_max": int(neigh_max),
            "neighbor_histogram_last_frame": hist.tolist(),
            "queue_overflow": int(overflow),
            "pass_ms": {n: round(m / a.steps, 4) for n, m in zip(abi.PASS_NAMES, pass_ms)},
            "frame_algorithmic_GBps": round(frame_bytes / (ms_per_frame / 1000.0) / 1e9, 2),
            "bvh_width": int(gscene.info.bvhWidth),
            "bvh_builder": a.bvh_builder,
            "bvh_max_leaf": a.bvh_max_leaf or 2,
            "bvh_build_ms": round(float(gscene.info.buildMs), 2),
            "trace_ms_per_frame": round(trace_ms / a.steps, 4),
            "tail_rays_per_frame": int(tail_rays.sum() / a.steps),
            "final_gather_pass": {"ms": round(fg_ms, 4), "algorithmic_bytes": int(fg_bytes),
                                  "algorithmic_GBps": round(fg_gbs, 2)},
            "rel_mse_vs_fp64_truth": relmse,
        },
        "roofline": {
            "kernel": roof_kernel,
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": traffic_from_profiles(roof_kernel, int(gscene.info.bvhWidth)),
            "avg_launch_ms": round(avg_launch_ms, 5),
            "algorithmic_bytes_per_launch": int(bytes_per_launch),
        },
        "cpu_baseline": None if cpu is None else {k: (round(v, 4) if isinstance(v, float) else v)
                                                  for k, v in cpu.items() if k != "ms_per_frame_at_sample"},
    }
    print(json.dumps(line), flush=True)
    if overflow:
        log("a wavefront queue overflowed (cx_frame_stats.queueOverflow): the frames are not valid")
        sys.exit(3)


if __name__ == "__main__":
    main()
