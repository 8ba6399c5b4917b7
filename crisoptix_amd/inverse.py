"""Differentiable ReSTIR PT workloads on one GPU: a gradient check and an inverse-rendering loop.

    python -m crisoptix_amd.inverse gradcheck [--width 512 --height 512]
    python -m crisoptix_amd.inverse optimize  [--steps 200 --views 8 --width 640 --height 480]
                                              [--exec graph|device|host]

gradcheck: the 100 k-triangle object grid (64 object materials + ground, 16 emitters), ReSTIR
PT (temporal and spatial reuse, max path length 4) at 512 x 512.  The loss is the L2 image
loss against a target rendered (64 accumulated frames) with seeded perturbed albedo (+-0.2) and
roughness (+-0.3) (per-pixel median of 8 batch means of frames from a reset history; pixels
above 10 x the target's 99.9th percentile are left out of the loss).  The gradient of one frame's loss with respect to 65 x 3 albedo, 65
roughness and 16 x 3 emitter radiances comes from cx_renderer_adjoint; it is checked against
central finite differences (h = 1e-3, and their Richardson extrapolation with h = 5e-4) of the
replayed estimate along seeded random directions and for the largest gradients.  Prints one
JSON line with timings and agreement.

optimize: 8 views at 640 x 480 (orbit camera at 8 angles) of the object grid with 16 x 16
objects on a finely tessellated ground (optimize_scene: 407 k triangles, 256 object materials +
ground).  Ground-truth parameters are the scene's; the loop starts from albedo and roughness
U[0.3, 0.7] and runs Adam (lr 1e-2, torch.optim) on the summed L2 loss of the 8 views; each step
renders per view one frame for the residual and an independent frame whose adjoint gives the
gradient (decorrelated, so the variance of the estimator is not minimised along with the
error).  Targets are 64 frames per
view at the true parameters (median of 8 batch means); the L2 loss leaves out pixels above
10 x the target's 99.9th percentile.  Each step renders from a reset history (the
reservoirs of the previous step were built with other parameters).  Reports parameter
relative error (all materials, and those covering at least --min-pixels primary hits over the
views) and ms per step.  --exec graph (default) records the device-resident step (DeviceStep)
once after 3 eager steps and replays it; --exec device enqueues the same step eagerly; --exec
host runs the views one after another with host round trips (materials, images and gradients
as host arrays).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import abi, runtime, scenes


def optimize_scene():
    """The inverse-rendering scene: 16 x 16 objects (icospheres of subdivision 3 and boxes, one
    material each) on a 290 x 290-quad ground, 407,120 triangles and 256 object materials +
    ground, lit by 16 light panels."""
    return scenes.object_grid(16, seed=3, ico_subdiv=3, ground_quads=290, trace_emitters=False, lights="panels")


def _params(size, w, h):
    p = abi.workload_params("restir_pt_cornell", w, h, size)
    p.maxPathLen = 4
    return p


def _accumulate(ren, cd, frames, batches=8):
    """Target image: per-pixel median of `batches` batch means of frames rendered, like each
    step of the loops below, from a reset history.  The reference's ReSTIR PT estimator is
    heavy-tailed (unclamped reconnection Jacobian evaluated at the primary hit,
    include/graph/restir/prefix_reuse.h:88-92): a plain mean of a few dozen frames can hold a
    1e30 firefly, and temporal reuse would carry one through many frames, so every frame
    starts from a reset reservoir history and the batch means are combined by their median."""
    per = max(1, frames // batches)
    means = []
    for b in range(batches):
        acc = None
        for f in range(per):
            ren.reset()
            img = ren.launch_frame(cd, True).numpy().astype(np.float64)
            acc = img if acc is None else acc + img
        means.append(acc / per)
    return np.median(np.stack(means), axis=0).astype(np.float32)


def _clip_level(target) -> float:
    """Pixel values above this are left out of the loss (their gradient is zero): 10 x the
    target's 99.9th percentile, far above any converged pixel and far below the estimator's
    fireflies."""
    return float(10.0 * np.percentile(target, 99.9))


def gradcheck(a) -> dict:
    ctx = runtime.Context(0)
    sc = scenes.inverse_scene()
    gs = runtime.Scene(ctx, sc)
    W, H = a.width, a.height
    p = _params(gs.size, W, H)
    cam = scenes.orbit_camera(sc, W, H)
    cd = abi.CameraData.from_dict(cam.data())
    ren = runtime.Renderer(W, H, ctx, cam, gs, params=p, seed=7)
    mats = sc.materials.astype(np.float32)
    emit = sc.emitter_color.astype(np.float32)
    tmats = scenes.perturbed_materials(mats, seed=8)
    gs.set_materials(tmats, 3)
    target = _accumulate(ren, cd, a.target_frames)
    gs.set_materials(mats, 3)
    ren.reset()
    ren.seed(7)
    for f in range(3):
        img = ren.launch_frame(cd, f == 0).numpy()
    clip = _clip_level(target)
    kept = img <= clip
    dLdI = np.where(kept, 2.0 * (img - target), 0.0).astype(np.float32)
    loss = float(np.sum(np.where(kept, (img.astype(np.float64) - target) ** 2, (clip - target) ** 2)))
    M, E = mats.shape[0], emit.shape[0]
    dl = runtime.DeviceArray.from_numpy(dLdI)
    gA, gR, gE = (runtime.DeviceArray(s, np.float64) for s in ((M, 3), (M,), (E, 3)))
    rep = runtime.DeviceArray((H, W, 3), np.float32)
    st = runtime.DeviceArray((3,), np.uint32)

    def adjoint(grad=True):
        for x in (gA, gR, gE, st):
            x.zero()
        ren.adjoint(dl if grad else None, gA if grad else None, gR if grad else None, gE if grad else None, rep, st)

    # timings: a forward frame and an adjoint pass, HIP events on the default stream
    t0, t1 = runtime.Event(), runtime.Event()
    fwd, adj = [], []
    adjoint()
    g_alb, g_rough, g_emit, status = gA.numpy(), gR.numpy(), gE.numpy(), st.numpy()
    replay0 = rep.numpy()
    for _ in range(a.reps):
        t0.record()
        ren.launch_frame(cd, False)
        t1.record()
        t1.synchronize()
        fwd.append(t0.elapsed_ms(t1))
        t0.record()
        adjoint()
        t1.record()
        t1.synchronize()
        adj.append(t0.elapsed_ms(t1))
    # restore the checked frame's state: render the same three frames again
    ren.reset()
    ren.seed(7)
    for f in range(3):
        img2 = ren.launch_frame(cd, f == 0).numpy()
    if not np.array_equal(img2.view(np.uint32), img.view(np.uint32)):
        raise RuntimeError("re-rendered frame differs")

    def J(m, e):
        gs.set_materials(m, 2)
        gs.set_emitter_colors(e, 2)
        adjoint(grad=False)
        return float(np.sum(rep.numpy().astype(np.float64) * dLdI))

    def central(dm, de, h):
        a_m, b_m = (mats + np.float32(h) * dm).astype(np.float32), (mats - np.float32(h) * dm).astype(np.float32)
        a_e, b_e = (emit + np.float32(h) * de).astype(np.float32), (emit - np.float32(h) * de).astype(np.float32)
        sm, se = (a_m.astype(np.float64) - b_m) / 2, (a_e.astype(np.float64) - b_e) / 2
        lin = np.sum(g_alb * sm[:, 0:3]) + np.sum(g_rough * sm[:, 9]) + np.sum(g_emit * se)
        mag = np.sum(np.abs(g_alb * sm[:, 0:3])) + np.sum(np.abs(g_rough * sm[:, 9])) + np.sum(np.abs(g_emit * se))
        return ((J(a_m, a_e) - J(b_m, b_e)) / 2 - lin) / mag   # error relative to sum |g_i d_i|

    rng = np.random.default_rng(9)
    dirs = []
    for _ in range(a.directions):
        dm = np.zeros_like(mats)
        dm[:, 0:3] = rng.standard_normal((M, 3))
        dm[:, 9] = rng.standard_normal(M)
        de = rng.standard_normal((E, 3)).astype(np.float32)
        r1, r2 = central(dm, de, 1e-3), central(dm, de, 5e-4)
        dirs.append({"central_h1e-3": abs(r1), "richardson": abs((4 * r2 - r1) / 3)})
    flat = np.concatenate([g_alb.ravel(), g_rough, g_emit.ravel()])
    order = np.argsort(-np.abs(flat))[:a.top]
    per = []
    for idx in order:
        dm, de = np.zeros_like(mats), np.zeros_like(emit)
        if idx < 3 * M:
            dm[idx // 3, idx % 3] = 1.0
        elif idx < 4 * M:
            dm[idx - 3 * M, 9] = 1.0
        else:
            de.reshape(-1)[idx - 4 * M] = 1.0
        r1, r2 = central(dm, de, 1e-3), central(dm, de, 5e-4)
        per.append({"param": int(idx), "central_h1e-3": abs(r1), "richardson": abs((4 * r2 - r1) / 3)})
    gs.set_materials(mats, 3)
    gs.set_emitter_colors(emit, 3)
    return {
        "workload": f"differentiable_restir_pt_{W}x{H}_inverse_scene",
        "triangles": sc.num_triangles, "materials": M, "emitters": E,
        "loss": loss, "loss_clip_level": clip, "pixels_clipped": int((~kept).any(-1).sum()),
        "pixels_replayed": int(status[0]), "replay_differs": int(status[1]), "skipped": int(status[2]),
        "replay_equals_frame": bool(np.array_equal(replay0.view(np.uint32), img.view(np.uint32))),
        "forward_ms": float(np.median(fwd)), "adjoint_ms": float(np.median(adj)),
        "fd_directional_max_rel_err": {k: max(d[k] for d in dirs) for k in ("central_h1e-3", "richardson")},
        "fd_top_gradients_max_rel_err": {k: max(d[k] for d in per) for k in ("central_h1e-3", "richardson")},
        "fd_top_gradients": len(per), "fd_directions": len(dirs),
    }


class DeviceStep:
    """One optimisation step of the inverse-rendering loop with every operand in device memory:
    the material rows are assembled from the parameter tensors and written with
    cx_scene_set_materials_device; each view renders a residual frame into a tensor, the masked
    residual becomes dL/dI on the GPU, an independent frame is rendered and its adjoint
    accumulates into one pair of float64 gradient tensors shared by all views; Adam
    (capturable) updates the parameters and they are clamped.  Each view runs on a stream of its
    own, forked from and joined back into `stream` (small views leave most of the GPU idle on
    their own; their adjoints add into the shared gradients with atomics).  Frames draw their
    random numbers on the device (cx_renderer_set_device_random), so the whole step can be
    recorded once with torch.cuda.graph and replayed; the fork and join are recorded with it.
    Nothing in a step synchronises with the host.

    views: list of (runtime.Renderer, abi.CameraData); targets: [H, W, 3] float32 tensors on the
    device; clips: per-view residual mask levels; base: [M, 10] material rows (specular and
    emission are kept, albedo and roughness come from the parameters)."""

    def __init__(self, gscene, views, targets, clips, base, albedo, rough, lr, seed=0):
        import torch
        self.torch = torch
        dev = torch.device("cuda")
        self.stream = torch.cuda.Stream()
        self.sp = self.stream.cuda_stream
        self.vstreams = [torch.cuda.Stream() for _ in views]
        self.gscene, self.views = gscene, views
        self.targets = [t.to(dev) for t in targets]
        self.clips = [float(c) for c in clips]
        self.base = torch.from_numpy(np.ascontiguousarray(base, np.float32)).to(dev)
        self.rows = self.base.clone()
        self.M = self.base.shape[0]
        self.albedo = torch.tensor(np.asarray(albedo, np.float32), device=dev)
        self.rough = torch.tensor(np.asarray(rough, np.float32), device=dev)
        self.albedo.grad = torch.zeros_like(self.albedo)
        self.rough.grad = torch.zeros_like(self.rough)
        self.opt = torch.optim.Adam([self.albedo, self.rough], lr=lr, capturable=True)
        H, W = self.targets[0].shape[0], self.targets[0].shape[1]
        self.npx = float(W * H * 3)
        self.gA = torch.zeros((self.M, 3), dtype=torch.float64, device=dev)
        self.gR = torch.zeros(self.M, dtype=torch.float64, device=dev)
        self.imgs = [torch.zeros((H, W, 3), dtype=torch.float32, device=dev) for _ in views]
        self.imgs2 = [torch.zeros_like(i) for i in self.imgs]
        self.dls = [torch.zeros_like(i) for i in self.imgs]
        self.zero = torch.zeros((), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()   # earlier work on the renderers ends before their generators are seeded
        for v, (ren, _) in enumerate(views):
            ren.set_device_random(True, seed + v)
        self.graph = None

    def step(self):
        torch = self.torch
        with torch.cuda.stream(self.stream), torch.no_grad():
            self.rows.copy_(self.base)
            self.rows[:, 0:3] = self.albedo
            self.rows[:, 9] = self.rough
            self.gscene.set_materials_device(self.rows.data_ptr(), self.M, 3, self.sp)
            self.gA.zero_()
            self.gR.zero_()
        for v, (ren, cam) in enumerate(self.views):
            vs = self.vstreams[v]
            vs.wait_stream(self.stream)
            sp = vs.cuda_stream
            with torch.cuda.stream(vs), torch.no_grad():
                # decorrelated estimate: residual from one frame, adjoint of an independent one
                ren.reset(stream=sp)
                ren.launch_frame(cam, False, output=runtime.DevicePtr(self.imgs[v].data_ptr()), stream=sp)
                img, tgt = self.imgs[v], self.targets[v]
                torch.where(img <= self.clips[v], (img - tgt) * (2.0 / self.npx), self.zero, out=self.dls[v])
                ren.reset(stream=sp)
                ren.launch_frame(cam, False, output=runtime.DevicePtr(self.imgs2[v].data_ptr()), stream=sp)
                ren.adjoint(self.dls[v].data_ptr(), self.gA.data_ptr(), self.gR.data_ptr(), None, stream=sp)
        for vs in self.vstreams:
            self.stream.wait_stream(vs)
        with torch.cuda.stream(self.stream), torch.no_grad():
            self.albedo.grad.copy_(self.gA)
            self.rough.grad.copy_(self.gR)
        with torch.cuda.stream(self.stream):
            self.opt.step()
            with torch.no_grad():
                self.albedo.clamp_(0.01, 0.99)
                self.rough.clamp_(0.02, 1.0)

    def capture(self):
        """Records step() into a graph (after at least one eager step has allocated everything)."""
        torch = self.torch
        self.stream.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self.step()

    def replay(self):
        self.graph.replay()

    def sync(self):
        """Waits for the enqueued steps; tensors of the step may then be read on any stream."""
        self.stream.synchronize()

    def loss(self) -> float:
        """Masked L2 loss of the last step's residual frames (synchronises)."""
        torch = self.torch
        with torch.cuda.stream(self.stream):
            tot = sum(float(torch.where(i <= c, (i - t) ** 2, self.zero).double().sum()) / self.npx
                      for i, t, c in zip(self.imgs, self.targets, self.clips))
        return tot


def optimize(a) -> dict:
    import torch

    from .autograd import DiffRenderer

    ctx = runtime.Context(0)
    sc = optimize_scene()
    gs = runtime.Scene(ctx, sc)
    W, H = a.width, a.height
    truth = sc.materials.astype(np.float32)
    M = truth.shape[0]
    views = []
    for v in range(a.views):
        cam = scenes.orbit_camera(sc, W, H)
        cam.rotate_around_origin([0, 1, 0], 2 * math.pi * v / a.views)
        views.append(DiffRenderer(ctx, sc, W, H, cam, params=_params(gs.size, W, H), seed=100 + v, gscene=gs))
    gs.set_materials(truth, 3)
    t0 = time.perf_counter()
    targets = [torch.from_numpy(_accumulate(dv.ren, dv.cam, a.target_frames)) for dv in views]
    clips = [_clip_level(t.numpy()) for t in targets]
    t_target = time.perf_counter() - t0
    rng = np.random.default_rng(12)
    albedo = torch.tensor(rng.uniform(0.3, 0.7, (M, 3)).astype(np.float32), requires_grad=True)
    rough = torch.tensor(rng.uniform(0.3, 0.7, M).astype(np.float32), requires_grad=True)
    emit = sc.emitter_color.astype(np.float32)
    opt = torch.optim.Adam([albedo, rough], lr=a.lr)
    # Materials seen by the views: pixels whose primary hit lies on each material.
    seen = np.zeros(M, np.int64)
    for dv in views:
        dv.ren.reset()
        dv.render()
        pre = dv.ren.prefix(dv.ren.state().frontBufferIdx)
        prim = pre["primary_prim"]
        ok = prim != 0xFFFFFFFF
        seen += np.bincount(sc.tri_material[prim[ok]], minlength=M)
    observed = seen >= a.min_pixels

    def rel_err(mask=None):
        m = np.ones(M, bool) if mask is None else mask
        al = albedo.detach().numpy()[m]
        ro = rough.detach().numpy()[m]
        ea = np.linalg.norm(al - truth[m, 0:3]) / np.linalg.norm(truth[m, 0:3])
        er = np.linalg.norm(ro - truth[m, 9]) / np.linalg.norm(truth[m, 9])
        return float(ea), float(er)

    e0, e0o = rel_err(), rel_err(observed)
    npx = W * H * 3
    hist = []
    if a.exec != "host":
        return _optimize_device(a, sc, gs, views, targets, clips, truth, albedo, rough, observed, t_target, e0, e0o)
    runtime.synchronize()
    t0 = time.perf_counter()
    for step in range(a.steps):
        opt.zero_grad()
        total = 0.0
        gA = np.zeros((M, 3))
        gR = np.zeros(M)
        al, ro = albedo.detach().numpy(), rough.detach().numpy()
        for dv, tgt, clip in zip(views, targets, clips):
            dv.set_parameters(al, ro, emit)
            # Decorrelated estimate: the residual comes from one frame, the adjoint replays an
            # independent one, so the product is unbiased for the gradient of the squared
            # error of the expected image (one frame for both would add the gradient of the
            # estimator's variance).
            dv.ren.reset()
            img = dv.render()
            tg = tgt.numpy()
            kept = img <= clip
            total += float(np.sum(np.where(kept, (img - tg) ** 2, 0.0)) / npx)
            dv.ren.reset()
            dv.render()
            ga, gr, _ = dv.adjoint(np.where(kept, 2.0 * (img - tg) / npx, 0.0).astype(np.float32))
            gA += ga
            gR += gr
        albedo.grad = torch.from_numpy(gA.astype(np.float32))
        rough.grad = torch.from_numpy(gR.astype(np.float32))
        opt.step()
        with torch.no_grad():
            albedo.clamp_(0.01, 0.99)
            rough.clamp_(0.02, 1.0)
        if step % a.log_every == 0 or step == a.steps - 1:
            hist.append({"step": step, "loss": total, "rel_err": rel_err(), "rel_err_observed": rel_err(observed)})
            print(f"step {step}: loss {total:.6f}, rel err albedo {hist[-1]['rel_err'][0]:.4f} "
                  f"roughness {hist[-1]['rel_err'][1]:.4f}; observed materials: albedo "
                  f"{hist[-1]['rel_err_observed'][0]:.4f} roughness {hist[-1]['rel_err_observed'][1]:.4f}",
                  file=sys.stderr, flush=True)
    runtime.synchronize()
    ms = (time.perf_counter() - t0) * 1000 / a.steps
    e1, e1o = rel_err(), rel_err(observed)
    return {
        "workload": f"inverse_rendering_{a.views}x{W}x{H}_objgrid16_subdiv3",
        "triangles": sc.num_triangles, "materials": M, "views": a.views, "steps": a.steps, "lr": a.lr,
        "target_frames": a.target_frames, "target_seconds": t_target,
        "ms_per_step": ms, "frames_per_step": 2 * a.views,
        "observed_materials": int(observed.sum()), "observed_min_pixels": a.min_pixels,
        "rel_err_initial": {"albedo": e0[0], "roughness": e0[1]},
        "rel_err_final": {"albedo": e1[0], "roughness": e1[1]},
        "rel_err_observed_initial": {"albedo": e0o[0], "roughness": e0o[1]},
        "rel_err_observed_final": {"albedo": e1o[0], "roughness": e1o[1]},
        "history": hist,
    }


def _optimize_device(a, sc, gs, views, targets, clips, truth, albedo, rough, observed, t_target, e0, e0o) -> dict:
    """The loop of optimize() with DeviceStep: eagerly (--exec device) or as one captured graph
    per step replayed after WARMUP eager steps (--exec graph)."""
    import torch
    M = truth.shape[0]
    ds = DeviceStep(gs, [(dv.ren, dv.cam) for dv in views], targets, clips, truth, albedo.detach().numpy(),
                    rough.detach().numpy(), a.lr, seed=1000)

    def rel_err(mask=None):
        m = np.ones(M, bool) if mask is None else mask
        al = ds.albedo.detach().cpu().numpy()[m]
        ro = ds.rough.detach().cpu().numpy()[m]
        ea = np.linalg.norm(al - truth[m, 0:3]) / np.linalg.norm(truth[m, 0:3])
        er = np.linalg.norm(ro - truth[m, 9]) / np.linalg.norm(truth[m, 9])
        return float(ea), float(er)

    hist = []
    warm = min(WARMUP, a.steps) if a.exec == "graph" else 0
    for _ in range(warm):
        ds.step()
    if a.exec == "graph":
        ds.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for step in range(warm, a.steps):
        if a.exec == "graph":
            ds.replay()
        else:
            ds.step()
        if step % a.log_every == 0 or step == a.steps - 1:
            hist.append({"step": step, "loss": ds.loss(), "rel_err": rel_err(), "rel_err_observed": rel_err(observed)})
            print(f"step {step}: loss {hist[-1]['loss']:.6f}, rel err albedo {hist[-1]['rel_err'][0]:.4f} "
                  f"roughness {hist[-1]['rel_err'][1]:.4f}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000 / max(1, a.steps - warm)
    e1, e1o = rel_err(), rel_err(observed)
    return {
        "workload": f"inverse_rendering_{a.views}x{a.width}x{a.height}_objgrid16_subdiv3",
        "exec": a.exec, "triangles": sc.num_triangles, "materials": M, "views": a.views, "steps": a.steps,
        "lr": a.lr, "target_frames": a.target_frames, "target_seconds": t_target, "ms_per_step": ms,
        "frames_per_step": 2 * a.views, "observed_materials": int(observed.sum()),
        "observed_min_pixels": a.min_pixels,
        "rel_err_initial": {"albedo": e0[0], "roughness": e0[1]},
        "rel_err_final": {"albedo": e1[0], "roughness": e1[1]},
        "rel_err_observed_initial": {"albedo": e0o[0], "roughness": e0o[1]},
        "rel_err_observed_final": {"albedo": e1o[0], "roughness": e1o[1]},
        "history": hist,
    }


WARMUP = 3   # eager steps before a step graph is captured


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["gradcheck", "optimize"])
    ap.add_argument("--width", type=int)
    ap.add_argument("--height", type=int)
    ap.add_argument("--target-frames", type=int, default=64)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--directions", type=int, default=4)
    ap.add_argument("--top", type=int, default=16)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--views", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--min-pixels", type=int, default=200)
    ap.add_argument("--log-every", type=int, default=20, help="optimize: steps between history entries")
    ap.add_argument("--exec", choices=["host", "device", "graph"], default="graph",
                    help="optimize: host round trips per view (host), a device-resident step (device), or that "
                         "step captured once and replayed (graph)")
    a = ap.parse_args(argv)
    if a.mode == "gradcheck":
        a.width, a.height = a.width or 512, a.height or 512
        out = gradcheck(a)
    else:
        a.width, a.height = a.width or 640, a.height or 480
        out = optimize(a)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
