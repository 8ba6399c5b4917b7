"""torch.autograd.Function over the HIP frame loop and its path-replay adjoint.

    dr = DiffRenderer(ctx, scene, width, height, camera, mode="restir_pt")
    img = dr(albedo, roughness, emission)      # [H, W, 3] float32 tensor
    loss = ((img - target) ** 2).sum()
    loss.backward()                            # albedo.grad [M, 3], roughness.grad [M], emission.grad [E, 3]

Forward uploads the parameters to both material tables of the scene (cx_scene_set_materials,
cx_scene_set_emitter_colors), renders one frame (cx_renderer_frame) and returns its estimate.
Backward runs cx_renderer_adjoint on the renderer's state of that frame with dL/dI from
autograd, so each backward must come before the next forward of the same renderer.  The
gradient is the path-replay estimate W * df(x)/dtheta of crisoptix_amd/csrc/cx_adjoint.hip (the
reservoir-selected path x and its contribution weight W held fixed).

Here tensors cross the boundary as host arrays (the parameters are a few kilobytes and an image
is a few megabytes), the simplest form for a module in a larger autograd graph.  A loop whose
parameters stay on the device, and that records each step as one graph, is
crisoptix_amd.inverse.DeviceStep: it hands device pointers and torch's stream to the library.
"""
from __future__ import annotations

import numpy as np
import torch

from . import abi, runtime


class DiffRenderer:
    def __init__(self, ctx: runtime.Context, scene, width: int, height: int, camera, mode: str = "restir_pt",
                 params: abi.LaunchParams | None = None, seed: int = 0, gscene: runtime.Scene | None = None):
        self.ctx = ctx
        self.host = scene
        self.gscene = gscene if gscene is not None else runtime.Scene(ctx, scene)
        self.W, self.H = width, height
        if params is None:
            params = abi.workload_params("restir_pt_cornell" if mode == "restir_pt" else "path_tracer", width, height,
                                         self.gscene.size)
        self.params = params
        self.camera = camera
        self.cam = abi.CameraData.from_dict(camera.data())
        self.ren = runtime.Renderer(width, height, ctx, camera, self.gscene, params=params, seed=seed)
        self.base = np.asarray(scene.materials, np.float32).copy()
        M, E = self.base.shape[0], scene.emitter_color.shape[0]
        self.M, self.E = M, E
        self._g = [runtime.DeviceArray((M, 3), np.float64), runtime.DeviceArray((M,), np.float64),
                   runtime.DeviceArray((E, 3), np.float64)]
        self._dl = runtime.DeviceArray((height, width, 3), np.float32)
        self._status = runtime.DeviceArray((3,), np.uint32)
        self.frames = 0
        self.last_status = None

    def set_parameters(self, albedo: np.ndarray, roughness: np.ndarray, emission: np.ndarray, tables: int = 3):
        m = self.base.copy()
        m[:, 0:3] = np.asarray(albedo, np.float32).reshape(self.M, 3)
        m[:, 9] = np.asarray(roughness, np.float32).reshape(self.M)
        self.gscene.set_materials(m, tables)
        self.gscene.set_emitter_colors(np.asarray(emission, np.float32).reshape(self.E, 3), tables)

    def render(self, reset_history: bool = False) -> np.ndarray:
        if reset_history:
            self.ren.reset()
        img = self.ren.launch_frame(self.cam, self.frames == 0).numpy()
        self.frames += 1
        return img

    def adjoint(self, dLdI: np.ndarray):
        """(g_albedo [M,3], g_rough [M], g_emit [E,3]) float64 for dL/dI of the last frame."""
        self._dl.upload(np.ascontiguousarray(dLdI, np.float32))
        for a in self._g:
            a.zero()
        self._status.zero()
        self.ren.adjoint(self._dl, *self._g, status=self._status)
        self.last_status = self._status.numpy()
        return tuple(a.numpy() for a in self._g)

    def __call__(self, albedo: torch.Tensor, roughness: torch.Tensor, emission: torch.Tensor) -> torch.Tensor:
        return _RenderFn.apply(self, albedo, roughness, emission)


class _RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dr: DiffRenderer, albedo, roughness, emission):
        dr.set_parameters(albedo.detach().cpu().numpy(), roughness.detach().cpu().numpy(),
                          emission.detach().cpu().numpy())
        img = dr.render()
        ctx.dr = dr
        ctx.devices = (albedo.device, roughness.device, emission.device)
        return torch.from_numpy(np.ascontiguousarray(img)).to(albedo.device)

    @staticmethod
    def backward(ctx, grad_img):
        dr: DiffRenderer = ctx.dr
        gA, gR, gE = dr.adjoint(grad_img.detach().cpu().float().numpy())
        da, dr_, de = ctx.devices
        return (None, torch.from_numpy(gA.astype(np.float32)).to(da), torch.from_numpy(gR.astype(np.float32)).to(dr_),
                torch.from_numpy(gE.astype(np.float32)).to(de))
