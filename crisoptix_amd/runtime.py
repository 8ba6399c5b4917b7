"""Host-side mirror of the reference's Renderer / Scene / Camera API over the C ABI.

``Renderer(width, height, ctx, camera, scene).launch_frame(output)`` stands in for
``Renderer(w, h, OptixDeviceContext, Camera&, const Scene&)`` and ``LaunchFrame(glm::vec3*)``
(include/graph/Renderer.h:11-21); ``Scene(ctx, scene_arrays)`` for
``Scene(ctx, instances, emitters)`` (include/model/Scene.h:12-24).  Device memory is handled
through the HIP runtime directly (ctypes), so no framework import is needed.

The HIP library must have been built (``python -m crisoptix_amd.build``); loading raises
:class:`LibraryNotBuilt` otherwise.  Errors from the C ABI raise :class:`CxError` with the
message of ``cx_last_error`` (the reference's debug-build ASSERT_CUDA/ASSERT_OPTIX behaviour,
include/util/custom_assert.h:26-66).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcrisoptix_amd.so")


class LibraryNotBuilt(RuntimeError):
    pass


class CxError(RuntimeError):
    pass


_hip = None
_lib = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise LibraryNotBuilt("HIP runtime libamdhip64.so not found")
        vp, sz, i32 = C.c_void_p, C.c_size_t, C.c_int
        _hip.hipMalloc.argtypes = [C.POINTER(vp), sz]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, sz, i32]
        _hip.hipMemset.argtypes = [vp, i32, sz]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(i32)]
        _hip.hipEventCreate.argtypes = [C.POINTER(vp)]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        _hip.hipEventDestroy.argtypes = [vp]
        _hip.hipStreamCreate.argtypes = [C.POINTER(vp)]
        _hip.hipStreamDestroy.argtypes = [vp]
        _hip.hipStreamSynchronize.argtypes = [vp]
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipGetErrorString.argtypes = [i32]
    return _hip


def _hcheck(rc, what):
    if rc != 0:
        raise CxError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def device_count() -> int:
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _hcheck(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DevicePtr:
    """A device pointer owned elsewhere (a torch tensor's data_ptr(), say), where the API takes
    a buffer object."""

    def __init__(self, ptr: int):
        self.ptr = int(ptr)


class DeviceArray:
    """A device allocation with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _hcheck(hip().hipMalloc(C.byref(p), max(self.nbytes, 16)), "hipMalloc")
        self.ptr = p.value

    @staticmethod
    def from_numpy(a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        if self.nbytes:
            _hcheck(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1), "hipMemcpy H2D")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _hcheck(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def zero(self):
        _hcheck(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download(ptr: int, nbytes: int, dtype) -> np.ndarray:
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
    if nbytes:
        _hcheck(hip().hipMemcpy(out.ctypes.data, ptr, nbytes, 2), "hipMemcpy D2H")
    return out


class Event:
    def __init__(self):
        e = C.c_void_p()
        _hcheck(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.e = e.value

    def record(self, stream=None):
        _hcheck(hip().hipEventRecord(self.e, stream), "hipEventRecord")

    def synchronize(self):
        _hcheck(hip().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _hcheck(hip().hipEventElapsedTime(C.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            hip().hipEventDestroy(self.e)
        except Exception:
            pass


def lib():
    """Loads libcrisoptix_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(f"{LIB_PATH} is missing: build it with `python -m crisoptix_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
        L.cx_abi_version.restype = i32
        if L.cx_abi_version() != abi.ABI_VERSION:
            raise LibraryNotBuilt(f"{LIB_PATH} has ABI {L.cx_abi_version()}, expected {abi.ABI_VERSION}: rebuild it")
        L.cx_bvh_host_trace.argtypes = [vp, vp, u32, i32, vp, vp, vp, vp, C.POINTER(u32)]
        L.cx_scene_set_materials.argtypes = [vp, vp, u32, i32]
        L.cx_scene_set_emitter_colors.argtypes = [vp, vp, u32, i32]
        L.cx_renderer_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.cx_tonemap.argtypes = [vp, vp, vp, u32, vp]
        L.cx_batch_frame.argtypes = [vp, vp, vp, vp, vp, u32, vp]
        L.cx_batch_sync.argtypes = [vp]
        L.cx_renderer_set_queue_grid.argtypes = [vp, u32]
        L.cx_renderer_set_graph_mode.argtypes = [vp, i32]
        L.cx_renderer_set_overlap.argtypes = [vp, i32]
        L.cx_renderer_set_wave_rounds.argtypes = [vp, u32]
        L.cx_renderer_set_device_random.argtypes = [vp, i32, u64]
        L.cx_scene_set_materials_device.argtypes = [vp, vp, u32, i32, vp]
        L.cx_renderer_graph_captures.argtypes = [vp]
        L.cx_ctx_create.argtypes = [i32, C.POINTER(vp)]
        L.cx_ctx_destroy.argtypes = [vp]
        L.cx_last_error.restype = C.c_char_p
        L.cx_last_error.argtypes = [vp]
        L.cx_scene_build.argtypes = [vp, C.POINTER(abi.SceneDesc), C.POINTER(vp)]
        L.cx_scene_destroy.argtypes = [vp]
        L.cx_scene_get_info.argtypes = [vp, C.POINTER(abi.SceneInfo)]
        L.cx_trace_closest.argtypes = [vp, vp, vp, u32, vp, vp, vp, vp, vp]
        L.cx_trace_occluded.argtypes = [vp, vp, vp, u32, vp, vp]
        L.cx_entries_create.argtypes = [vp, u32, C.POINTER(vp)]
        L.cx_entries_destroy.argtypes = [vp]
        L.cx_entries_build.argtypes = [vp, vp, vp, vp, vp, u32, vp, vp, vp]
        L.cx_entries_query.argtypes = [vp, vp, vp, u32, u32, vp, vp, vp, vp]
        L.cx_entries_search_static.argtypes = [vp, vp, vp, vp, u32, vp, vp, vp, u32, u32, vp, vp, vp, vp]
        L.cx_default_params.argtypes = [vp, u32, u32, C.POINTER(abi.LaunchParams)]
        L.cx_renderer_create.argtypes = [vp, vp, u32, u32, C.POINTER(vp)]
        L.cx_renderer_destroy.argtypes = [vp]
        L.cx_renderer_set_params.argtypes = [vp, C.POINTER(abi.LaunchParams)]
        L.cx_renderer_get_params.argtypes = [vp, C.POINTER(abi.LaunchParams)]
        L.cx_renderer_seed.argtypes = [vp, u64]
        L.cx_renderer_reset.argtypes = [vp, vp]
        L.cx_renderer_frame.argtypes = [vp, C.POINTER(abi.CameraData), i32, vp, vp]
        L.cx_renderer_begin_frame.argtypes = [vp, C.POINTER(abi.CameraData), i32, vp]
        L.cx_pass_launch.argtypes = [vp, i32, C.POINTER(abi.LaunchParams), vp]
        L.cx_renderer_enable_timing.argtypes = [vp, i32]
        L.cx_renderer_set_exec_mode.argtypes = [vp, i32]
        L.cx_renderer_get_exec_mode.argtypes = [vp]
        L.cx_renderer_get_stats.argtypes = [vp, C.POINTER(abi.FrameStats)]
        L.cx_renderer_state.argtypes = [vp, C.POINTER(abi.StateView)]
        _lib = L
    return _lib


def host_trace(scene, rays: np.ndarray, any_hit: bool = False, width: int = 4, max_leaf: int = 0):
    """cx_bvh_host_trace: the BVH of `scene` at `width` built and traversed on the CPU with the
    device traversal code.  Returns (prim, t, u, v, width_used), or (occluded, width_used)."""
    keep: list = []
    desc = abi.scene_desc(scene, keep)
    desc.bvhMaxLeafSize = max_leaf
    desc.bvhWidth = width
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    prim = np.empty(n, np.uint32)
    t, u, v = (np.empty(n, np.float32) for _ in range(3))
    used = C.c_uint32(0)
    rc = lib().cx_bvh_host_trace(C.byref(desc), rays.ctypes.data, n, 1 if any_hit else 0, prim.ctypes.data,
                                 t.ctypes.data, u.ctypes.data, v.ctypes.data, C.byref(used))
    if rc != 0:
        raise CxError(f"cx_bvh_host_trace failed ({rc})")
    if any_hit:
        return prim, used.value
    return prim, t, u, v, used.value


def tonemap(ctx: "Context", hdr: "DeviceArray", stream=None) -> np.ndarray:
    """cx_tonemap of a float image on the device; returns the sRGB bytes (same shape)."""
    ldr = DeviceArray(hdr.shape, np.uint8)
    n = int(np.prod(hdr.shape))
    ctx.check(lib().cx_tonemap(ctx.h, hdr.ptr, ldr.ptr, n, stream), "cx_tonemap")
    return ldr.numpy()


class Batch:
    """cx_batch_frame over several renderers (views of any scenes and resolutions)."""

    def __init__(self, ctx: "Context", renderers: list):
        self.ctx = ctx
        self.rs = list(renderers)
        n = len(self.rs)
        self._h = (C.c_void_p * n)(*[r.h for r in self.rs])
        self._outs = (C.c_void_p * n)(*[r.output.ptr for r in self.rs])
        self._cams = (abi.CameraData * n)()
        self._chg = (C.c_int * n)()

    def frame(self, cams: list, changed: list):
        for i, (c, ch) in enumerate(zip(cams, changed)):
            self._cams[i] = c
            self._chg[i] = 1 if ch else 0
        self.ctx.check(lib().cx_batch_frame(self.ctx.h, self._h, self._cams, self._chg, self._outs, len(self.rs),
                                            None), "cx_batch_frame")

    def sync(self):
        self.ctx.check(lib().cx_batch_sync(self.ctx.h), "cx_batch_sync")


class Context:
    """Replaces InitCuda/InitOptix (src/main.cpp:35-65)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib().cx_ctx_create(device, C.byref(h))
        if rc != 0:
            raise CxError(f"cx_ctx_create failed ({rc}): no usable HIP device {device}")
        self.h = h.value

    def check(self, rc: int, what: str):
        if rc != 0:
            raise CxError(f"{what} failed ({rc}): {lib().cx_last_error(self.h).decode()}")

    def __del__(self):
        if getattr(self, "h", None):
            lib().cx_ctx_destroy(self.h)
            self.h = None


class Scene:
    """Device scene: flattened triangles, BVH, materials and explicit emitters."""

    def __init__(self, ctx: Context, scene, max_leaf: int = 0, width: int = 0, builder: str = "host"):
        """builder: "host" (binned SAH on the CPU) or "device" (PLOC on the GPU); same hits."""
        self.ctx = ctx
        self._keep: list = []
        desc = abi.scene_desc(scene, self._keep)
        desc.bvhMaxLeafSize = max_leaf
        desc.bvhWidth = width
        desc.bvhBuilder = abi.BVH_BUILDERS[builder]
        h = C.c_void_p()
        ctx.check(lib().cx_scene_build(ctx.h, C.byref(desc), C.byref(h)), "cx_scene_build")
        self.h = h.value
        self.host = scene
        info = abi.SceneInfo()
        ctx.check(lib().cx_scene_get_info(self.h, C.byref(info)), "cx_scene_get_info")
        self.info = info

    @property
    def size(self) -> float:
        return self.info.sceneSize

    TABLE_SAMPLING, TABLE_EVAL, TABLE_BOTH = 1, 2, 3

    def set_materials(self, materials: np.ndarray, tables: int = 3):
        """materials: [M, 10] float32 rows of cx_material (diffuse rgb, specF0 rgb, emissive rgb,
        roughness).  tables: 1 sampling, 2 evaluation, 3 both."""
        m = np.ascontiguousarray(materials, np.float32).reshape(-1, 10)
        self.ctx.check(lib().cx_scene_set_materials(self.h, m.ctypes.data, m.shape[0], tables),
                       "cx_scene_set_materials")

    def set_materials_device(self, rows_ptr: int, n: int, tables: int = 3, stream=None):
        """Material rows (n x 10 float32, cx_material layout) from device memory, enqueued on
        `stream` (cx_scene_set_materials_device)."""
        self.ctx.check(lib().cx_scene_set_materials_device(self.h, rows_ptr, n, tables, stream),
                       "cx_scene_set_materials_device")

    def set_emitter_colors(self, colors: np.ndarray, tables: int = 3):
        c = np.ascontiguousarray(colors, np.float32).reshape(-1, 3)
        self.ctx.check(lib().cx_scene_set_emitter_colors(self.h, c.ctypes.data, c.shape[0], tables),
                       "cx_scene_set_emitter_colors")

    def trace_closest(self, rays: np.ndarray, stream=None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        d_rays = DeviceArray.from_numpy(rays)
        prim, t, u, v = (DeviceArray((n,), dt) for dt in (np.uint32, np.float32, np.float32, np.float32))
        self.ctx.check(lib().cx_trace_closest(self.ctx.h, self.h, d_rays.ptr, n, prim.ptr, t.ptr, u.ptr, v.ptr, stream),
                       "cx_trace_closest")
        synchronize()
        return prim.numpy(), t.numpy(), u.numpy(), v.numpy()

    def trace_occluded(self, rays: np.ndarray, stream=None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        d_rays = DeviceArray.from_numpy(rays)
        occ = DeviceArray((n,), np.uint32)
        self.ctx.check(lib().cx_trace_occluded(self.ctx.h, self.h, d_rays.ptr, n, occ.ptr, stream), "cx_trace_occluded")
        synchronize()
        return occ.numpy()

    def __del__(self):
        if getattr(self, "h", None):
            lib().cx_scene_destroy(self.h)
            self.h = None


class Entries:
    """Prefix-entry search structure (PrefixAccelStruct, src/PrefixAccelStruct.cpp)."""

    def __init__(self, ctx: Context, capacity: int):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().cx_entries_create(ctx.h, capacity, C.byref(h)), "cx_entries_create")
        self.h = h.value
        self.capacity = capacity

    def build(self, pos, radius, valid, scene_min, scene_max):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        self._d = [DeviceArray.from_numpy(pos), DeviceArray.from_numpy(np.ascontiguousarray(radius, np.float32)),
                   DeviceArray.from_numpy(np.ascontiguousarray(valid, np.uint8))]
        mn = np.ascontiguousarray(scene_min, np.float32)
        mx = np.ascontiguousarray(scene_max, np.float32)
        self.ctx.check(lib().cx_entries_build(self.ctx.h, self.h, self._d[0].ptr, self._d[1].ptr, self._d[2].ptr, n,
                                              mn.ctypes.data, mx.ctypes.data, None), "cx_entries_build")

    def query(self, q, k: int):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        nq = q.shape[0]
        dq = DeviceArray.from_numpy(q)
        idx = DeviceArray((nq, k), np.uint32)
        dist = DeviceArray((nq, k), np.float32)
        cnt = DeviceArray((nq,), np.uint32)
        self.ctx.check(lib().cx_entries_query(self.ctx.h, self.h, dq.ptr, nq, k, idx.ptr, dist.ptr, cnt.ptr, None),
                       "cx_entries_query")
        synchronize()
        return idx.numpy(), dist.numpy(), cnt.numpy()

    def __del__(self):
        if getattr(self, "h", None):
            lib().cx_entries_destroy(self.h)
            self.h = None


def entries_search_static(ctx: Context, pos, radius, valid, scene_min, scene_max, q, k: int):
    """cx_entries_search_static: the entry structure built and queried in one call that
    allocates no device memory in the library (the arrays here are the caller's).  Returns
    (idx [nq, k], dist [nq, k], count [nq]) as Entries.build + Entries.query do."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
    n, nq = pos.shape[0], q.shape[0]
    d = [DeviceArray.from_numpy(pos), DeviceArray.from_numpy(np.ascontiguousarray(radius, np.float32)),
         DeviceArray.from_numpy(np.ascontiguousarray(valid, np.uint8)), DeviceArray.from_numpy(q)]
    idx = DeviceArray((nq, k), np.uint32)
    dist = DeviceArray((nq, k), np.float32)
    cnt = DeviceArray((nq,), np.uint32)
    mn = np.ascontiguousarray(scene_min, np.float32)
    mx = np.ascontiguousarray(scene_max, np.float32)
    ctx.check(lib().cx_entries_search_static(ctx.h, d[0].ptr, d[1].ptr, d[2].ptr, n, mn.ctypes.data, mx.ctypes.data,
                                             d[3].ptr, nq, k, idx.ptr, dist.ptr, cnt.ptr, None),
              "cx_entries_search_static")
    synchronize()
    return idx.numpy(), dist.numpy(), cnt.numpy()


class Renderer:
    """Renderer(w, h, ctx, Camera&, const Scene&) and LaunchFrame (src/Renderer.cpp:13-406)."""

    def __init__(self, width: int, height: int, ctx: Context, camera, scene: Scene,
                 params: abi.LaunchParams | None = None, seed: int = 0):
        self.ctx = ctx
        self.scene = scene
        self.camera = camera
        self.W, self.H = width, height
        h = C.c_void_p()
        ctx.check(lib().cx_renderer_create(ctx.h, scene.h, width, height, C.byref(h)), "cx_renderer_create")
        self.h = h.value
        if params is not None:
            self.set_params(params)
        lib().cx_renderer_seed(self.h, seed)
        self.output = DeviceArray((height, width, 3), np.float32)
        self.output.zero()

    def set_params(self, params: abi.LaunchParams):
        self.ctx.check(lib().cx_renderer_set_params(self.h, C.byref(params)), "cx_renderer_set_params")

    def params(self) -> abi.LaunchParams:
        p = abi.LaunchParams()
        self.ctx.check(lib().cx_renderer_get_params(self.h, C.byref(p)), "cx_renderer_get_params")
        return p

    def seed(self, seed: int):
        lib().cx_renderer_seed(self.h, seed)

    def reset(self, stream=None):
        self.ctx.check(lib().cx_renderer_reset(self.h, stream), "cx_renderer_reset")

    def enable_timing(self, on: bool = True):
        lib().cx_renderer_enable_timing(self.h, 1 if on else 0)

    def set_exec_mode(self, mode: str):
        """'wavefront' (default) or 'megakernel'; both give identical results."""
        self.ctx.check(lib().cx_renderer_set_exec_mode(self.h, abi.EXEC_MODES[mode]), "cx_renderer_set_exec_mode")

    def set_queue_grid(self, max_blocks: int):
        self.ctx.check(lib().cx_renderer_set_queue_grid(self.h, max_blocks), "cx_renderer_set_queue_grid")

    def set_graph_mode(self, enable: bool = True):
        """Replay each frame from a captured hipGraph (cx_renderer_set_graph_mode); same images."""
        self.ctx.check(lib().cx_renderer_set_graph_mode(self.h, 1 if enable else 0), "cx_renderer_set_graph_mode")

    def set_overlap(self, enable: bool = True):
        """Entry build and neighbour query beside the gather walks (cx_renderer_set_overlap)."""
        self.ctx.check(lib().cx_renderer_set_overlap(self.h, 1 if enable else 0), "cx_renderer_set_overlap")

    def set_wave_rounds(self, rounds: int):
        """Trace/step rounds per machine before the in-place tail kernel (0: all rounds)."""
        self.ctx.check(lib().cx_renderer_set_wave_rounds(self.h, rounds), "cx_renderer_set_wave_rounds")

    def set_device_random(self, enable: bool = True, seed: int = 0):
        """Frames draw their random number on the device (cx_renderer_set_device_random), so a
        frame recorded into a caller's graph draws a new one on every replay."""
        self.ctx.check(lib().cx_renderer_set_device_random(self.h, 1 if enable else 0, seed),
                       "cx_renderer_set_device_random")

    def graph_captures(self) -> int:
        return int(lib().cx_renderer_graph_captures(self.h))

    def exec_mode(self) -> str:
        m = lib().cx_renderer_get_exec_mode(self.h)
        return {v: k for k, v in abi.EXEC_MODES.items()}[m]

    def launch_frame(self, cam: abi.CameraData | None = None, changed: bool | None = None, output=None,
                     stream=None):
        if cam is None:
            cam = abi.CameraData.from_dict(self.camera.data())
        if changed is None:
            changed = bool(getattr(self.camera, "changed", False))
        out = self.output if output is None else output
        self.ctx.check(lib().cx_renderer_frame(self.h, C.byref(cam), 1 if changed else 0, out.ptr, stream),
                       "cx_renderer_frame")
        return out

    def adjoint(self, dLdI=None, g_albedo=None, g_rough=None, g_emit=None, replay=None, status=None, stream=None):
        """cx_renderer_adjoint over device buffers (DeviceArray or raw device pointers)."""
        ptr = lambda a: None if a is None else (a if isinstance(a, int) else a.ptr)
        self.ctx.check(lib().cx_renderer_adjoint(self.h, ptr(dLdI), ptr(g_albedo), ptr(g_rough), ptr(g_emit),
                                                 ptr(replay), ptr(status), stream), "cx_renderer_adjoint")

    def stats(self) -> abi.FrameStats:
        st = abi.FrameStats()
        self.ctx.check(lib().cx_renderer_get_stats(self.h, C.byref(st)), "cx_renderer_get_stats")
        return st

    def state(self) -> abi.StateView:
        v = abi.StateView()
        self.ctx.check(lib().cx_renderer_state(self.h, C.byref(v)), "cx_renderer_state")
        return v

    def prefix(self, which: int) -> dict:
        v = self.state()
        P = self.W * self.H
        raw = download(v.prefixRes[which], abi.PREFIX_CHUNKS * P * 16, np.uint32)
        return abi.decode_prefix(raw, P)

    def suffix(self, which: int) -> dict:
        v = self.state()
        P = self.W * self.H
        raw = download(v.suffixRes[which], abi.SUFFIX_CHUNKS * P * 16, np.uint32)
        return abi.decode_suffix(raw, P)

    def gbuffer(self) -> dict:
        v = self.state()
        P = self.W * self.H
        return abi.decode_gbuffer(download(v.gbuffer, P * 16, np.uint32), P)

    def entries(self):
        v = self.state()
        P = self.W * self.H
        e = download(v.entryPos, P * 16, np.float32).reshape(P, 4)
        val = download(v.entryValid, P, np.uint8)
        return e[:, :3], e[:, 3], val

    def neighbors(self):
        v = self.state()
        P = self.W * self.H
        k = max(1, v.k)
        idx = download(v.neighIdx, P * k * 4, np.uint32).reshape(P, k)
        dist = download(v.neighDist, P * k * 4, np.float32).reshape(P, k)
        cnt = download(v.neighCount, P * 4, np.uint32)
        return idx[:, :v.k], dist[:, :v.k], cnt

    def __del__(self):
        if getattr(self, "h", None):
            lib().cx_renderer_destroy(self.h)
            self.h = None
