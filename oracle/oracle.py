"""ORACLE -- test infrastructure only.

ctypes wrapper of the host C++ restatement of the reference shader logic
(oracle/crisoptix_oracle.cpp).  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
of bench.py import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from crisoptix_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborc.so")

PACKED_DT = np.dtype([("prim", "<u4"), ("bu", "<f4"), ("bv", "<f4"), ("inDir", "<f4", 3), ("valid", "<u4")])
PREFIX_DT = np.dtype([
    ("rngState", "<u8"), ("flags", "<u4"), ("f", "<f4", 3), ("postReconF", "<f4", 3), ("p", "<f4"),
    ("pathLen", "<f4"), ("reconOutDir", "<f4", 3), ("wSum", "<f4"), ("confidence", "<f4"),
    ("primaryInt", PACKED_DT), ("reconInt", PACKED_DT), ("lastInt", PACKED_DT)], align=True)
SUFFIX_DT = np.dtype([
    ("rngState", "<u8"), ("flags", "<u4"), ("f", "<f4", 3), ("postReconF", "<f4", 3), ("p", "<f4"),
    ("reconOutDir", "<f4", 3), ("wSum", "<f4"), ("confidence", "<f4"),
    ("reconInt", PACKED_DT), ("lastPrefixInt", PACKED_DT)], align=True)


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, u32, u64, f32, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float, C.c_int
        L.orc_scene_create.restype = vp
        L.orc_gather_combine.argtypes = [vp, C.c_float, vp, vp, vp, u32, vp]
        L.orc_scene_create.argtypes = [C.POINTER(abi.SceneDesc), i32]
        L.orc_scene_destroy.argtypes = [vp]
        L.orc_scene_size.restype = f32
        L.orc_scene_size.argtypes = [vp]
        L.orc_trace_closest.argtypes = [vp, vp, u32, vp, vp, vp, vp]
        L.orc_trace_occluded.argtypes = [vp, vp, u32, vp]
        L.orc_tea32.restype = u32
        L.orc_tea32.argtypes = [u32, u32, u32]
        L.orc_tea64.restype = u64
        L.orc_tea64.argtypes = [u32, u32, u32]
        L.orc_tea_keyed.argtypes = [vp, vp, u32]
        L.orc_pcg32_stream.argtypes = [u64, u64, i32, u32, vp]
        L.orc_pcg32_state.argtypes = [u64, u64, vp]
        L.orc_sincos.argtypes = [vp, u32, vp, vp]
        L.orc_ggx_eval.argtypes = [vp, vp, vp, u32, vp]
        L.orc_ggx_sample.argtypes = [vp, vp, vp, u32, vp]
        L.orc_sample_emitter.argtypes = [vp, vp, u32, vp]
        L.orc_jacobian.restype = f32
        L.orc_jacobian.argtypes = [vp, vp, vp, vp]
        L.orc_pairwise_mis.argtypes = [vp, vp, f32, f32, f32, f32, vp, vp]
        L.orc_reservoir_stream.restype = i32
        L.orc_reservoir_stream.argtypes = [vp, vp, u32, u64, vp, vp]
        L.orc_identity_shift_prefix.argtypes = [vp, C.POINTER(abi.LaunchParams), vp, vp, u64, vp]
        L.orc_identity_shift_suffix.argtypes = [vp, C.POINTER(abi.LaunchParams), vp, vp, u64, vp]
        L.orc_renderer_create.restype = vp
        L.orc_renderer_create.argtypes = [vp, u32, u32]
        L.orc_renderer_destroy.argtypes = [vp]
        L.orc_renderer_set_params.argtypes = [vp, C.POINTER(abi.LaunchParams)]
        L.orc_renderer_seed.argtypes = [vp, u64]
        L.orc_renderer_reset.argtypes = [vp]
        L.orc_renderer_frame.argtypes = [vp, C.POINTER(abi.CameraData), i32, vp]
        L.orc_renderer_pass_ms.argtypes = [vp, vp]
        L.orc_renderer_rays.argtypes = [vp, vp]
        L.orc_renderer_buffers.argtypes = [vp, vp, vp]
        L.orc_dump_prefix.argtypes = [vp, i32, vp]
        L.orc_dump_suffix.argtypes = [vp, i32, vp]
        L.orc_dump_gbuffer.argtypes = [vp, vp]
        L.orc_dump_entries.argtypes = [vp, vp, vp, vp]
        L.orc_dump_neighbors.restype = u32
        L.orc_dump_neighbors.argtypes = [vp, vp, vp, vp]
        L.orc_entries_query.argtypes = [vp, vp, vp, u32, vp, u32, u32, i32, vp, vp, vp]
        L.orc_truth_render.argtypes = [vp, C.POINTER(abi.CameraData), u32, u32, u32, C.c_double, u32, u64, vp, vp]
        L.orc_renderer_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.orc_scene_set_materials.argtypes = [vp, vp, u32, i32]
        L.orc_scene_set_emitter_colors.argtypes = [vp, vp, u32, i32]
        L.orc_tonemap.argtypes = [vp, vp, u32]
        L.orc_texture_sample.argtypes = [vp, vp, u32, vp]
        L.orc_num_threads.restype = i32
        L.orc_set_num_threads.argtypes = [i32]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(n: int):
    lib().orc_set_num_threads(n)


def tonemap(hdr: np.ndarray) -> np.ndarray:
    h = np.ascontiguousarray(hdr, np.float32)
    out = np.zeros(h.shape, np.uint8)
    lib().orc_tonemap(_p(h), _p(out), h.size)
    return out


def texture_sample(rgba: np.ndarray, uv: np.ndarray) -> np.ndarray:
    """The canonical bilinear wrap sampler on one RGBA8 image [H, W, 4] at uv [n, 2]; [n, 3]."""
    img = np.ascontiguousarray(rgba, np.uint8)
    t = abi.Texture(img.ctypes.data, img.shape[1], img.shape[0])
    q = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    out = np.zeros((q.shape[0], 3), np.float32)
    lib().orc_texture_sample(C.byref(t), _p(q), q.shape[0], _p(out))
    return out


def tea32(a, b, rounds=4):
    return lib().orc_tea32(a, b, rounds)


def tea64(a, b, rounds=4):
    return lib().orc_tea64(a, b, rounds)


def tea_keyed(v0, v1, key, rounds):
    v = np.asarray([v0, v1], np.uint32)
    k = np.asarray(key, np.uint32)
    lib().orc_tea_keyed(_p(v), _p(k), rounds)
    return int(v[0]), int(v[1])


def pcg32_stream(initstate, initseq, n, kind=0):
    out = np.zeros(n * (2 if kind == 2 else 1), np.uint32)
    lib().orc_pcg32_stream(initstate, initseq, kind, n, _p(out))
    if kind == 1:
        return out.view(np.float32)
    if kind == 2:
        return out.view(np.uint64)
    return out


def pcg32_state(initstate, initseq=0xDA3E39CB94B95BDB):
    o = np.zeros(2, np.uint64)
    lib().orc_pcg32_state(initstate, initseq, _p(o))
    return int(o[0]), int(o[1])


def sincos(x):
    x = np.ascontiguousarray(x, np.float32)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    lib().orc_sincos(_p(x), x.size, _p(s), _p(c))
    return s, c


class OracleScene:
    def __init__(self, scene, brute_force: bool = False):
        self._keep = []
        self.desc = abi.scene_desc(scene, self._keep)
        self.h = lib().orc_scene_create(C.byref(self.desc), 1 if brute_force else 0)
        self.scene = scene

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    @property
    def size(self) -> float:
        return lib().orc_scene_size(self.h)

    def trace_closest(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        lib().orc_trace_closest(self.h, _p(rays), n, _p(prim), _p(t), _p(u), _p(v))
        return prim, t, u, v

    def trace_occluded(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        occ = np.zeros(rays.shape[0], np.uint32)
        lib().orc_trace_occluded(self.h, _p(rays), rays.shape[0], _p(occ))
        return occ

    def ggx_eval(self, inter: np.ndarray, out_dir: np.ndarray):
        inter = np.ascontiguousarray(inter, np.float32)
        out_dir = np.ascontiguousarray(out_dir, np.float32)
        n = inter.shape[0]
        out = np.zeros((n, 8), np.float32)
        lib().orc_ggx_eval(self.h, _p(inter), _p(out_dir), n, _p(out))
        return out

    def ggx_sample(self, inter: np.ndarray, rng_state: np.ndarray):
        inter = np.ascontiguousarray(inter, np.float32)
        st = np.ascontiguousarray(rng_state, np.uint64).copy()
        n = inter.shape[0]
        out = np.zeros((n, 8), np.float32)
        lib().orc_ggx_sample(self.h, _p(inter), _p(st), n, _p(out))
        return out, st

    def set_materials(self, materials: np.ndarray, tables: int = 3):
        """[M, 10] cx_material rows; tables: 1 sampling, 2 evaluation (the FP64 adjoint's), 3 both."""
        m = np.ascontiguousarray(materials, np.float32).reshape(-1, 10)
        lib().orc_scene_set_materials(self.h, _p(m), m.shape[0], tables)

    def set_emitter_colors(self, colors: np.ndarray, tables: int = 3):
        c = np.ascontiguousarray(colors, np.float32).reshape(-1, 3)
        lib().orc_scene_set_emitter_colors(self.h, _p(c), c.shape[0], tables)

    def sample_emitter(self, rng_state: np.ndarray):
        st = np.ascontiguousarray(rng_state, np.uint64).copy()
        out = np.zeros((st.size, 4), np.float32)
        lib().orc_sample_emitter(self.h, _p(st), st.size, _p(out))
        return out, st

    def identity_shift(self, params: abi.LaunchParams, origin, direction, rng_state: int, which="prefix"):
        o = np.ascontiguousarray(origin, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        # f, shifted f, J, valid; the suffix's reconIdx and length, the suffix vertex of a BSDF
        # reconnection (0: none) and the prefix length
        out = np.zeros(12, np.float32)
        fn = lib().orc_identity_shift_prefix if which == "prefix" else lib().orc_identity_shift_suffix
        fn(self.h, C.byref(params), _p(o), _p(d), rng_state, _p(out))
        return out


def truth_render(oscene: "OracleScene", cam: abi.CameraData, width: int, height: int, max_path_len: int,
                 nee_prob: float, spp: int, seed: int = 0):
    """FP64 truth of the PathTracer estimator (path_gen.h:389-472): (mean, variance of the
    mean), both [height][width][3] float64."""
    out = np.zeros((height, width, 3), np.float64)
    var = np.zeros((height, width, 3), np.float64)
    lib().orc_truth_render(oscene.h, C.byref(cam), width, height, max_path_len, float(nee_prob), spp, seed,
                           _p(out), _p(var))
    return out, var


def rel_mse(img, ref) -> float:
    """mean over pixels and channels of (I - I_ref)^2 / (I_ref^2 + 1e-2)."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


def jacobian(old, new, tp, tn) -> float:
    a = [np.ascontiguousarray(x, np.float32) for x in (old, new, tp, tn)]
    return lib().orc_jacobian(*[_p(x) for x in a])


def pairwise_mis(f_basis, f_neigh, J, K, canon_m, neigh_m):
    a = np.ascontiguousarray(f_basis, np.float32)
    b = np.ascontiguousarray(f_neigh, np.float32)
    cw = C.c_float()
    nw = C.c_float()
    lib().orc_pairwise_mis(_p(a), _p(b), J, K, canon_m, neigh_m, C.byref(cw), C.byref(nw))
    return cw.value, nw.value


def gather_combine(canon_pre_f, canon_pre_p, canon_suf_f, canon_contrib, terms):
    """The final gather's pairwise-MIS combination (final_gather.cu:135-187) over injected neighbour
    terms: rows of fFromCanonOfNeigh[3], jNeighToCanon, neighF[3], neighWSum, neighConf,
    fFromNeighOfCanon[3], jCanonToNeigh."""
    t = np.ascontiguousarray(terms, np.float32).reshape(-1, 13)
    a = [np.ascontiguousarray(x, np.float32) for x in (canon_pre_f, canon_suf_f, canon_contrib)]
    out = np.zeros(3, np.float32)
    lib().orc_gather_combine(_p(a[0]), C.c_float(canon_pre_p), _p(a[1]), _p(a[2]), _p(t), t.shape[0], _p(out))
    return out


def reservoir_stream(weights, confidences, rng_state: int):
    w = np.ascontiguousarray(weights, np.float32)
    c = np.ascontiguousarray(confidences, np.float32)
    ws = C.c_float()
    cf = C.c_float()
    idx = lib().orc_reservoir_stream(_p(w), _p(c), w.size, rng_state, C.byref(ws), C.byref(cf))
    return idx, ws.value, cf.value


def entries_query(pos, radius, valid, queries, k, brute=False):
    pos = np.ascontiguousarray(pos, np.float32)
    radius = np.ascontiguousarray(radius, np.float32)
    valid = np.ascontiguousarray(valid, np.uint8)
    q = np.ascontiguousarray(queries, np.float32)
    nq = q.shape[0]
    idx = np.zeros((nq, max(k, 1)), np.uint32)
    dist = np.zeros((nq, max(k, 1)), np.float32)
    cnt = np.zeros(nq, np.uint32)
    lib().orc_entries_query(_p(pos), _p(radius), _p(valid), pos.shape[0], _p(q), nq, k, 1 if brute else 0,
                            _p(idx), _p(dist), _p(cnt))
    return idx[:, :k], dist[:, :k], cnt


class OracleRenderer:
    """Renderer(w, h, ctx, Camera&, const Scene&) + LaunchFrame (src/Renderer.cpp), on the CPU."""

    def __init__(self, oscene: OracleScene, width: int, height: int, params: abi.LaunchParams, seed: int = 0):
        self.oscene = oscene
        self.W, self.H = width, height
        self.h = lib().orc_renderer_create(oscene.h, width, height)
        self.params = params.copy()
        lib().orc_renderer_set_params(self.h, C.byref(self.params))
        lib().orc_renderer_seed(self.h, seed)
        self.output = np.zeros((height, width, 3), np.float32)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_renderer_destroy(self.h)
            self.h = None

    def set_params(self, params: abi.LaunchParams):
        self.params = params.copy()
        lib().orc_renderer_set_params(self.h, C.byref(self.params))

    def frame(self, cam: abi.CameraData, changed: bool = False) -> np.ndarray:
        lib().orc_renderer_frame(self.h, C.byref(cam), 1 if changed else 0, _p(self.output))
        return self.output

    def pass_ms(self) -> np.ndarray:
        ms = np.zeros(abi.CX_PASS_COUNT, np.float64)
        lib().orc_renderer_pass_ms(self.h, _p(ms))
        return ms

    def rays(self) -> np.ndarray:
        """[pass][closest, shadow] ray counts of the last frame."""
        out = np.zeros(2 * abi.CX_PASS_COUNT, np.uint64)
        lib().orc_renderer_rays(self.h, _p(out))
        return out.reshape(abi.CX_PASS_COUNT, 2)

    def buffers(self):
        fb = np.zeros(2, np.uint32)
        fr = np.zeros(2, np.uint64)
        lib().orc_renderer_buffers(self.h, _p(fb), _p(fr))
        return int(fb[0]), int(fb[1]), int(fr[0]), int(fr[1])

    def prefix(self, which: int) -> np.ndarray:
        out = np.zeros(self.W * self.H, PREFIX_DT)
        lib().orc_dump_prefix(self.h, which, _p(out))
        return out

    def suffix(self, which: int) -> np.ndarray:
        out = np.zeros(self.W * self.H, SUFFIX_DT)
        lib().orc_dump_suffix(self.h, which, _p(out))
        return out

    def gbuffer(self) -> np.ndarray:
        out = np.zeros((self.W * self.H, 4), np.uint32)
        lib().orc_dump_gbuffer(self.h, _p(out))
        return out

    def entries(self):
        P = self.W * self.H
        pos = np.zeros((P, 3), np.float32)
        rad = np.zeros(P, np.float32)
        val = np.zeros(P, np.uint8)
        lib().orc_dump_entries(self.h, _p(pos), _p(rad), _p(val))
        return pos, rad, val

    def adjoint(self, dLdI=None):
        """FP64 path-replay adjoint of the last frame: (g_albedo [M,3], g_rough [M], g_emit [E,3],
        replayed estimate [H,W,3] float64, status [replayed, -, skipped])."""
        sc = self.oscene.scene
        M, E = sc.materials.shape[0], sc.emitter_color.shape[0]
        gA, gR, gE = np.zeros((M, 3)), np.zeros(M), np.zeros((E, 3))
        rep = np.zeros((self.H, self.W, 3))
        st = np.zeros(3, np.uint32)
        g = None if dLdI is None else np.ascontiguousarray(dLdI, np.float32)
        lib().orc_renderer_adjoint(self.h, None if g is None else _p(g), _p(gA), _p(gR), _p(gE), _p(rep), _p(st))
        return gA, gR, gE, rep, st

    def neighbors(self):
        P = self.W * self.H
        k = max(1, self.params.gatherM - 1)
        idx = np.zeros((P, k), np.uint32)
        dist = np.zeros((P, k), np.float32)
        cnt = np.zeros(P, np.uint32)
        kk = lib().orc_dump_neighbors(self.h, _p(idx), _p(dist), _p(cnt))
        return idx[:, :kk], dist[:, :kk], cnt
