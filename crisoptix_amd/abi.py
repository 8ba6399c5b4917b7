"""ctypes mirrors of the POD structs declared in include/crisoptix.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

CX_PATH_TRACER, CX_RESTIR_PT, CX_CONDITIONAL_RESTIR = 0, 1, 2
CX_RADIUS_CONSTANT, CX_RADIUS_PATH_LENGTH = 0, 1
CX_EXEC_WAVEFRONT, CX_EXEC_MEGAKERNEL = 0, 1
EXEC_MODES = {"wavefront": CX_EXEC_WAVEFRONT, "megakernel": CX_EXEC_MEGAKERNEL}
PASS_NAMES = ["prefix_gen_temp", "prefix_spatial", "suffix_gen_temp", "suffix_spatial",
              "prefix_store_entries", "entries_build", "final_gather"]
CX_PASS_COUNT = len(PASS_NAMES)


class LaunchParams(C.Structure):
    _fields_ = [
        ("random", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("windowRadius", C.c_uint32),
        ("frameIdx", C.c_uint64), ("neeProb", C.c_float), ("enableAccum", C.c_int32),
        ("maxPathLen", C.c_int32), ("rendererType", C.c_int32),
        ("reconMinRoughness", C.c_float), ("reconMinDistance", C.c_float), ("spatialRoundIdx", C.c_int32),
        ("prefixLen", C.c_int32), ("prefixEnableTemporal", C.c_int32), ("prefixEnableSpatial", C.c_int32),
        ("prefixSpatialCount", C.c_int32), ("prefixSpatialRounds", C.c_int32),
        ("suffixEnableTemporal", C.c_int32), ("suffixEnableSpatial", C.c_int32),
        ("suffixSpatialCount", C.c_int32), ("suffixSpatialRounds", C.c_int32),
        ("gatherN", C.c_int32), ("gatherM", C.c_int32), ("gatherRadius", C.c_float),
        ("gatherRadiusType", C.c_int32), ("trackPrefixStats", C.c_int32), ("pad0", C.c_int32),
    ]

    def copy(self) -> "LaunchParams":
        p = LaunchParams()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(LaunchParams))
        return p

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class CameraData(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("U", C.c_float * 3), ("V", C.c_float * 3), ("W", C.c_float * 3)]

    @staticmethod
    def from_dict(d: dict) -> "CameraData":
        c = CameraData()
        for k in ("pos", "U", "V", "W"):
            getattr(c, k)[:] = [float(x) for x in np.asarray(d[k], np.float32)]
        return c


class Material(C.Structure):
    _fields_ = [("diffColor", C.c_float * 3), ("specF0", C.c_float * 3), ("emissiveColor", C.c_float * 3),
                ("roughness", C.c_float)]


ABI_VERSION = 10   # CX_ABI_VERSION of include/crisoptix.h
CX_SPEC_TEX_COLOR = 0
CX_SPEC_TEX_OCC_ROUGH_METAL = 1
CX_WAVE_ROUNDS_DEFAULT = 0xFFFFFFFF   # per-machine default wavefront rounds
CX_VIS_SURFACE, CX_VIS_OCCLUSION, CX_VIS_ALL = 1, 2, 3   # cx_scene_desc.triMask bits
CX_BVH_BUILD_HOST = 0
CX_BVH_BUILD_DEVICE = 1
BVH_BUILDERS = {"host": CX_BVH_BUILD_HOST, "device": CX_BVH_BUILD_DEVICE}


class SceneDesc(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p), ("normals", C.c_void_p), ("numVertices", C.c_uint32),
        ("indices", C.c_void_p), ("triMaterial", C.c_void_p), ("numTriangles", C.c_uint32),
        ("materials", C.c_void_p), ("numMaterials", C.c_uint32),
        ("emitterFaces", C.c_void_p), ("emitterFaceOffset", C.c_void_p), ("emitterColor", C.c_void_p),
        ("numEmitters", C.c_uint32), ("bvhMaxLeafSize", C.c_uint32), ("bvhWidth", C.c_uint32),
        ("uvs", C.c_void_p), ("materialTex", C.c_void_p), ("textures", C.c_void_p), ("numTextures", C.c_uint32),
        ("bvhBuilder", C.c_uint32), ("triMask", C.c_void_p),
    ]


class Texture(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32)]


class MaterialTex(C.Structure):
    """One row of Scene.material_tex (an int32 [M, 4] array is passed as this struct's array)."""
    _fields_ = [("diffTex", C.c_int32), ("specTex", C.c_int32), ("roughTex", C.c_int32),
                ("specTexUsage", C.c_uint32)]


class SceneInfo(C.Structure):
    _fields_ = [
        ("numTriangles", C.c_uint32), ("numVertices", C.c_uint32), ("numNodes", C.c_uint32),
        ("maxDepth", C.c_uint32), ("numEmitters", C.c_uint32), ("numEmitterFaces", C.c_uint32),
        ("sceneSize", C.c_float), ("boundsMin", C.c_float * 3), ("boundsMax", C.c_float * 3),
        ("deviceBytes", C.c_uint64), ("bvhWidth", C.c_uint32), ("numWideNodes", C.c_uint32),
        ("stackBound", C.c_uint32), ("numTextures", C.c_uint32), ("bvhBuilder", C.c_uint32),
        ("buildMs", C.c_float),
    ]


class FrameStats(C.Structure):
    _fields_ = [
        ("passMs", C.c_float * CX_PASS_COUNT), ("totalMs", C.c_float), ("closestRays", C.c_uint64),
        ("shadowRays", C.c_uint64), ("passClosestRays", C.c_uint64 * CX_PASS_COUNT),
        ("passShadowRays", C.c_uint64 * CX_PASS_COUNT), ("minNeighCount", C.c_uint32), ("maxNeighCount", C.c_uint32),
        ("totalNeighCount", C.c_uint64), ("numEntries", C.c_uint32), ("traceLaunches", C.c_uint32),
        ("traceMs", C.c_float), ("queueOverflow", C.c_uint32),
        ("tailClosestRays", C.c_uint64), ("tailShadowRays", C.c_uint64),
    ]


class StateView(C.Structure):
    _fields_ = [
        ("prefixRes", C.c_void_p * 2), ("suffixRes", C.c_void_p * 2), ("gbuffer", C.c_void_p),
        ("neighIdx", C.c_void_p), ("neighDist", C.c_void_p), ("neighCount", C.c_void_p),
        ("entryPos", C.c_void_p), ("entryValid", C.c_void_p),
        ("frontBufferIdx", C.c_uint32), ("backBufferIdx", C.c_uint32), ("prefixChunks", C.c_uint32),
        ("suffixChunks", C.c_uint32), ("k", C.c_uint32), ("pad0", C.c_uint32),
    ]


def scene_desc(scene, keep: list) -> SceneDesc:
    """Builds a SceneDesc over the arrays of a scenes.Scene; `keep` holds references."""
    arrs = {
        "positions": np.ascontiguousarray(scene.positions, np.float32),
        "normals": np.ascontiguousarray(scene.normals, np.float32),
        "indices": np.ascontiguousarray(scene.indices, np.uint32),
        "triMaterial": np.ascontiguousarray(scene.tri_material, np.uint32),
        "materials": np.ascontiguousarray(scene.materials, np.float32),
        "emitterFaces": np.ascontiguousarray(scene.emitter_faces, np.float32),
        "emitterFaceOffset": np.ascontiguousarray(scene.emitter_face_offset, np.uint32),
        "emitterColor": np.ascontiguousarray(scene.emitter_color, np.float32),
    }
    keep.append(arrs)
    d = SceneDesc()
    for k, a in arrs.items():
        setattr(d, k, a.ctypes.data)
    d.numVertices = arrs["positions"].shape[0]
    d.numTriangles = arrs["indices"].shape[0]
    d.numMaterials = arrs["materials"].shape[0]
    d.numEmitters = arrs["emitterColor"].shape[0]
    d.bvhMaxLeafSize = 0
    d.bvhWidth = 0
    tri_mask = getattr(scene, "tri_mask", None)
    if tri_mask is not None:
        arrs["triMask"] = np.ascontiguousarray(tri_mask, np.uint8).reshape(-1)
        if arrs["triMask"].shape[0] != d.numTriangles:
            raise ValueError("tri_mask must have one entry per triangle")
        d.triMask = arrs["triMask"].ctypes.data
    uvs = getattr(scene, "uvs", None)
    if uvs is not None:
        arrs["uvs"] = np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
        if arrs["uvs"].shape[0] != d.numVertices:
            raise ValueError("uvs must have one (u, v) per vertex")
        d.uvs = arrs["uvs"].ctypes.data
    mt = getattr(scene, "material_tex", None)
    if mt is not None:
        arrs["materialTex"] = np.ascontiguousarray(mt, np.int32).reshape(-1, 4)
        if arrs["materialTex"].shape[0] != d.numMaterials:
            raise ValueError("material_tex must have one row per material")
        d.materialTex = arrs["materialTex"].ctypes.data
    texs = list(getattr(scene, "textures", None) or [])
    if texs:
        imgs = [np.ascontiguousarray(t, np.uint8) for t in texs]
        for im in imgs:
            if im.ndim != 3 or im.shape[2] != 4:
                raise ValueError("textures are [height, width, 4] uint8 RGBA arrays")
        tab = (Texture * len(imgs))()
        for i, im in enumerate(imgs):
            tab[i].rgba = im.ctypes.data
            tab[i].width = im.shape[1]
            tab[i].height = im.shape[0]
        arrs["_textures"] = (imgs, tab)
        d.textures = C.addressof(tab)
        d.numTextures = len(imgs)
    return d


def default_params(width: int, height: int, scene_size: float) -> LaunchParams:
    """Renderer defaults (src/Renderer.cpp:34-58) with WINDOW_RADIUS = 10
    (include/graph/restir/path_gen.h:12) and accumulation off (Renderer.h:32 value-initialises
    LaunchParams)."""
    p = LaunchParams()
    p.width, p.height = width, height
    p.windowRadius = 10
    p.neeProb = 0.3
    p.maxPathLen = 8
    p.rendererType = CX_PATH_TRACER
    p.reconMinDistance = np.float32(0.01) * np.float32(scene_size)
    p.reconMinRoughness = 0.2
    p.prefixLen = 2
    p.prefixEnableTemporal = 1
    p.prefixEnableSpatial = 0
    p.prefixSpatialCount = 1
    p.prefixSpatialRounds = 1
    p.suffixEnableTemporal = 1
    p.suffixEnableSpatial = 1
    p.suffixSpatialCount = 1
    p.suffixSpatialRounds = 1
    p.gatherN = 1
    p.gatherM = 1
    p.gatherRadius = 0.01
    p.gatherRadiusType = CX_RADIUS_PATH_LENGTH
    p.trackPrefixStats = 0
    p.enableAccum = 0
    return p


def workload_params(name: str, width: int, height: int, scene_size: float) -> LaunchParams:
    """Renderer settings of the benchmark workloads."""
    p = default_params(width, height, scene_size)
    if name == "restir_pt_cornell":
        # ReSTIR PT, max path length 4, temporal + spatial reuse with 3 neighbours in 30 px.
        p.rendererType = CX_RESTIR_PT
        p.maxPathLen = 4
        p.prefixEnableTemporal = 1
        p.prefixEnableSpatial = 1
        p.prefixSpatialCount = 3
        p.prefixSpatialRounds = 1
        p.windowRadius = 30
    elif name == "conditional_restir":
        # Conditional ReSTIR, max path length 8, prefix length 2, 8 gathered suffixes.
        p.rendererType = CX_CONDITIONAL_RESTIR
        p.maxPathLen = 8
        p.prefixLen = 2
        p.gatherM = 9
        p.gatherN = 1
        p.gatherRadius = np.float32(0.01) * np.float32(scene_size)
    elif name == "path_tracer":
        p.rendererType = CX_PATH_TRACER
    else:
        raise KeyError(name)
    return p


# Record layouts of the renderer's per-pixel state (chunked structure-of-arrays: chunk c of
# pixel p lives at base + (c * P + p) * 16).  Field order within the chunks is documented in
# crisoptix_amd/csrc/cx_records.h.
PREFIX_CHUNKS = 9
SUFFIX_CHUNKS = 7
GBUF_CHUNKS = 1


def decode_prefix(raw: np.ndarray, P: int) -> dict:
    """raw: uint32 array of PREFIX_CHUNKS * P * 4 words -> dict of per-pixel fields."""
    w = raw.reshape(PREFIX_CHUNKS, P, 4)
    f = w.view(np.float32)
    out = {
        "rngState": w[0, :, 0].astype(np.uint64) | (w[0, :, 1].astype(np.uint64) << np.uint64(32)),
        "flags": w[0, :, 2], "p": f[0, :, 3],
        "f": f[1, :, 0:3], "pathLen": f[1, :, 3],
        "postReconF": f[2, :, 0:3], "wSum": f[2, :, 3],
        "reconOutDir": f[3, :, 0:3], "confidence": f[3, :, 3],
        "primary_prim": w[4, :, 0], "primary_bu": f[4, :, 1], "primary_bv": f[4, :, 2],
        "primary_dir": np.stack([f[4, :, 3], f[5, :, 0], f[5, :, 1]], 1),
        "recon_prim": w[5, :, 2], "recon_bu": f[5, :, 3], "recon_bv": f[6, :, 0],
        "recon_dir": f[6, :, 1:4],
        "last_prim": w[7, :, 0], "last_bu": f[7, :, 1], "last_bv": f[7, :, 2],
        "last_dir": np.stack([f[7, :, 3], f[8, :, 0], f[8, :, 1]], 1),
    }
    return out


def decode_suffix(raw: np.ndarray, P: int) -> dict:
    w = raw.reshape(SUFFIX_CHUNKS, P, 4)
    f = w.view(np.float32)
    return {
        "rngState": w[0, :, 0].astype(np.uint64) | (w[0, :, 1].astype(np.uint64) << np.uint64(32)),
        "flags": w[0, :, 2], "p": f[0, :, 3],
        "f": f[1, :, 0:3], "wSum": f[1, :, 3],
        "postReconF": f[2, :, 0:3], "confidence": f[2, :, 3],
        "reconOutDir": f[3, :, 0:3], "recon_prim": w[3, :, 3],
        "recon_bu": f[4, :, 0], "recon_bv": f[4, :, 1], "recon_dir": np.stack([f[4, :, 2], f[4, :, 3], f[5, :, 0]], 1),
        "last_prim": w[5, :, 1], "last_bu": f[5, :, 2], "last_bv": f[5, :, 3],
        "last_dir": f[6, :, 0:3],
    }


def decode_gbuffer(raw: np.ndarray, P: int) -> dict:
    w = raw.reshape(P, 4)
    return {
        "rngState": w[:, 0].astype(np.uint64) | (w[:, 1].astype(np.uint64) << np.uint64(32)),
        "primaryIntValid": w[:, 2],
        "canonFlags": w[:, 3],
    }
