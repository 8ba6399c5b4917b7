"""Seeded procedural scenes and cameras for the benchmark workloads.

Every random draw comes from a PCG32 stream (the generator of include/util/random.h:45-119,
restated in :class:`Pcg32`) seeded from the workload seed, so the same arrays are produced
on every machine.  Geometry is generated in float64 and rounded once to float32.

Scenes follow the flattening of the renderer's C ABI: one triangle list with per-triangle
material ids (the reference's Model/Mesh/ModelInstance tree always rehydrates with
transforms[0], include/graph/Interaction.h:143), plus explicit emitters that NEE samples
(Scene::BuildEmitterTable, src/Scene.cpp:176-204).  With ``trace_emitters`` the emitter
triangles are also traced geometry, as the reference's scenes have them (implicit emitters are
mesh instances, src/Scene.cpp:184-197; an explicit Emitter wraps a traced mesh,
src/main.cpp:164-167): closest-hit rays see them and occlusion rays pass over them (their
``tri_mask`` is CX_VIS_SURFACE), so a shadow ray that ends on a light is not blocked by it.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1
PCG32_MULT = 0x5851F42D4C957F2D
PCG32_DEFAULT_STREAM = 0xDA3E39CB94B95BDB


class Pcg32:
    """PCG32 XSH-RR with pcg32_srandom seeding (include/util/random.h:52-81)."""

    def __init__(self, initstate: int, initseq: int = PCG32_DEFAULT_STREAM):
        self.state = 0
        self.inc = ((initseq << 1) | 1) & MASK64
        self.next_u32()
        self.state = (self.state + initstate) & MASK64
        self.next_u32()

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * PCG32_MULT + self.inc) & MASK64
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((32 - rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * (self.next_u32() / 4294967296.0)

    def normal(self) -> float:
        u1 = max(self.uniform(), 1e-12)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


@dataclass
class Scene:
    positions: np.ndarray          # float32 [V,3]
    normals: np.ndarray            # float32 [V,3]
    indices: np.ndarray            # uint32 [T,3]
    tri_material: np.ndarray       # uint32 [T]
    materials: np.ndarray          # float32 [M,10] diff rgb, specF0 rgb, emissive rgb, roughness
    emitter_faces: np.ndarray      # float32 [F,3,3] world-space emitter triangles
    emitter_face_offset: np.ndarray  # uint32 [E+1]
    emitter_color: np.ndarray      # float32 [E,3]
    name: str = ""
    meta: dict = field(default_factory=dict)
    uvs: np.ndarray | None = None          # float32 [V,2] texture coordinates, or None (all 0)
    material_tex: np.ndarray | None = None  # int32 [M,4] diffuse, specular, roughness texture
    #                                         (-1: none), specular usage (0 Color, 1 OccRoughMetal)
    textures: list = field(default_factory=list)  # uint8 [H,W,4] RGBA images
    tri_mask: np.ndarray | None = None     # uint8 [T] CX_VIS_* bits per triangle, or None (all 3)

    @property
    def num_triangles(self) -> int:
        return int(self.indices.shape[0])

    def scene_size(self) -> float:
        """Scene::GetSceneSize (src/Scene.cpp:48-71): bbox of all vertices and the origin."""
        mn = np.minimum(self.positions.min(axis=0), 0.0).astype(np.float32)
        mx = np.maximum(self.positions.max(axis=0), 0.0).astype(np.float32)
        d = (mx - mn).astype(np.float32)
        return float(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1]) + np.float32(d[2] * d[2])))


class _Builder:
    def __init__(self):
        self.pos: list[np.ndarray] = []
        self.nrm: list[np.ndarray] = []
        self.idx: list[np.ndarray] = []
        self.mat: list[np.ndarray] = []
        self.vis: list[np.ndarray] = []
        self.nv = 0
        self.materials: list[list[float]] = []

    def add_material(self, diff, rough, spec=(0.0, 0.0, 0.0), emissive=(0.0, 0.0, 0.0)) -> int:
        self.materials.append([*diff, *spec, *emissive, rough])
        return len(self.materials) - 1

    def add_mesh(self, pos: np.ndarray, nrm: np.ndarray, idx: np.ndarray, mat: int, vis: int = 3):
        self.pos.append(np.asarray(pos, np.float64))
        self.nrm.append(np.asarray(nrm, np.float64))
        self.idx.append(np.asarray(idx, np.int64) + self.nv)
        self.mat.append(np.full(len(idx), mat, np.uint32))
        self.vis.append(np.full(len(idx), vis, np.uint8))
        self.nv += len(pos)

    def add_emitter_geometry(self, emitters, mat: int):
        """The emitters' triangles as traced geometry seen only by closest-hit rays (flat,
        normal (v1 - v0) x (v2 - v0), the side the reference's emitters light)."""
        for tris, _ in emitters:
            tris = np.asarray(tris, np.float64).reshape(-1, 3, 3)
            for tri in tris:
                n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
                n = n / np.linalg.norm(n)
                self.add_mesh(tri, np.tile(n, (3, 1)), np.asarray([[0, 1, 2]]), mat, vis=1)

    def add_grid_quad(self, origin, eu, ev, n, normal, mat):
        """An n x n tessellated parallelogram origin + s*eu + t*ev, flat normal."""
        origin, eu, ev = (np.asarray(a, np.float64) for a in (origin, eu, ev))
        s = np.linspace(0.0, 1.0, n + 1)
        ss, tt = np.meshgrid(s, s, indexing="ij")
        pos = origin + ss.reshape(-1, 1) * eu + tt.reshape(-1, 1) * ev
        nrm = np.tile(np.asarray(normal, np.float64), (len(pos), 1))
        i = np.arange(n)
        a = (i[:, None] * (n + 1) + i[None, :]).reshape(-1)
        b, c, d = a + (n + 1), a + (n + 2), a + 1
        # winding so the geometric normal (v1-v0)x(v2-v0) matches `normal`
        tri = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
        gn = np.cross(pos[tri[0, 1]] - pos[tri[0, 0]], pos[tri[0, 2]] - pos[tri[0, 0]])
        if np.dot(gn, normal) < 0:
            tri = tri[:, [0, 2, 1]]
        self.add_mesh(pos, nrm, tri, mat)

    def add_box(self, center, half, rot, n, mat, outward=True):
        center = np.asarray(center, np.float64)
        half = np.asarray(half, np.float64)
        faces = []
        for axis in range(3):
            for sgn in (-1.0, 1.0):
                nrm = np.zeros(3)
                nrm[axis] = sgn
                u = np.zeros(3)
                v = np.zeros(3)
                u[(axis + 1) % 3] = 2 * half[(axis + 1) % 3]
                v[(axis + 2) % 3] = 2 * half[(axis + 2) % 3]
                o = -half.copy()
                o[axis] = sgn * half[axis]
                faces.append((o, u, v, nrm if outward else -nrm))
        for o, u, v, nrm in faces:
            self.add_grid_quad(center + rot @ o, rot @ u, rot @ v, n, rot @ nrm, mat)

    def build(self, emitters, name="", meta=None) -> Scene:
        pos = np.concatenate(self.pos).astype(np.float32)
        nrm = np.concatenate(self.nrm).astype(np.float32)
        idx = np.concatenate(self.idx).astype(np.uint32)
        mat = np.concatenate(self.mat).astype(np.uint32)
        vis = np.concatenate(self.vis).astype(np.uint8)
        faces = []
        offs = [0]
        cols = []
        for tris, color in emitters:
            tris = np.asarray(tris, np.float64).reshape(-1, 3, 3)
            faces.append(tris)
            offs.append(offs[-1] + len(tris))
            cols.append(color)
        return Scene(
            positions=pos, normals=nrm, indices=idx, tri_material=mat,
            materials=np.asarray(self.materials, np.float32).reshape(-1, 10),
            emitter_faces=np.concatenate(faces).astype(np.float32),
            emitter_face_offset=np.asarray(offs, np.uint32),
            emitter_color=np.asarray(cols, np.float32).reshape(-1, 3),
            name=name, meta=meta or {}, tri_mask=None if (vis == 3).all() else vis)


def _icosphere(subdiv: int):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.asarray(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache: dict = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.asarray(verts), np.asarray(faces, np.int64)


_ICO_CACHE: dict = {}


def icosphere(subdiv: int):
    if subdiv not in _ICO_CACHE:
        _ICO_CACHE[subdiv] = _icosphere(subdiv)
    return _ICO_CACHE[subdiv]


def _random_rotation(rng: Pcg32) -> np.ndarray:
    q = np.asarray([rng.normal() for _ in range(4)])
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.asarray([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def cornell_box(seed: int = 1, tess: int = 8, trace_emitters: bool = False) -> Scene:
    """Closed Cornell-style box [-1,1] x [0,2] x [-3,1]: 6 wall quads and 2 axis-aligned boxes,
    each face tessellated tess x tess (2,304 triangles at tess = 8), and one 2-triangle emitter
    of radiance 15 hanging 0.2 below the ceiling (the reference's emitter has no cosine term,
    so a light flush with the ceiling would give the ceiling unbounded 1/d^2 weights).
    Materials: albedo U[0.2, 0.8]^3 per surface, roughness 1.  trace_emitters adds the light's
    two triangles as geometry for closest-hit rays (white diffuse panel, albedo 0.8)."""
    rng = Pcg32(seed)
    b = _Builder()

    def mat():
        return b.add_material([rng.uniform(0.2, 0.8) for _ in range(3)], 1.0)

    z0, z1 = -3.0, 1.0
    d = z1 - z0
    b.add_grid_quad((-1, 0, z0), (2, 0, 0), (0, 0, d), tess, (0, 1, 0), mat())    # floor
    b.add_grid_quad((-1, 2, z0), (2, 0, 0), (0, 0, d), tess, (0, -1, 0), mat())   # ceiling
    b.add_grid_quad((-1, 0, z0), (2, 0, 0), (0, 2, 0), tess, (0, 0, 1), mat())    # back
    b.add_grid_quad((-1, 0, z1), (2, 0, 0), (0, 2, 0), tess, (0, 0, -1), mat())   # front
    b.add_grid_quad((-1, 0, z0), (0, 0, d), (0, 2, 0), tess, (1, 0, 0), mat())    # left
    b.add_grid_quad((1, 0, z0), (0, 0, d), (0, 2, 0), tess, (-1, 0, 0), mat())    # right
    eye = np.eye(3)
    b.add_box((-0.4, 0.6, -2.2), (0.3, 0.6, 0.3), eye, tess, mat())
    b.add_box((0.4, 0.3, -1.5), (0.3, 0.3, 0.3), eye, tess, mat())
    y = 1.8
    lz = -1.5
    light = [[(-0.3, y, lz - 0.3), (0.3, y, lz - 0.3), (0.3, y, lz + 0.3)],
             [(-0.3, y, lz - 0.3), (0.3, y, lz + 0.3), (-0.3, y, lz + 0.3)]]
    emitters = [(light, (15.0, 15.0, 15.0))]
    if trace_emitters:
        b.add_emitter_geometry(emitters, b.add_material([0.8, 0.8, 0.8], 1.0))
    return b.build(emitters, name=f"cornell_seed{seed}_tess{tess}" + ("_tl" if trace_emitters else ""))


def object_grid(grid: int = 16, seed: int = 2, ground_quads: int = 200, ico_subdiv: int = 4,
                ico_prob: float = 0.75, num_lights: int = 16, spacing: float = 4.0,
                trace_emitters: bool = True, lights: str = "tent") -> Scene:
    """grid x grid seeded icospheres (subdivision `ico_subdiv`) and boxes, scale U[0.3, 1.5],
    random rotation, on a ground plane of ground_quads^2 quads (half-width `half` = grid *
    spacing / 2 + spacing).  GGX roughness U[0.05, 1], albedo U[0.1, 0.9], specF0 0.04.
    `num_lights` emitter triangles of radiance lognormal(mu=2, sigma=0.5), one emitter each:
      lights="tent":   the triangles are the faces of a regular pyramid over the grid (apex at
                       height `half` above the centre, base a num_lights-gon of radius `half`
                       on the ground), facing inward, so ground and tent close the scene as the
                       reference's Cornell box is closed: a path ends by missing geometry only
                       where it leaves between the faces' edges, never into an open sky;
      lights="panels": small triangles facing down from height 4 * spacing + 4 at seeded
                       positions (an open scene: most bounces escape and end the path).
    With trace_emitters (the default) the triangles are also geometry for closest-hit rays, a
    white diffuse material (albedo 0.8, roughness 1, no emission: light reaches paths through
    NEE), the last material.  grid 16 gives about 1.06 M triangles."""
    rng = Pcg32(seed)
    b = _Builder()
    half = grid * spacing / 2.0 + spacing
    ground = b.add_material([0.5, 0.5, 0.5], 0.8, spec=(0.04, 0.04, 0.04))
    b.add_grid_quad((-half, 0, -half), (0, 0, 2 * half), (2 * half, 0, 0), ground_quads, (0, 1, 0), ground)
    iv, ifc = icosphere(ico_subdiv)
    for gx in range(grid):
        for gz in range(grid):
            scale = rng.uniform(0.3, 1.5)
            rot = _random_rotation(rng)
            m = b.add_material([rng.uniform(0.1, 0.9) for _ in range(3)], rng.uniform(0.05, 1.0),
                               spec=(0.04, 0.04, 0.04))
            cx = (gx - (grid - 1) / 2.0) * spacing
            cz = (gz - (grid - 1) / 2.0) * spacing
            center = np.asarray([cx, scale, cz])
            if rng.uniform() < ico_prob:
                pos = center + (iv @ rot.T) * scale
                nrm = iv @ rot.T
                b.add_mesh(pos, nrm, ifc, m)
            else:
                b.add_box(center, (scale, scale, scale), rot, 1, m)
    emitters = []
    if lights == "tent":
        if num_lights < 3:
            raise ValueError("a tent needs at least 3 light triangles")
        top = np.asarray([0.0, half, 0.0])
        for i in range(num_lights):
            a0, a1 = 2.0 * math.pi * i / num_lights, 2.0 * math.pi * (i + 1) / num_lights
            p0 = np.asarray([half * math.cos(a0), 0.0, half * math.sin(a0)])
            p1 = np.asarray([half * math.cos(a1), 0.0, half * math.sin(a1)])
            # wound so that (v1 - v0) x (v2 - v0) points into the tent
            tri = [top, p1, p0] if np.dot(np.cross(p0 - top, p1 - top), p0 + p1) > 0 else [top, p0, p1]
            radiance = math.exp(2.0 + 0.5 * rng.normal())
            emitters.append(([tri], (radiance, radiance, radiance)))
    elif lights == "panels":
        h = 4.0 * spacing + 4.0
        for _ in range(num_lights):
            cx = rng.uniform(-half * 0.7, half * 0.7)
            cz = rng.uniform(-half * 0.7, half * 0.7)
            s = spacing * 0.75
            radiance = math.exp(2.0 + 0.5 * rng.normal())
            tri = [(cx - s, h, cz - s), (cx + s, h, cz - s), (cx, h, cz + s)]
            emitters.append(([tri], (radiance, radiance, radiance)))
    else:
        raise ValueError(f"unknown lights {lights!r}")
    if trace_emitters:
        b.add_emitter_geometry(emitters, b.add_material([0.8, 0.8, 0.8], 1.0))
    name = f"objgrid{grid}_seed{seed}" + ("" if lights == "tent" else "_" + lights) + \
        ("" if trace_emitters else "_untraced")
    return b.build(emitters, name=name, meta={"grid": grid, "spacing": spacing, "half": half,
                                              "trace_emitters": trace_emitters, "lights": lights})


def inverse_scene(seed: int = 4) -> Scene:
    """The differentiable-rendering scene: the object grid at 8 x 8 objects (icospheres of
    subdivision 3 and boxes) on a 140 x 140-quad ground, about 100 k triangles, with 64 object
    materials plus the ground's and 16 emitter triangles (hanging panels, not traced)."""
    return object_grid(grid=8, seed=seed, ground_quads=140, ico_subdiv=3, num_lights=16, trace_emitters=False,
                       lights="panels")


def with_textures(scene: Scene, seed: int = 7, scale: float = 0.3) -> Scene:
    """The scene with seeded procedural textures, exercising every texture path of the BSDF
    (shader/brdf.cu:121-138):
      * uvs: planar from the vertex positions, u = scale (x + 0.37 y), v = scale (z - 0.21 y),
        so they run far outside [0, 1) and exercise the wrap addressing;
      * texture 0, 16 x 16 random colours (channels 20..255): diffuse of material 0 (the
        ground) and of every material m = 1 mod 4;
      * texture 1, 8 x 8 random roughness (red 26..230): roughness of materials m = 2 mod 4;
      * texture 2, 4 x 8 (height x width) OccRoughMetal (green: roughness 26..230, blue: metal
        0..255): specular of materials m = 3 mod 4, with SpecTexUsage OccRoughMetal;
      * the last material maps OccRoughMetal from its constant specular value (0, 0.5, 0.7):
        roughness 0.5, metal 0.7, no texture."""
    rng = np.random.default_rng(seed)
    p = scene.positions.astype(np.float64)
    uv = np.stack([scale * (p[:, 0] + 0.37 * p[:, 1]), scale * (p[:, 2] - 0.21 * p[:, 1])], 1).astype(np.float32)
    t0 = rng.integers(20, 256, (16, 16, 4)).astype(np.uint8)
    t1 = rng.integers(26, 231, (8, 8, 4)).astype(np.uint8)
    t2 = rng.integers(0, 256, (4, 8, 4)).astype(np.uint8)
    t2[..., 1] = rng.integers(26, 231, (4, 8))
    M = scene.materials.shape[0]
    mt = np.full((M, 4), -1, np.int32)
    mt[:, 3] = 0
    mats = np.array(scene.materials, np.float32, copy=True)
    for m in range(M):
        if m == 0 or m % 4 == 1:
            mt[m, 0] = 0
        elif m % 4 == 2:
            mt[m, 2] = 1
        else:
            mt[m, 1] = 2
            mt[m, 3] = 1
    mt[M - 1] = (-1, -1, -1, 1)
    mats[M - 1, 3:6] = (0.0, 0.5, 0.7)
    return Scene(positions=scene.positions, normals=scene.normals, indices=scene.indices,
                 tri_material=scene.tri_material, materials=mats, emitter_faces=scene.emitter_faces,
                 emitter_face_offset=scene.emitter_face_offset, emitter_color=scene.emitter_color,
                 name=scene.name + "_textured", meta=dict(scene.meta), uvs=uv, material_tex=mt,
                 textures=[t0, t1, t2], tri_mask=scene.tri_mask)


def perturbed_materials(materials: np.ndarray, seed: int, albedo: float = 0.2, roughness: float = 0.3) -> np.ndarray:
    """Seeded target parameters for inverse rendering: albedo + U[-albedo, albedo] per channel
    (clamped to [0.02, 0.98]) and roughness + U[-roughness, roughness] (clamped to [0.05, 1])."""
    rng = np.random.default_rng(seed)
    m = np.array(materials, np.float32, copy=True)
    n = m.shape[0]
    m[:, 0:3] = np.clip(m[:, 0:3] + rng.uniform(-albedo, albedo, (n, 3)), 0.02, 0.98)
    m[:, 9] = np.clip(m[:, 9] + rng.uniform(-roughness, roughness, n), 0.05, 1.0)
    return m.astype(np.float32)


@dataclass
class Camera:
    """Camera (include/graph/Camera.h:25-79, src/Camera.cpp): position, view direction, up,
    aspect and vertical field of view in radians."""
    pos: np.ndarray
    view_dir: np.ndarray
    up: np.ndarray
    aspect: float
    fov: float
    changed: bool = False

    def data(self) -> dict:
        """Camera::GetData (src/Camera.cpp:20-32) for the orthonormal view of glm::lookAt:
        U = s * tan(fov/2) * aspect, V = -u * tan(fov/2), W = f, with s = normalize(f x up),
        u = s x f.  Computed in float64 and rounded once to float32."""
        f = np.asarray(self.view_dir, np.float64)
        f = f / np.linalg.norm(f)
        s = np.cross(f, np.asarray(self.up, np.float64))
        s /= np.linalg.norm(s)
        u = np.cross(s, f)
        th = math.tan(0.5 * self.fov)
        return {
            "pos": np.asarray(self.pos, np.float32),
            "U": (s * th * self.aspect).astype(np.float32),
            "V": (-u / np.linalg.norm(u) * th).astype(np.float32),
            "W": f.astype(np.float32),
        }

    def rotate_around_origin(self, axis, angle: float):
        """Camera::RotateAroundOrigin (src/Camera.cpp:34-38)."""
        axis = np.asarray(axis, np.float64)
        axis = axis / np.linalg.norm(axis)
        c, s = math.cos(angle), math.sin(angle)
        x, y, z = axis
        rot = np.asarray([
            [c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
            [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
            [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)],
        ])
        self.pos = (rot @ np.asarray(self.pos, np.float64)).astype(np.float32)
        p = np.asarray(self.pos, np.float64)
        self.view_dir = (-p / np.linalg.norm(p)).astype(np.float32)
        self.changed = True


def cornell_camera(width: int, height: int) -> Camera:
    """Pinhole camera inside the closed box, 40 degree vertical field of view."""
    return Camera(pos=np.asarray([0.0, 1.0, 0.95], np.float32), view_dir=np.asarray([0.0, 0.0, -1.0], np.float32),
                  up=np.asarray([0.0, 1.0, 0.0], np.float32), aspect=width / height, fov=math.radians(40.0))


def orbit_camera(scene: Scene, width: int, height: int) -> Camera:
    """Camera of the orbit workloads, aimed at the origin (Camera::RotateAroundOrigin turns it):
    inside the tent of an object grid lit by one (half the tent's radius out, a quarter of its
    height up), otherwise above and beside the grid."""
    half = scene.meta.get("half", 34.0)
    if scene.meta.get("lights") == "tent":
        pos = np.asarray([0.0, 0.25 * half, 0.5 * half], np.float64)
    else:
        pos = np.asarray([0.0, 0.55 * half, 1.25 * half], np.float64)
    return Camera(pos=pos.astype(np.float32), view_dir=(-pos / np.linalg.norm(pos)).astype(np.float32),
                  up=np.asarray([0.0, 1.0, 0.0], np.float32), aspect=width / height, fov=math.radians(50.0))


def workload(name: str):
    """Named workloads: returns (scene, camera factory, renderer settings)."""
    if name == "cornell":
        return cornell_box(1, 8)
    if name == "objgrid16":
        return object_grid(16, 2)
    raise KeyError(name)
