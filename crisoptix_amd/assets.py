"""Wavefront OBJ / MTL scenes, the model format the reference loads through assimp
(src/Model.cpp:9-374, e.g. data/model/Cornell/CornellBox-Original.obj in src/main.cpp:176).

Geometry: `v`, `vt`, `vn`, `f` (polygons fan-triangulated like aiProcess_Triangulate;
`v/vt/vn`, `v//vn`, `v/vt` and negative, relative indices), `usemtl`, `mtllib`.  Faces without
normals get their flat face normal (aiProcess_GenNormals); corners without `vt` get uv (0, 0)
(src/Model.cpp:79-85); `flip_uv` is aiProcess_FlipUVs (v -> 1 - v).

Materials (Model::LoadMaterials, src/Model.cpp:209-374): diffuse `Kd`, specular `Ks` (the
specF0 slot), emissive `Ke`, roughness `Pr` (AI_MATKEY_ROUGHNESS_FACTOR; 1 when absent, as the
reference's default).  A material the file names but does not define, or faces before any
`usemtl`, get assimp's OBJ default material (diffuse 0.6).  Texture maps: `map_Kd` (diffuse),
`map_Ks` (specular) and `map_Pr` (roughness) are bound to the material, each file loaded once
per path (Model::m_Textures) as RGBA8 with `.dds` read as the `.png` of the same name
(src/Texture.cpp:9-21); the metalness map is loaded by the reference but never bound, so it is
not read.  `spec_tex_usage` is the model's SpecTexUsage (Color, or OccRoughMetal for glTF-style
occlusion/roughness/metal maps).  Image files are read by image.read_png (8-bit PNG).

Emitters (Scene::BuildEmitterTable, src/Scene.cpp:176-204): every group of faces whose material
has |Ke| > 0.01 becomes an implicit emitter over those faces with colour (1, 1, 1) (the reference
forces implicit emitters to unit colour, Scene.cpp:194); `emitter_color` overrides that.  The
reference also traces those faces as geometry, and so does the loader by default: closest-hit
rays see them and occlusion rays pass over them (tri_mask CX_VIS_SURFACE), so a shadow ray aimed
at a point of a light is not reported occluded by that light whenever t rounds below the
distance.  `trace_emitters=False` keeps them out of the traced triangles.
"""
from __future__ import annotations

import os

import numpy as np

from .image import read_png
from .scenes import Scene

DEFAULT_MATERIAL = {"Kd": (0.6, 0.6, 0.6), "Ks": (0.0, 0.0, 0.0), "Ke": (0.0, 0.0, 0.0), "Pr": 1.0}
TEXTURE_KEYS = ("map_Kd", "map_Ks", "map_Pr")   # diffuse, specular, roughness slots


def parse_mtl(text: str) -> dict:
    mats: dict = {}
    cur = None
    for line in text.splitlines():
        t = line.split("#", 1)[0].split()
        if not t:
            continue
        key, args = t[0], t[1:]
        if key == "newmtl":
            cur = " ".join(args)
            mats[cur] = dict(DEFAULT_MATERIAL)
        elif cur is None:
            continue
        elif key in ("Kd", "Ks", "Ke") and len(args) >= 3:
            mats[cur][key] = tuple(float(x) for x in args[:3])
        elif key == "Pr" and args:
            mats[cur]["Pr"] = float(args[0])
        elif key in TEXTURE_KEYS and args:
            mats[cur][key] = args[-1]     # options (-s, -o, ...) precede the file name
    return mats


def parse_obj(text: str, mtl_loader=None):
    """Returns (positions [V,3], normals [V,3], uvs [V,2], triangles [T,3], material name per
    triangle, material table).  Vertices are emitted per face corner (no sharing), so flat
    normals and the file's per-corner normals and uvs are all kept."""
    vs, vns, vts = [], [], []
    pos, nrm, tex, tri_mat = [], [], [], []
    mats: dict = {}
    cur = None

    def idx(tok: str, n: int) -> int:
        i = int(tok)
        return i - 1 if i > 0 else n + i

    for line in text.splitlines():
        t = line.split("#", 1)[0].split()
        if not t:
            continue
        key, args = t[0], t[1:]
        if key == "v":
            vs.append([float(x) for x in args[:3]])
        elif key == "vn":
            vns.append([float(x) for x in args[:3]])
        elif key == "vt":
            vts.append([float(args[0]), float(args[1]) if len(args) > 1 else 0.0])
        elif key == "mtllib" and mtl_loader is not None:
            for name in args:
                mats.update(parse_mtl(mtl_loader(name)))
        elif key == "usemtl":
            cur = " ".join(args)
        elif key == "f":
            corners = []
            for c in args:
                parts = c.split("/")
                vi = idx(parts[0], len(vs))
                ti = idx(parts[1], len(vts)) if len(parts) > 1 and parts[1] else None
                ni = idx(parts[2], len(vns)) if len(parts) > 2 and parts[2] else None
                corners.append((vi, ni, ti))
            for k in range(1, len(corners) - 1):   # fan triangulation
                face = [corners[0], corners[k], corners[k + 1]]
                p = np.asarray([vs[c[0]] for c in face], np.float64)
                if any(c[1] is None for c in face):
                    n = np.cross(p[1] - p[0], p[2] - p[0])
                    ln = np.linalg.norm(n)
                    n = n / ln if ln > 0 else np.asarray([0.0, 1.0, 0.0])
                    ns = [n, n, n]
                else:
                    ns = [np.asarray(vns[c[1]], np.float64) for c in face]
                pos.extend(p)
                nrm.extend(ns)
                tex.extend([vts[c[2]] if c[2] is not None else [0.0, 0.0] for c in face])
                tri_mat.append(cur)
    P = np.asarray(pos, np.float32).reshape(-1, 3)
    N = np.asarray(nrm, np.float32).reshape(-1, 3)
    UV = np.asarray(tex, np.float32).reshape(-1, 2)
    T = np.arange(len(P), dtype=np.uint32).reshape(-1, 3)
    return P, N, UV, T, tri_mat, mats


def load_texture(path: str) -> np.ndarray:
    """A texture file as [H, W, 4] RGBA8 (Texture::Texture, src/Texture.cpp:7-27): a `.dds` name
    is read as the `.png` of the same name."""
    i = path.find(".dds")
    if i >= 0:
        path = path[:i] + ".png" + path[i + 4:]
    return read_png(path, rgba=True)


def load_obj(path: str, trace_emitters: bool = True, emitter_color=(1.0, 1.0, 1.0), flip_uv: bool = False,
             spec_tex_usage: int = 0) -> Scene:
    """An OBJ file (with its MTL libraries and texture files next to it) as a flattened scene
    (one instance); the arguments after the path are those of Model::Model
    (src/Model.cpp:9-40) plus the emitter options of scene_from_obj."""
    base = os.path.dirname(os.path.abspath(path))

    def loader(name):
        with open(os.path.join(base, name)) as f:
            return f.read()

    with open(path) as f:
        text = f.read()
    return scene_from_obj(text, loader, trace_emitters, emitter_color, name=os.path.basename(path),
                          texture_loader=lambda name: load_texture(os.path.join(base, name)), flip_uv=flip_uv,
                          spec_tex_usage=spec_tex_usage)


def scene_from_obj(text: str, mtl_loader=None, trace_emitters: bool = True, emitter_color=(1.0, 1.0, 1.0),
                   name: str = "obj", texture_loader=None, flip_uv: bool = False, spec_tex_usage: int = 0) -> Scene:
    """`texture_loader(name)` returns the RGBA8 image of a texture file named by the MTL (None:
    texture maps are not bound)."""
    P, N, UV, T, tri_mat, mats = parse_obj(text, mtl_loader)
    if flip_uv:
        UV = UV.copy()
        UV[:, 1] = np.float32(1.0) - UV[:, 1]
    names = []
    for m in tri_mat:
        key = m if m in mats else None
        if key not in names:
            names.append(key)
    table, mtex = [], []
    textures, tex_index = [], {}
    for key in names:
        m = mats.get(key, DEFAULT_MATERIAL) if key is not None else DEFAULT_MATERIAL
        table.append([*m["Kd"], *m["Ks"], *m["Ke"], m["Pr"]])
        row = [-1, -1, -1, int(spec_tex_usage)]
        for slot, tk in enumerate(TEXTURE_KEYS):
            fname = m.get(tk)
            if fname and texture_loader is not None:
                if fname not in tex_index:          # one texture per file (Model::m_Textures)
                    tex_index[fname] = len(textures)
                    textures.append(np.ascontiguousarray(texture_loader(fname), np.uint8))
                row[slot] = tex_index[fname]
        mtex.append(row)
    mat_of = np.asarray([names.index(m if m in mats else None) for m in tri_mat], np.uint32)
    materials = np.asarray(table, np.float32).reshape(-1, 10)
    emissive = np.linalg.norm(materials[:, 6:9], axis=1) > 0.01
    faces, offs, cols = [], [0], []
    for mi in np.flatnonzero(emissive):
        tris = T[mat_of == mi]
        faces.append(P[tris])
        offs.append(offs[-1] + len(tris))
        cols.append(emitter_color)
    keep = np.ones(len(T), bool) if trace_emitters else ~emissive[mat_of]
    if not faces:
        raise ValueError("the scene has no emissive material (the frame loop needs an emitter)")
    T2 = T[keep]
    used = np.unique(T2.reshape(-1))
    remap = np.full(len(P), 0, np.uint32)
    remap[used] = np.arange(len(used), dtype=np.uint32)
    mt = np.asarray(mtex, np.int32).reshape(-1, 4)
    any_tex = bool(textures) or bool(np.any(mt[:, 3] != 0))
    # traced emissive faces are seen by closest-hit rays only (CX_VIS_SURFACE)
    vis = np.where(emissive[mat_of[keep]], 1, 3).astype(np.uint8)
    return Scene(
        positions=P[used], normals=N[used], indices=remap[T2], tri_material=mat_of[keep],
        materials=materials, emitter_faces=np.concatenate(faces).astype(np.float32),
        emitter_face_offset=np.asarray(offs, np.uint32),
        emitter_color=np.asarray(cols, np.float32).reshape(-1, 3), name=name,
        meta={"half": float(np.abs(P).max()) if len(P) else 1.0},
        uvs=UV[used] if any_tex else None, material_tex=mt if any_tex else None, textures=textures,
        tri_mask=vis if (vis != 3).any() else None)


def cornell_box_obj(textured: bool = False):
    """An OBJ + MTL pair of a Cornell box in the layout of the classic one (white floor,
    ceiling and back wall, red left wall, green right wall, two white blocks, a ceiling light
    quad), written here from its published room and block dimensions: the test asset for the
    loader, which the reference's own CornellBox-Original.obj is not (it is not in its tree).

    textured: every quad gets uvs; the floor (material `floor`, `map_Kd checker.dds`, read as
    checker.png) repeats an 8 x 8 checker 3 times per side; the blocks (material `block`,
    `map_Pr rough.png`) read their roughness from a 4 x 4 pattern.  Returns (obj, mtl,
    {file name: RGBA8 image}) then."""
    mtl = """# Cornell box materials
newmtl white
Kd 0.725 0.71 0.68
Ks 0 0 0
newmtl red
Kd 0.63 0.065 0.05
newmtl green
Kd 0.14 0.45 0.091
newmtl light
Kd 0.78 0.78 0.78
Ke 17 12 4
Pr 1.0
"""
    if textured:
        mtl += """newmtl floor
Kd 1 1 1
map_Kd checker.dds
newmtl block
Kd 0.725 0.71 0.68
Pr 0.5
map_Pr -bm 1.0 rough.png
"""
    v = []
    f = []

    def quad(a, b, c, d, mat, uvscale=1.0):
        base = len(v)
        v.extend([a, b, c, d])
        f.append((mat, [base + 1, base + 2, base + 3, base + 4], uvscale))

    S = 5.56
    quad((S, 0, 0), (0, 0, 0), (0, 0, S), (S, 0, S), "floor" if textured else "white", 3.0)   # floor (+y)
    quad((S, S, 0), (S, S, S), (0, S, S), (0, S, 0), "white")          # ceiling (normal -y)
    quad((S, 0, S), (0, 0, S), (0, S, S), (S, S, S), "white")          # back wall (normal -z)
    quad((S, 0, 0), (S, 0, S), (S, S, S), (S, S, 0), "red")            # left wall (normal -x)
    quad((0, 0, S), (0, 0, 0), (0, S, 0), (0, S, S), "green")          # right wall (normal +x)
    quad((3.43, S - 0.001, 2.27), (3.43, S - 0.001, 3.32), (2.13, S - 0.001, 3.32), (2.13, S - 0.001, 2.27),
         "light")

    def block(x0, z0, w, d, h, ang):
        c, s_ = np.cos(ang), np.sin(ang)
        cx, cz = x0 + w / 2, z0 + d / 2
        pts = []
        for (dx, dz) in ((-w / 2, -d / 2), (w / 2, -d / 2), (w / 2, d / 2), (-w / 2, d / 2)):
            pts.append((cx + c * dx - s_ * dz, cz + s_ * dx + c * dz))
        lo = [(p[0], 0.0, p[1]) for p in pts]
        hi = [(p[0], h, p[1]) for p in pts]
        mat = "block" if textured else "white"
        quad(hi[0], hi[3], hi[2], hi[1], mat)
        for i in range(4):
            j = (i + 1) % 4
            quad(lo[i], hi[i], hi[j], lo[j], mat)

    block(1.3, 0.65, 1.6, 1.6, 1.65, -0.29)
    block(2.65, 2.96, 1.6, 1.6, 3.3, 0.29)
    lines = ["# Cornell box", "mtllib cornell.mtl"]
    lines += [f"v {a:.4f} {b:.4f} {c:.4f}" for a, b, c in v]
    if textured:
        for k in (1.0, 3.0):
            lines += [f"vt 0 0", f"vt {k:g} 0", f"vt {k:g} {k:g}", f"vt 0 {k:g}"]
    cur = None
    for mat, idx, uvscale in f:
        if mat != cur:
            lines.append(f"usemtl {mat}")
            cur = mat
        if textured:
            t0 = 1 if uvscale == 1.0 else 5
            lines.append("f " + " ".join(f"{i}/{t0 + k}" for k, i in enumerate(idx)))
        else:
            lines.append("f " + " ".join(str(i) for i in idx))
    obj = "\n".join(lines) + "\n"
    if not textured:
        return obj, mtl
    yy, xx = np.mgrid[0:8, 0:8]
    checker = np.zeros((8, 8, 4), np.uint8)
    checker[..., 3] = 255
    dark = ((xx + yy) % 2) == 0
    checker[dark] = (30, 50, 200, 255)
    checker[~dark] = (120, 140, 250, 255)
    rough = np.full((4, 4, 4), 255, np.uint8)
    rough[..., 0] = (np.arange(16).reshape(4, 4) * 14 + 25).astype(np.uint8)
    return obj, mtl, {"checker.png": checker, "rough.png": rough}
