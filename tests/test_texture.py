"""The oracle's texture sampler (tex2D<float4> on the RGBA8 textures of src/Texture.cpp:44-76:
wrap addressing, linear filtering, normalized coordinates, normalized-float reads) pinned by
what the filter's definition fixes: constants, texel centres, midpoints, wrap periodicity,
exact reproduction of bilinear ramps (which a transposed or shifted index breaks), the 1/256
weight grid, and the textured BSDF's channel mapping (shader/brdf.cu:121-138)."""
import numpy as np
import pytest

from crisoptix_amd import abi, scenes


def rand_tex(h, w, seed=0):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 4)).astype(np.uint8)


def test_constant_texture_reads_its_value(orc):
    for c in (0, 1, 77, 128, 254, 255):
        t = np.full((5, 3, 4), c, np.uint8)
        uv = np.random.default_rng(c).uniform(-7, 7, (500, 2))
        out = orc.texture_sample(t, uv)
        assert np.all(out == np.float32(c) / np.float32(255)), c


def test_texel_centres_read_the_texel(orc):
    # power-of-two sizes: (i + 0.5) / W * W - 0.5 = i exactly; H != W catches a transposition
    t = rand_tex(4, 8, 1)
    for j in range(4):
        for i in range(8):
            out = orc.texture_sample(t, [[(i + 0.5) / 8, (j + 0.5) / 4]])[0]
            assert np.array_equal(out, t[j, i, :3].astype(np.float32) / np.float32(255)), (i, j)


def test_midpoints_average_neighbours_and_wrap(orc):
    t = rand_tex(4, 8, 2)
    for i in range(8):
        # u = i / 8: x = i - 0.5, halfway between texels i - 1 and i (texel -1 wraps to 7)
        out = orc.texture_sample(t, [[i / 8, 0.5 / 4]])[0]
        want = (t[0, (i - 1) % 8, :3].astype(np.float64) + t[0, i, :3]) / 510.0
        assert np.allclose(out, want, rtol=1e-7, atol=0), i
    for j in range(4):
        out = orc.texture_sample(t, [[0.5 / 8, j / 4]])[0]
        want = (t[(j - 1) % 4, 0, :3].astype(np.float64) + t[j, 0, :3]) / 510.0
        assert np.allclose(out, want, rtol=1e-7, atol=0), j


def test_wrap_is_periodic(orc):
    t = rand_tex(6, 10, 3)
    # multiples of 1/4096, so a shift by whole periods keeps the fraction exactly
    uv = (np.random.default_rng(4).integers(0, 4096, (300, 2)) / 4096.0).astype(np.float32)
    base = orc.texture_sample(t, uv)
    for du, dv in [(1, 0), (0, 1), (-3, 2), (5, -4), (-1, -1)]:
        out = orc.texture_sample(t, (uv.astype(np.float64) + [du, dv]).astype(np.float32))
        assert np.array_equal(out, base), (du, dv)


def test_bilinear_ramp_is_reproduced(orc):
    # texel (i, j) = 3 i + 11 j + 5 on an 8 x 16 (h x w) image: bilinear filtering returns
    # 3 x + 11 y + 5 at the quantised position, away from the wrap seam
    h, w = 8, 16
    jj, ii = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    t = np.zeros((h, w, 4), np.uint8)
    t[..., 0] = 3 * ii + 11 * jj + 5
    t[..., 1] = 11 * ii + 3 * jj + 5
    t[..., 2] = 100
    rng = np.random.default_rng(5)
    x = rng.integers(0, (w - 1) * 256, 400) / 256.0      # positions on the 1/256 grid
    y = rng.integers(0, (h - 1) * 256, 400) / 256.0
    uv = np.stack([(x + 0.5) / w, (y + 0.5) / h], 1).astype(np.float32)
    out = orc.texture_sample(t, uv).astype(np.float64) * 255
    assert np.allclose(out[:, 0], 3 * x + 11 * y + 5, atol=2e-4)
    assert np.allclose(out[:, 1], 11 * x + 3 * y + 5, atol=2e-4)
    assert np.allclose(out[:, 2], 100, atol=2e-5)


def test_weights_are_held_to_eight_fractional_bits(orc):
    t = np.zeros((1, 2, 4), np.uint8)
    t[0, 1, 0] = 255           # red ramps from 0 at texel 0 to 1 at texel 1
    for frac, want in [(0.3 / 256, 0.0), (0.7 / 256, 1 / 256), (0.5, 0.5), (255.4 / 256, 255 / 256),
                       (255.6 / 256, 1.0)]:
        u = (0.5 + frac) / 2
        out = orc.texture_sample(t, [[u, 0.5]])[0]
        assert out[0] == np.float32(want), (frac, out[0])


def test_non_finite_coordinates_read_as_zero(orc):
    t = rand_tex(3, 5, 6)
    zero = orc.texture_sample(t, [[0.0, 0.0]])[0]
    for u, v in [(np.nan, 0.0), (0.0, np.inf), (-np.inf, np.nan)]:
        assert np.array_equal(orc.texture_sample(t, [[u, v]])[0], zero)


def _bsdf_probe(orc, sc, prim, out_dir):
    os_ = orc.OracleScene(sc)
    inter = np.zeros((1, 8), np.float32)
    inter[0, 0] = np.array([prim], np.uint32).view(np.float32)[0]
    inter[0, 1:3] = (0.25, 0.25)
    i0, i1, i2 = sc.indices[prim]
    n = sc.normals[i0] + sc.normals[i1] + sc.normals[i2]
    n = n / np.linalg.norm(n)
    inter[0, 3:6] = -n       # looking straight down the normal
    return os_.ggx_eval(inter, np.asarray([out_dir], np.float32))[0], n


def test_textured_bsdf_channel_mapping(orc):
    """A textured material evaluates like the untextured material whose constants are the
    texture's values at the hit (diffuse colour; roughness from red; OccRoughMetal: roughness
    from green, metal from blue, specF0 = mix(diffuse, 1, metal))."""
    base = scenes.object_grid(grid=2, seed=3, ground_quads=4, ico_subdiv=1, num_lights=1, trace_emitters=False,
                              lights="panels")
    tex = scenes.with_textures(base)
    M = tex.materials.shape[0]
    rng = np.random.default_rng(7)
    for m in range(M):
        prim = int(np.flatnonzero(tex.tri_material == m)[0])
        i = tex.indices[prim]
        uv = (0.5 * tex.uvs[i[0]] + 0.25 * tex.uvs[i[1]]) + 0.25 * tex.uvs[i[2]]
        mt = tex.material_tex[m]
        const = np.array(tex.materials, np.float32, copy=True)
        spec = const[m, 3:6].copy()
        if mt[0] >= 0:
            const[m, 0:3] = orc.texture_sample(tex.textures[mt[0]], [uv])[0]
        if mt[1] >= 0:
            spec = orc.texture_sample(tex.textures[mt[1]], [uv])[0]
        if mt[2] >= 0:
            const[m, 9] = orc.texture_sample(tex.textures[mt[2]], [uv])[0][0]
        metal = 0.0
        if mt[3] == abi.CX_SPEC_TEX_OCC_ROUGH_METAL:
            const[m, 9] = spec[1]
            metal = float(spec[2])
        plain = scenes.Scene(**{k: getattr(tex, k) for k in (
            "positions", "normals", "indices", "tri_material", "emitter_faces", "emitter_face_offset",
            "emitter_color")}, materials=const)
        d = rng.normal(size=3)
        _, n = _bsdf_probe(orc, tex, prim, [0, 0, 1])
        d = n + 0.6 * d / np.linalg.norm(d)
        d = d / np.linalg.norm(d)
        got, _ = _bsdf_probe(orc, tex, prim, d)
        ref, _ = _bsdf_probe(orc, plain, prim, d)
        if metal == 0.0:
            assert np.array_equal(got, ref), m
        else:
            # the untextured material has no metal: compare against the closed form instead
            assert np.array_equal(got[7], np.float32(const[m, 9])), m
            assert got[3] > 0


def _quad_scene(mat, mat_tex=None, textures=()):
    P = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    N = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    em = np.array([[[-0.2, 2, -0.2], [0.2, 2, -0.2], [0.2, 2, 0.2]]], np.float32)
    return scenes.Scene(positions=P, normals=N, indices=np.array([[0, 2, 1], [0, 3, 2]], np.uint32),
                        tri_material=np.zeros(2, np.uint32), materials=np.asarray([mat], np.float32),
                        emitter_faces=em, emitter_face_offset=np.array([0, 1], np.uint32),
                        emitter_color=np.ones((1, 3), np.float32), uvs=uv,
                        material_tex=None if mat_tex is None else np.asarray([mat_tex], np.int32),
                        textures=list(textures))


def test_occ_rough_metal_closed_form_at_normal_incidence(orc):
    """At l = v = n: D = 1/(pi r^2), V = 1/4, Fresnel = F0, so
    brdf = diff/pi + F0/(4 pi r^2) with F0 = diff (1 - metal) + metal, r = specular green."""
    diff = np.array([0.3, 0.6, 0.9], np.float32)
    g, b = np.float32(0.5), np.float32(0.7)
    mat = [*diff, 0.0, g, b, 0, 0, 0, 0.9]
    for sc in (_quad_scene(mat, (-1, -1, -1, 1)),
               # the same values from a constant specular texture (green 128, blue 179)
               _quad_scene([*diff, 0, 0, 0, 0, 0, 0, 0.9], (-1, 0, -1, 1),
                           [np.tile(np.array([[[9, 128, 179, 255]]], np.uint8), (2, 2, 1))])):
        os_ = orc.OracleScene(sc)
        inter = np.zeros((1, 8), np.float32)
        inter[0, 0] = np.array([1], np.uint32).view(np.float32)[0]
        inter[0, 1:3] = (0.2, 0.3)
        inter[0, 3:6] = (0, -1, 0)
        out = os_.ggx_eval(inter, np.array([[0, 1, 0]], np.float32))[0]
        tex = sc.material_tex[0, 1] >= 0
        r = 128 / 255 if tex else float(g)
        m = 179 / 255 if tex else float(b)
        f0 = diff.astype(np.float64) * (1 - m) + m
        want = diff / np.pi + f0 / (4 * np.pi * r * r)
        assert np.allclose(out[:3], want, rtol=2e-6), (out[:3], want)
        assert np.isclose(out[7], r, rtol=1e-7)
