"""Closest-hit and occlusion parity of the HIP BVH traversal against the oracle (bit-exact
primitive index, t and barycentrics), replacing optixTrace (include/graph/trace.h:25-88)."""
import numpy as np
import pytest

from crisoptix_amd import scenes
from parity_util import bits, random_rays, random_soup, shadow_rays, small_grid_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["cornell", "grid3"])
def test_closest_hit_bit_exact(name, orc, cx_runtime, cx_ctx):
    sc = scenes.cornell_box(1, 8) if name == "cornell" else small_grid_scene()
    rays = random_rays(sc, 60000, seed=11)
    oscene = orc.OracleScene(sc)
    gscene = cx_runtime.Scene(cx_ctx, sc)
    po, to, uo, vo = oscene.trace_closest(rays)
    pg, tg, ug, vg = gscene.trace_closest(rays)
    print(f"{name}: hits {np.mean(po != 0xFFFFFFFF):.3f}, prim mismatches {(po != pg).sum()}")
    assert np.array_equal(po, pg)
    assert np.array_equal(bits(to), bits(tg))
    assert np.array_equal(bits(uo), bits(ug))
    assert np.array_equal(bits(vo), bits(vg))


@pytest.mark.parametrize("name", ["cornell", "grid3"])
def test_occlusion_bit_exact(name, orc, cx_runtime, cx_ctx):
    sc = scenes.cornell_box(1, 8) if name == "cornell" else small_grid_scene()
    rays = shadow_rays(sc, 60000, seed=12)
    oscene = orc.OracleScene(sc)
    gscene = cx_runtime.Scene(cx_ctx, sc)
    oo = oscene.trace_occluded(rays)
    og = gscene.trace_occluded(rays)
    print(f"{name}: occluded fraction {oo.mean():.3f}")
    assert np.array_equal(oo, og)


def test_leaf_size_independence(orc, cx_runtime, cx_ctx):
    """Different BVHs (leaf sizes 1 and 8) give identical hits: results do not depend on the tree."""
    sc = small_grid_scene()
    rays = random_rays(sc, 20000, seed=13)
    a = cx_runtime.Scene(cx_ctx, sc, max_leaf=1).trace_closest(rays)
    b = cx_runtime.Scene(cx_ctx, sc, max_leaf=8).trace_closest(rays)
    for x, y in zip(a, b):
        assert np.array_equal(bits(x), bits(y))


@pytest.mark.parametrize("name", ["cornell", "grid3"])
def test_bvh_width_independence(name, orc, cx_runtime, cx_ctx):
    """Binary, 4-wide and 8-wide node arrays give the oracle's hits and occlusion bit for bit."""
    sc = scenes.cornell_box(1, 8) if name == "cornell" else small_grid_scene()
    rays = random_rays(sc, 30000, seed=14)
    srays = shadow_rays(sc, 30000, seed=15)
    oscene = orc.OracleScene(sc)
    ref = oscene.trace_closest(rays)
    occ = oscene.trace_occluded(srays)
    for width in (2, 4, 8):
        gscene = cx_runtime.Scene(cx_ctx, sc, width=width)
        assert gscene.info.bvhWidth == width
        got = gscene.trace_closest(rays)
        for x, y in zip(ref, got):
            assert np.array_equal(bits(x), bits(y)), width
        assert np.array_equal(occ, gscene.trace_occluded(srays)), width


def centroid_rays(scene, seed):
    """One ray per triangle, from a seeded point of the scene's box towards the triangle's
    centroid: a triangle missing from a tree (or a box that does not cover it) changes an answer."""
    rng = np.random.default_rng(seed)
    pos = scene.positions.astype(np.float64)
    lo, hi = pos.min(0), pos.max(0)
    c = pos[scene.indices].mean(axis=1)
    o = lo + rng.random(c.shape) * (hi - lo)
    d = c - o
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-12)
    rays = np.zeros((c.shape[0], 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 0.0
    rays[:, 7] = 1e16
    return rays


def _check_against_oracle(sc, gscene, oscene, rays, srays):
    ref = oscene.trace_closest(rays)
    got = gscene.trace_closest(rays)
    for x, y in zip(ref, got):
        assert np.array_equal(bits(x), bits(y))
    assert np.array_equal(oscene.trace_occluded(srays), gscene.trace_occluded(srays))


@pytest.mark.parametrize("name", ["cornell", "grid3", "grid8"])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_device_build_bit_exact(name, width, orc, cx_runtime, cx_ctx):
    """BVH built on the GPU (Morton sort, PLOC clustering, SAH leaves, wide collapse): closest-hit
    and occlusion answers equal the oracle's bit for bit, also for one ray per triangle."""
    if name == "cornell":
        sc = scenes.cornell_box(1, 8)
    elif name == "grid3":
        sc = small_grid_scene()
    else:
        sc = scenes.object_grid(grid=8, seed=9, ground_quads=40, ico_subdiv=3, num_lights=4)
    rays = np.concatenate([random_rays(sc, 30000, seed=16), centroid_rays(sc, seed=17)])
    srays = shadow_rays(sc, 30000, seed=18)
    gscene = cx_runtime.Scene(cx_ctx, sc, width=width, builder="device")
    info = gscene.info
    print(f"{name} width {width}: {sc.indices.shape[0]} triangles, {info.numNodes} binary nodes, depth "
          f"{info.maxDepth}, {info.numWideNodes} wide nodes, stack bound {info.stackBound}, {info.buildMs:.2f} ms")
    assert info.bvhBuilder == 1 and info.bvhWidth == width
    assert 0 < info.numNodes < sc.indices.shape[0]
    assert info.stackBound + 1 <= 64
    _check_against_oracle(sc, gscene, orc.OracleScene(sc), rays, srays)


def _subset(sc, k):
    import dataclasses
    return dataclasses.replace(sc, indices=sc.indices[:k].copy(), tri_material=sc.tri_material[:k].copy())


@pytest.mark.parametrize("k", [0, 1, 2, 3, 5, 17])
def test_device_build_tiny_scenes(k, orc, cx_runtime, cx_ctx):
    """0 to 17 triangles: no geometry, a single leaf holding every triangle (k <= max leaf) and
    small trees."""
    full = scenes.cornell_box(1, 8)
    sc = _subset(full, k)
    rays = random_rays(full, 3000, seed=19)   # aimed at the full box (its vertices stay in the scene)
    if k:
        rays = np.concatenate([rays, centroid_rays(sc, seed=20)])
    srays = shadow_rays(sc, 3000, seed=21)
    oscene = orc.OracleScene(sc)
    for width in (2, 4, 8):
        for max_leaf in (1, 4, 8):
            gscene = cx_runtime.Scene(cx_ctx, sc, width=width, max_leaf=max_leaf, builder="device")
            if k <= 1:
                assert gscene.info.numNodes == 0
            _check_against_oracle(sc, gscene, oscene, rays, srays)


def test_device_build_coincident_triangles(orc, cx_runtime, cx_ctx):
    """64 copies of one triangle (one Morton key, equal areas) among the Cornell box: the copies
    tie on every comparison of the clustering; the lowest primitive index must still win."""
    import dataclasses
    sc = scenes.cornell_box(1, 8)
    dup = np.repeat(sc.indices[5:6], 64, axis=0)
    sc = dataclasses.replace(sc, indices=np.concatenate([sc.indices, dup]).astype(np.uint32),
                             tri_material=np.concatenate([sc.tri_material, np.repeat(sc.tri_material[5:6], 64)])
                             .astype(np.uint32))
    rays = np.concatenate([random_rays(sc, 20000, seed=22), centroid_rays(sc, seed=23)])
    srays = shadow_rays(sc, 20000, seed=24)
    oscene = orc.OracleScene(sc)
    for width in (2, 4, 8):
        gscene = cx_runtime.Scene(cx_ctx, sc, width=width, builder="device")
        assert gscene.info.bvhBuilder == 1   # the copies merge pairwise: no chain, no fallback
        _check_against_oracle(sc, gscene, oscene, rays, srays)


def test_device_build_falls_back_when_too_deep(orc, cx_runtime, cx_ctx):
    """100 nested triangles around one centre (each contains the smaller ones): every merge
    candidate of a triangle costs its own box, so the clustering can only peel one triangle per
    iteration and the tree would be a chain deeper than the traversal stack.  The scene build
    then takes the host build (bvhBuilder reports it) and answers stay exact."""
    import dataclasses
    sc = scenes.cornell_box(1, 8)
    shape = np.array([[-1.0, -0.6, 0.0], [1.0, -0.6, 0.0], [0.0, 1.2, 0.0]])
    centre = sc.positions.astype(np.float64).mean(axis=0)
    tri = np.concatenate([centre + 0.002 * 1.05 ** k * shape for k in range(100)]).astype(np.float32)
    nv = sc.positions.shape[0]
    nrm = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (tri.shape[0], 1))
    sc = dataclasses.replace(
        sc, positions=np.concatenate([sc.positions, tri]), normals=np.concatenate([sc.normals, nrm]),
        indices=np.concatenate([sc.indices, nv + np.arange(tri.shape[0], dtype=np.uint32).reshape(-1, 3)]),
        tri_material=np.concatenate([sc.tri_material, np.zeros(100, np.uint32)]))
    if sc.uvs is not None:
        sc = dataclasses.replace(sc, uvs=np.concatenate([sc.uvs, np.zeros((tri.shape[0], 2), np.float32)]))
    rays = np.concatenate([random_rays(sc, 20000, seed=26), centroid_rays(sc, seed=27)])
    srays = shadow_rays(sc, 20000, seed=28)
    gscene = cx_runtime.Scene(cx_ctx, sc, width=4, builder="device")
    assert gscene.info.bvhBuilder == 0
    _check_against_oracle(sc, gscene, orc.OracleScene(sc), rays, srays)


def test_device_build_deterministic(cx_runtime, cx_ctx):
    """Two device builds of one scene give the same tree shape (the build orders by scans, not
    by atomics) and the same answers as the host-built tree."""
    sc = small_grid_scene()
    a = cx_runtime.Scene(cx_ctx, sc, builder="device")
    b = cx_runtime.Scene(cx_ctx, sc, builder="device")
    for f in ("numNodes", "maxDepth", "numWideNodes", "stackBound", "deviceBytes"):
        assert getattr(a.info, f) == getattr(b.info, f), f
    rays = random_rays(sc, 20000, seed=25)
    h = cx_runtime.Scene(cx_ctx, sc, builder="host").trace_closest(rays)
    for x, y, z in zip(h, a.trace_closest(rays), b.trace_closest(rays)):
        assert np.array_equal(bits(x), bits(y)) and np.array_equal(bits(y), bits(z))


@pytest.mark.parametrize("seed", range(4))
def test_device_build_random_soups(seed, orc, cx_runtime, cx_ctx):
    """Random soups with zero-area, collinear and duplicated triangles at mixed scales: the device
    build's answers equal the oracle's at every width."""
    sc = random_soup(seed)
    rays = np.concatenate([random_rays(sc, 3000, seed=60 + seed), centroid_rays(sc, seed=70 + seed)])
    srays = shadow_rays(sc, 2000, seed=80 + seed)
    oscene = orc.OracleScene(sc)
    for width in (2, 4, 8):
        gscene = cx_runtime.Scene(cx_ctx, sc, width=width, builder="device")
        _check_against_oracle(sc, gscene, oscene, rays, srays)
