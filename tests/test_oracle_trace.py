"""The oracle's traversal: its BVH gives the same closest hits and occlusion answers as a
brute-force loop over every triangle (the C5 rule makes the result tree-independent), and the
watertight test does not leak through shared edges of a closed mesh."""
import numpy as np

from crisoptix_amd import scenes
from parity_util import bits, random_rays, shadow_rays, small_grid_scene


def test_bvh_equals_brute_force_closest(orc):
    for sc in (scenes.cornell_box(1, 4), small_grid_scene()):
        rays = random_rays(sc, 6000, seed=21)
        a = orc.OracleScene(sc).trace_closest(rays)
        b = orc.OracleScene(sc, brute_force=True).trace_closest(rays)
        for x, y in zip(a, b):
            assert np.array_equal(bits(x), bits(y))


def test_bvh_equals_brute_force_occlusion(orc):
    sc = small_grid_scene()
    rays = shadow_rays(sc, 6000, seed=22)
    assert np.array_equal(orc.OracleScene(sc).trace_occluded(rays),
                          orc.OracleScene(sc, brute_force=True).trace_occluded(rays))


def test_watertight_closed_sphere(orc):
    b = scenes._Builder()
    m = b.add_material([0.5, 0.5, 0.5], 0.5)
    v, f = scenes.icosphere(3)
    b.add_mesh(v * 2.0, v, f, m)
    sc = b.build([([[(0, 9, 0), (1, 9, 0), (0, 9, 1)]], (1.0, 1.0, 1.0))])
    rng = np.random.default_rng(23)
    n = 20000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # aim a third of the rays exactly at vertices (shared by 5-6 triangles)
    d[: n // 3] = v[rng.integers(0, len(v), n // 3)]
    rays = np.zeros((n, 8), np.float32)
    rays[:, 3:6] = d
    rays[:, 6] = 0.0
    rays[:, 7] = 1e16
    prim, t, u, w = orc.OracleScene(sc).trace_closest(rays)
    assert (prim != 0xFFFFFFFF).all()
    assert np.all(np.abs(t - 2.0) < 0.02)
    assert (u >= 0).all() and (w >= 0).all() and ((u + w) <= 1.0 + 1e-6).all()


def test_interval_is_inclusive(orc):
    """A hit exactly at tmax counts (optixTrace's [tmin, tmax]); one past it does not."""
    sc = scenes.cornell_box(1, 2)
    oscene = orc.OracleScene(sc)
    rays = np.asarray([[0, 1, -1, 0, 0, -1, 1e-3, 1e16]], np.float32)
    prim, t, _, _ = oscene.trace_closest(rays)
    assert prim[0] != 0xFFFFFFFF
    hit_at = t[0]
    r2 = rays.copy()
    r2[0, 7] = hit_at
    assert oscene.trace_occluded(r2)[0] == 1
    r2[0, 7] = np.nextafter(hit_at, np.float32(0))
    assert oscene.trace_occluded(r2)[0] == 0
