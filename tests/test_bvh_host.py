"""The wide (4- and 8-child) BVH collapse and its traversal, checked on the CPU: the device
traversal code compiled for the host (cx_bvh_host_trace) over binary, 4-wide and 8-wide node
arrays gives the oracle's closest hits (primitive, t, barycentrics) and occlusion bit for bit.
Results are a function of the ray and the triangle set alone, so every BVH shape must agree."""
import numpy as np
import pytest

from crisoptix_amd import runtime, scenes
from parity_util import bits, random_rays, random_soup, shadow_rays, small_grid_scene

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(runtime.LIB_PATH), reason="HIP library not built")


@pytest.mark.parametrize("name", ["cornell", "grid3"])
@pytest.mark.parametrize("width,max_leaf", [(2, 4), (4, 4), (8, 4), (4, 1), (8, 8)])
def test_host_traversal_matches_oracle(name, width, max_leaf, orc):
    sc = scenes.cornell_box(1, 8) if name == "cornell" else small_grid_scene()
    rays = random_rays(sc, 6000, seed=21)
    po, to, uo, vo = orc.OracleScene(sc).trace_closest(rays)
    pg, tg, ug, vg, used = runtime.host_trace(sc, rays, width=width, max_leaf=max_leaf)
    assert used == width
    assert np.array_equal(po, pg)
    assert np.array_equal(bits(to), bits(tg))
    assert np.array_equal(bits(uo), bits(ug))
    assert np.array_equal(bits(vo), bits(vg))
    srays = shadow_rays(sc, 6000, seed=22)
    oo = orc.OracleScene(sc).trace_occluded(srays)
    og, _ = runtime.host_trace(sc, srays, any_hit=True, width=width, max_leaf=max_leaf)
    assert np.array_equal(oo.astype(np.uint32), og)


def test_wide8_collapse_falls_back_to_greedy_on_large_scene():
    """The 1.06 M-triangle object grid of the 1080p benchmark: the least-area collapse at width 8
    needs more than 64 traversal-stack entries, so the collapse takes the greedy rule, which fits;
    width 8 is kept (not the binary fallback) and its answers equal the 4-wide tree's."""
    sc = scenes.object_grid(16, 2)
    rays = random_rays(sc, 3000, seed=31)
    *a8, used8 = runtime.host_trace(sc, rays, width=8)
    *a4, used4 = runtime.host_trace(sc, rays, width=4)
    assert (used8, used4) == (8, 4)
    assert np.mean(a4[0] != 0xFFFFFFFF) > 0.3
    for x, y in zip(a8, a4):
        assert np.array_equal(bits(x), bits(y))


def test_empty_geometry_misses():
    """A scene with no triangles (explicit emitters only): every query misses, at every width."""
    import dataclasses
    sc = scenes.cornell_box(1, 8)
    rays = random_rays(sc, 200, seed=32)
    sc = dataclasses.replace(sc, indices=np.zeros((0, 3), np.uint32), tri_material=np.zeros(0, np.uint32))
    for width in (2, 4, 8):
        prim, t, u, v, _ = runtime.host_trace(sc, rays, width=width)
        assert (prim == 0xFFFFFFFF).all() and (t == 0).all()
        occ, _ = runtime.host_trace(sc, rays, any_hit=True, width=width)
        assert not occ.any()


@pytest.mark.parametrize("seed", range(6))
def test_random_soups_match_oracle(seed, orc):
    """Random soups with degenerate triangles: every width and leaf size gives the oracle's
    closest hits and occlusion bit for bit (the builds and both collapses on odd trees)."""
    sc = random_soup(seed)
    rays = random_rays(sc, 3000, seed=40 + seed)
    srays = shadow_rays(sc, 2000, seed=50 + seed)
    oscene = orc.OracleScene(sc)
    ref = oscene.trace_closest(rays)
    assert np.mean(ref[0] != 0xFFFFFFFF) > 0.1   # about a quarter of the rays hit
    occ = oscene.trace_occluded(srays).astype(np.uint32)
    for width in (2, 4, 8):
        for max_leaf in (1, 2, 4, 15):
            *got, _ = runtime.host_trace(sc, rays, width=width, max_leaf=max_leaf)
            for x, y in zip(ref, got):
                assert np.array_equal(bits(x), bits(y)), (width, max_leaf)
            og, _ = runtime.host_trace(sc, srays, any_hit=True, width=width, max_leaf=max_leaf)
            assert np.array_equal(occ, og), (width, max_leaf)
