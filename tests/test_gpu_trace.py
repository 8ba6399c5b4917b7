"""Closest-hit and occlusion parity of the HIP BVH traversal against the oracle (bit-exact
primitive index, t and barycentrics), replacing optixTrace (include/graph/trace.h:25-88)."""
import numpy as np
import pytest

from crisoptix_amd import scenes
from parity_util import bits, random_rays, shadow_rays, small_grid_scene

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
