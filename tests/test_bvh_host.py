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
    width 8 is kept (not the binary fallback) and its answers equal the 4-wide tree's.  (With its
    16 light panels traced as well, the greedy tree does not fit either and width 8 falls back
    to the binary tree, which test_traced_emitters_* cover.)"""
    sc = scenes.object_grid(16, 2, trace_emitters=False)
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


@pytest.mark.parametrize("width", [2, 4, 8])
def test_traced_emitters_are_invisible_to_occlusion(width, orc):
    """C9': the same scene with and without its light panels as geometry.  Occlusion rays pass
    over triangles that lack CX_VIS_OCCLUSION, so every occlusion answer is the same for both
    scenes; closest-hit rays see the panels, so a closest hit changes only where the traced
    scene's hit is a panel triangle no farther than the untraced scene's hit.  Host traversal
    and oracle agree bit for bit on the traced scene."""
    traced = scenes.object_grid(grid=3, seed=5, ground_quads=20, ico_subdiv=2, num_lights=4)
    plain = scenes.object_grid(grid=3, seed=5, ground_quads=20, ico_subdiv=2, num_lights=4, trace_emitters=False)
    nplain = plain.num_triangles
    assert traced.num_triangles == nplain + 4
    assert (traced.tri_mask[:nplain] == 3).all() and (traced.tri_mask[nplain:] == 1).all()
    # shadow rays from random ground / object points to random points of the lights
    rng = np.random.default_rng(61)
    n = 4000
    tri = plain.indices[rng.integers(0, nplain, n)]
    bary = rng.dirichlet((1, 1, 1), n)
    src = np.einsum("nk,nkd->nd", bary, plain.positions[tri].astype(np.float64))
    faces = traced.emitter_faces.astype(np.float64)
    fb = rng.dirichlet((1, 1, 1), n)
    dst = np.einsum("nk,nkd->nd", fb, faces[rng.integers(0, faces.shape[0], n)])
    v = dst - src
    dist = np.linalg.norm(v, axis=1)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = src
    rays[:, 3:6] = v / dist[:, None]
    rays[:, 6] = 1e-3
    rays[:, 7] = dist
    occ_t, _ = runtime.host_trace(traced, rays, any_hit=True, width=width)
    occ_p, _ = runtime.host_trace(plain, rays, any_hit=True, width=width)
    assert np.array_equal(occ_t, occ_p)
    assert 0.05 < occ_t.mean() < 0.95
    assert np.array_equal(orc.OracleScene(traced).trace_occluded(rays).astype(np.uint32), occ_t)
    # the same rays as closest-hit queries without the length limit: the unoccluded ones end on
    # the light they aim at
    rays[:, 7] = 1e16
    pt, tt, ut, vt, _ = runtime.host_trace(traced, rays, width=width)
    pp, tp, _, _, _ = runtime.host_trace(plain, rays, width=width)
    ref = orc.OracleScene(traced).trace_closest(rays)
    for x, y in zip(ref, (pt, tt, ut, vt)):
        assert np.array_equal(bits(x), bits(y))
    panel = pt >= nplain
    assert panel[occ_t == 0].mean() > 0.99
    same = pt == pp
    assert (same | panel).all()
    assert (tt[panel & (pp != 0xFFFFFFFF)] <= tp[panel & (pp != 0xFFFFFFFF)]).all()


@pytest.mark.parametrize("width", [2, 4, 8])
def test_overflowing_box_areas_and_non_finite_positions(width, orc):
    """Box areas that overflow to +inf leave no collapse cost below +inf (the least-area collapse
    then has no split to follow): the collapse must still produce a tree, or fall back, whose
    answers are the oracle's.  Two scenes: an object grid scaled by 3e18 (every area is +inf; the
    triangle test overflows as well, in the oracle too, so every ray misses), and the unscaled
    grid with two small triangles 1e19 away on either side (the upper levels' areas are +inf,
    rays through the grid hit).  A NaN or infinite vertex coordinate is refused."""
    import dataclasses
    base = scenes.object_grid(4, 1)
    huge = dataclasses.replace(base, positions=(base.positions.astype(np.float64) * 3e18).astype(np.float32))
    far = np.asarray([[1e19, 0, 0], [1e19, 1, 0], [1e19, 0, 1], [-1e19, 0, 0], [-1e19, 1, 0], [-1e19, 0, 1]],
                     np.float32)
    nv = len(base.positions)
    spread = dataclasses.replace(
        base, positions=np.concatenate([base.positions, far]),
        normals=np.concatenate([base.normals, np.tile([[1, 0, 0]], (6, 1)).astype(np.float32)]),
        indices=np.concatenate([base.indices, np.asarray([[nv, nv + 1, nv + 2], [nv + 3, nv + 4, nv + 5]],
                                                         np.uint32)]),
        tri_material=np.concatenate([base.tri_material, base.tri_material[:2]]),
        tri_mask=None if base.tri_mask is None else np.concatenate([base.tri_mask, [3, 3]]).astype(np.uint8))
    for sc, rays, hit_frac in ((huge, random_rays(huge, 2000, seed=33), 0.0),
                               (spread, random_rays(base, 2000, seed=34), 0.3)):
        rays[:, 7] = 1e30
        po, to, uo, vo = orc.OracleScene(sc).trace_closest(rays)
        pg, tg, ug, vg, used = runtime.host_trace(sc, rays, width=width)
        assert used in (2, width)
        assert np.mean(po != 0xFFFFFFFF) >= hit_frac
        assert np.array_equal(po, pg)
        for a, b in ((to, tg), (uo, ug), (vo, vg)):
            assert np.array_equal(bits(a), bits(b))
    for bad in (np.nan, np.inf, -np.inf):
        pos = base.positions.copy()
        pos[5, 1] = bad
        with pytest.raises(runtime.CxError):
            runtime.host_trace(dataclasses.replace(base, positions=pos), random_rays(base, 10, seed=1), width=width)
