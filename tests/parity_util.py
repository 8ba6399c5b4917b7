"""Shared helpers of the parity tests: seeded ray batches and comparisons."""
import numpy as np

from crisoptix_amd import scenes


def small_grid_scene():
    """A reduced object grid: 3x3 objects (icosphere subdivision 2 or boxes) on a 20x20 ground,
    under a tent of 4 traced light triangles."""
    return scenes.object_grid(grid=3, seed=5, ground_quads=20, ico_subdiv=2, num_lights=4)


def tent_grid_scene():
    """The Conditional ReSTIR workload's scene reduced: 3x3 objects on a 20x20 ground under the
    tent of 16 traced light triangles of the full scene, a closed scene in which about a third
    of the pixels hold a valid suffix."""
    return scenes.object_grid(grid=3, seed=5, ground_quads=20, ico_subdiv=2, num_lights=16)


def random_rays(scene, n, seed, kind="mixed"):
    """Rays [n][8] (origin, dir, tmin, tmax).  A third aim exactly at triangle vertices and edge
    midpoints, where a non-watertight test would leak or double-hit."""
    rng = np.random.default_rng(seed)
    pos = scene.positions.astype(np.float64)
    lo, hi = pos.min(0), pos.max(0)
    ext = hi - lo
    o = lo + rng.random((n, 3)) * ext
    o[:, 1] = lo[1] + 0.05 * ext[1] + rng.random(n) * 0.9 * ext[1]
    d = rng.normal(size=(n, 3))
    m = n // 3
    tri = scene.indices[rng.integers(0, scene.indices.shape[0], m)]
    w = rng.integers(0, 3, m)
    tgt = pos[tri[np.arange(m), w]]
    e2 = pos[tri[np.arange(m), (w + 1) % 3]]
    half = rng.random(m) < 0.5
    tgt[half] = 0.5 * (tgt[half] + e2[half])
    d[:m] = tgt - o[:m]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 1e-3
    rays[:, 7] = 1e16
    return rays


def shadow_rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    pos = scene.positions.astype(np.float64)
    lo, hi = pos.min(0), pos.max(0)
    a = lo + rng.random((n, 3)) * (hi - lo)
    b = lo + rng.random((n, 3)) * (hi - lo)
    v = b - a
    L = np.linalg.norm(v, axis=1)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = a
    rays[:, 3:6] = v / L[:, None]
    rays[:, 6] = 1e-3
    rays[:, 7] = L
    return rays


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def rel_mismatch(a, b, rtol=1e-4, atol=1e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    bad = np.abs(a - b) > rtol * np.maximum(np.abs(a), np.abs(b)) + atol
    bad |= np.isnan(a) != np.isnan(b)
    return bad


def random_soup(seed):
    """A random triangle soup with degenerate triangles mixed in: zero-area (repeated vertex),
    collinear, exact duplicates, and clusters at very different scales."""
    import dataclasses
    rng = np.random.default_rng(seed)
    base = scenes.cornell_box(1, 8)
    n = int(rng.integers(1, 600))
    c = rng.normal(size=(n, 1, 3)) * rng.choice([1e-3, 1.0, 50.0], size=(n, 1, 1))
    tri = c + rng.normal(size=(n, 3, 3)) * rng.choice([1e-4, 0.05, 2.0], size=(n, 1, 1))
    k = n // 8
    tri[:k, 2] = tri[:k, 1]                                          # zero area
    tri[k:2 * k, 2] = 2.0 * tri[k:2 * k, 1] - tri[k:2 * k, 0]        # collinear
    if n > 4 * k:
        tri[2 * k:3 * k] = tri[3 * k:4 * k]                          # duplicates
    pos = tri.reshape(-1, 3).astype(np.float32)
    return dataclasses.replace(
        base, positions=pos, normals=np.tile(np.float32([[0, 0, 1]]), (pos.shape[0], 1)),
        indices=np.arange(pos.shape[0], dtype=np.uint32).reshape(-1, 3),
        tri_material=np.zeros(n, np.uint32), uvs=None, material_tex=None, textures=[])
