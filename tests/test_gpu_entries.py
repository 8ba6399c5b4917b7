"""Neighbour-search parity: the device entry tree (Morton sort + implicit tree) against the
oracle's brute-force O(n^2) search under the (distance, index) rule, bit-exact indices and
distances.  Stands in for PrefixAccelStruct + __intersection__prefix_entry
(src/PrefixAccelStruct.cpp:66-221, shader/prefix_search_hit.cu:9-73)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_entries(n, seed):
    rng = np.random.default_rng(seed)
    centers = rng.random((16, 3)) * 10
    pos = centers[rng.integers(0, 16, n)] + rng.normal(scale=0.6, size=(n, 3))
    radius = rng.uniform(0.05, 0.5, n)
    valid = (rng.random(n) < 0.8).astype(np.uint8)
    q = pos[rng.integers(0, n, n // 2)] + rng.normal(scale=0.2, size=(n // 2, 3))
    return pos.astype(np.float32), radius.astype(np.float32), valid, q.astype(np.float32)


@pytest.mark.parametrize("k", [1, 8, 31])
def test_neighbor_sets_bit_exact(k, orc, cx_runtime, cx_ctx):
    pos, radius, valid, q = make_entries(6000, seed=k)
    io, do, co = orc.entries_query(pos, radius, valid, q, k, brute=True)
    E = cx_runtime.Entries(cx_ctx, pos.shape[0])
    E.build(pos, radius, valid, pos.min(0), pos.max(0))
    ig, dg, cg = E.query(q, k)
    print(f"k={k}: mean neighbours {co.mean():.2f}, full {np.mean(co == k):.2f}")
    assert np.array_equal(co, cg)
    for j in range(q.shape[0]):
        c = co[j]
        assert np.array_equal(io[j, :c], ig[j, :c]), j
        assert np.array_equal(do[j, :c].view(np.uint32), dg[j, :c].view(np.uint32)), j


def test_large_set_scan_spans_several_chunks(orc, cx_runtime, cx_ctx):
    """3 M entry slots, 80 % valid: every pass of the radix sort has more than 2048 tiles of 1024
    keys, so its per-digit scan carries across several block-wide chunks.  The oracle's grid
    search (pinned against brute force in test_oracle_knn.py) gives the expected sets."""
    pos, radius, valid, q = make_entries(3_000_000, seed=5)
    q = q[:4000]
    assert int(valid.sum()) > 2048 * 1024
    io, do, co = orc.entries_query(pos, radius, valid, q, 8, brute=False)
    E = cx_runtime.Entries(cx_ctx, pos.shape[0])
    E.build(pos, radius, valid, pos.min(0), pos.max(0))
    ig, dg, cg = E.query(q, 8)
    print(f"3 M entries: mean neighbours {co.mean():.2f}, full {np.mean(co == 8):.3f}")
    assert co.mean() > 7.5
    assert np.array_equal(co, cg)
    for j in range(q.shape[0]):
        c = co[j]
        assert np.array_equal(io[j, :c], ig[j, :c]), j
        assert np.array_equal(do[j, :c].view(np.uint32), dg[j, :c].view(np.uint32)), j


def test_empty_and_all_invalid(cx_runtime, cx_ctx):
    pos = np.zeros((10, 3), np.float32)
    E = cx_runtime.Entries(cx_ctx, 10)
    E.build(pos, np.ones(10, np.float32), np.zeros(10, np.uint8), np.zeros(3), np.ones(3))
    _, _, c = E.query(np.zeros((5, 3), np.float32), 4)
    assert (c == 0).all()
