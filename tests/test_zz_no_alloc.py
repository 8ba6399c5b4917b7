"""cx_entries_search_static (the entry build and query that allocates nothing) against
cx_entries_build + cx_entries_query on the same inputs: neighbour indices, distances and counts
bit for bit, on two calls in a row whose second is smaller (the static scratch keeps the first
call's contents), plus the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_entries(n, nq, seed, centers=16, spread=0.6):
    rng = np.random.default_rng(seed)
    c = rng.random((centers, 3)) * 10
    pos = c[rng.integers(0, centers, n)] + rng.normal(scale=spread, size=(n, 3))
    radius = rng.uniform(0.05, 0.5, n)
    valid = (rng.random(n) < 0.8).astype(np.uint8)
    q = pos[rng.integers(0, n, nq)] + rng.normal(scale=0.2, size=(nq, 3))
    return pos.astype(np.float32), radius.astype(np.float32), valid, q.astype(np.float32)


def reference(rt, ctx, pos, radius, valid, q, k):
    E = rt.Entries(ctx, pos.shape[0])
    E.build(pos, radius, valid, pos.min(0), pos.max(0))
    return E.query(q, k)


def check_equal(a, b):
    ia, da, ca = a
    ib, db, cb = b
    assert np.array_equal(ca, cb)
    for j in range(ca.shape[0]):
        c = ca[j]
        assert np.array_equal(ia[j, :c], ib[j, :c]), j
        assert np.array_equal(da[j, :c].view(np.uint32), db[j, :c].view(np.uint32)), j


@pytest.mark.parametrize("k", [8, 31])
def test_static_search_matches_entries_build_and_query(k, cx_runtime, cx_ctx):
    # first call near the 65536-entry limit, second one much smaller: the second must not see
    # anything the first left in the shared scratch
    for n, nq, seed in ((65536, 20000, 11 + k), (3000, 1500, 12 + k)):
        pos, radius, valid, q = make_entries(n, nq, seed)
        want = reference(cx_runtime, cx_ctx, pos, radius, valid, q, k)
        got = cx_runtime.entries_search_static(cx_ctx, pos, radius, valid, pos.min(0), pos.max(0), q, k)
        print(f"n={n} k={k}: mean neighbours {want[2].mean():.2f}", flush=True)
        assert want[2].max() > 0
        check_equal(want, got)


def test_static_search_empty_and_all_invalid(cx_runtime, cx_ctx):
    pos, radius, valid, q = make_entries(5000, 100, 3)
    cx_runtime.entries_search_static(cx_ctx, pos, radius, valid, pos.min(0), pos.max(0), q, 4)
    z = np.zeros((10, 3), np.float32)
    _, _, c = cx_runtime.entries_search_static(cx_ctx, z, np.ones(10, np.float32), np.zeros(10, np.uint8),
                                               np.zeros(3), np.ones(3), np.zeros((5, 3), np.float32), 4)
    assert (c == 0).all()
    _, _, c = cx_runtime.entries_search_static(cx_ctx, np.zeros((0, 3), np.float32), np.zeros(0, np.float32),
                                               np.zeros(0, np.uint8), np.zeros(3), np.ones(3),
                                               np.zeros((5, 3), np.float32), 4)
    assert (c == 0).all()


def test_static_search_rejects_more_than_its_capacity(cx_runtime, cx_ctx):
    pos, radius, valid, q = make_entries(65537, 10, 5)
    with pytest.raises(cx_runtime.CxError, match="65536"):
        cx_runtime.entries_search_static(cx_ctx, pos, radius, valid, pos.min(0), pos.max(0), q, 8)
