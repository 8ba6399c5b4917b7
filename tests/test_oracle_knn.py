"""Neighbour search of the oracle (prefix_search_hit.cu under the (distance, index) rule):
the grid search equals the brute-force search, and the candidate rule is pinned."""
import numpy as np


def test_grid_equals_brute_force(orc):
    rng = np.random.default_rng(31)
    n = 4000
    pos = (rng.random((n, 3)) * 4).astype(np.float32)
    r = rng.uniform(0.05, 0.4, n).astype(np.float32)
    valid = (rng.random(n) < 0.85).astype(np.uint8)
    q = (rng.random((1500, 3)) * 4).astype(np.float32)
    for k in (1, 8, 31):
        a = orc.entries_query(pos, r, valid, q, k, brute=False)
        b = orc.entries_query(pos, r, valid, q, k, brute=True)
        assert np.array_equal(a[2], b[2])
        for j in range(q.shape[0]):
            c = a[2][j]
            assert np.array_equal(a[0][j, :c], b[0][j, :c])
            assert np.array_equal(a[1][j, :c], b[1][j, :c])


def test_candidate_rule_and_order(orc):
    pos = np.asarray([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 2], [-1, 0, 0]], np.float32)
    r = np.asarray([1.0, 1.0, 0.5, 5.0, 1.0], np.float32)
    valid = np.asarray([1, 1, 1, 1, 0], np.uint8)
    q = np.asarray([[0, 0, 0]], np.float32)
    idx, dist, cnt = orc.entries_query(pos, r, valid, q, 8, brute=True)
    # entry 2 is at distance 1 > its own radius 0.5; entry 4 is invalid; the query's own
    # radius plays no part (prefix_search_hit.cu:33-42 uses the entry's radius)
    assert cnt[0] == 3
    assert list(idx[0, :3]) == [0, 1, 3]
    assert list(dist[0, :3]) == [0.0, 1.0, 2.0]
    # ties at equal distance are ordered by index, and k truncates
    pos2 = np.asarray([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0]], np.float32)
    idx2, _, cnt2 = orc.entries_query(pos2, np.full(4, 2.0, np.float32), np.ones(4, np.uint8), q, 2, brute=False)
    assert cnt2[0] == 2 and list(idx2[0]) == [0, 1]


def test_radius_boundary_inclusive(orc):
    pos = np.asarray([[0.5, 0, 0]], np.float32)
    r = np.asarray([0.5], np.float32)
    q = np.asarray([[0, 0, 0], [np.nextafter(np.float32(0), np.float32(-1)) - np.float32(1e-7), 0, 0]], np.float32)
    _, _, cnt = orc.entries_query(pos, r, np.ones(1, np.uint8), q, 4, brute=True)
    assert cnt[0] == 1 and cnt[1] == 0
