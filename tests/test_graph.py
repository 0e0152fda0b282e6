"""The slicing oracle (oracle/graph.py, graph_partition.f90:282-571) pinned
by closed forms (chain, ring, grid), by breadth-first distances from scipy's
shortest paths, and by the defining property of a slicing (no coupling
beyond the neighbouring slice) on random connected graphs, with a planted
violation that the checker must catch."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import shortest_path

from oracle import graph as G


def _chain(n):
    return G.graph_from_pairs(n, [(i, i + 1) for i in range(n - 1)])


def _random_connected(rng, n):
    perm = rng.permutation(n)
    pairs = [(perm[i], perm[rng.integers(0, i)]) for i in range(1, n)]
    pairs += [tuple(e) for e in rng.integers(0, n, (int(rng.integers(0, 2 * n)), 2))]
    return pairs


def test_graph_from_pairs_drops_self_pairs_and_repeats():
    g = G.graph_from_pairs(4, [(0, 1), (1, 0), (2, 2), (1, 2), (1, 2), (3, 1)])
    assert [sorted(x) for x in g] == [[1], [0, 2, 3], [1], [1]]


def test_chain_closed_forms():
    g = _chain(10)
    assert G.slice_1contact(g, [0]) == [[i] for i in range(10)]
    assert G.slice_1contact(g, [9]) == [[9 - i] for i in range(10)]
    assert G.slice_2contacts(g, [0], [9], 1) == [[i] for i in range(10)]
    # first cut at 4 | 5 (points strictly closer to the left contact go left),
    # then 0 1 | 2 3 4 and 5 6 | 7 8 9 (ties at distance 2 and 2 go right)
    assert G.slice_2contacts(g, [0], [9], 3) == [[0, 1], [2, 3, 4], [5, 6], [7, 8, 9]]
    assert G.slice_2contacts(g, [0], [9], 10) == [[0, 1, 2, 3, 4], [5, 6, 7, 8, 9]]


def test_ring_layers():
    n = 8  # ring of 2n points, contact at 0: layers {k, 2n - k}
    g = G.graph_from_pairs(2 * n, [(i, (i + 1) % (2 * n)) for i in range(2 * n)])
    want = [[0]] + [sorted([k, 2 * n - k]) for k in range(1, n)] + [[n]]
    assert G.slice_1contact(g, [0]) == want


def test_grid_columns():
    nx, ny = 7, 4
    pid = lambda x, y: x + nx * y
    pairs = [(pid(x, y), pid(x + 1, y)) for y in range(ny) for x in range(nx - 1)]
    pairs += [(pid(x, y), pid(x, y + 1)) for y in range(ny - 1) for x in range(nx)]
    g = G.graph_from_pairs(nx * ny, pairs)
    cols = [[pid(x, y) for y in range(ny)] for x in range(nx)]
    assert G.slice_1contact(g, cols[0]) == cols
    assert G.slice_2contacts(g, cols[0], cols[-1], 1) == cols


def test_shared_contact_and_disconnected():
    g = _chain(5)
    assert G.slice_2contacts(g, [0, 2], [2, 4], 1) == [[0, 1, 2, 3, 4]]
    g2 = G.graph_from_pairs(5, [(0, 1), (3, 4)])
    with pytest.raises(ValueError):
        G.slice_1contact(g2, [0])
    with pytest.raises(ValueError):
        G.slice_2contacts(g2, [0], [4], 1)


def test_distances_equal_shortest_paths():
    rng = np.random.default_rng(4)
    for _ in range(20):
        n = int(rng.integers(2, 120))
        pairs = _random_connected(rng, n)
        g = G.graph_from_pairs(n, pairs)
        src = sorted(rng.choice(n, int(rng.integers(1, 4)), replace=False).tolist())
        r = [i for i, j in pairs if i != j] + [j for i, j in pairs if i != j]
        c = [j for i, j in pairs if i != j] + [i for i, j in pairs if i != j]
        A = coo_matrix((np.ones(len(r)), (r, c)), shape=(n, n)).tocsr()
        sp = shortest_path(A, unweighted=True, indices=src).min(axis=0)
        assert np.array_equal(G.dist(g, src), sp.astype(np.int64))


@pytest.mark.parametrize("seed", range(5))
def test_random_slicings_are_valid(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(30):
        n = int(rng.integers(2, 250))
        g = G.graph_from_pairs(n, _random_connected(rng, n))
        left = sorted(rng.choice(n, int(rng.integers(1, max(2, n // 10))), replace=False).tolist())
        right = sorted(rng.choice(n, int(rng.integers(1, max(2, n // 10))), replace=False).tolist())
        nmax = int(rng.integers(1, 60))
        for s in (G.slice_1contact(g, left), G.slice_2contacts(g, left, right, nmax)):
            assert sorted(p for sl in s for p in sl) == list(range(n))  # a partition
            assert all(sl == sorted(sl) for sl in s)
            assert G.check_slicing(g, s) == 0
        s = G.slice_2contacts(g, left, right, nmax)
        if not set(left) & set(right):
            assert set(left) <= set(s[0]) and set(right) <= set(s[-1])


def test_checker_catches_a_point_two_slices_away():
    g = _chain(6)
    s = G.slice_1contact(g, [0])
    assert G.check_slicing(g, s) == 0
    bad = [list(x) for x in s]
    bad[2].remove(2)
    bad[4].append(2)  # point 2 touches 1 and 3, now two slices from slice 1
    assert G.check_slicing(g, bad) == 2
    bad[4].remove(2)
    assert G.check_slicing(g, bad) == -1
