"""CPU restatement of the reference's device slicing (graph_partition.f90).

TEST INFRASTRUCTURE ONLY: the checker of the library's qt_graph_slice;
imported by tests/, never by the product.

A device graph (one point per orbital, an edge per nonzero coupling) is cut
into slices that connect only to their neighbouring slices, the block
structure the RGF sweeps need.  Points are 0-based here (the reference counts
from 1); slices are lists of point ids in ascending order, as the reference's
slice tables hold them.

* ``dist``: breadth-first distance of every point to a set of edge points
  (graph_partition.f90:282-316); unreachable points keep UNREACHED.
* ``slice_1contact`` (:426-450): slice i holds the points at distance i from
  the contact.
* ``slice_2contacts`` (:452-541): recursive bisection between two contacts:
  a point goes to the left part when it is strictly closer to the left edge
  (:318-328, ties go right), each part's new edge is its points with a
  neighbour in the other part (:347-367), a part is sliced again while it
  has more than nmax points and more points than its outer edge, and the
  slices of the left part come first.  Contacts that share a point make the
  whole graph one slice.
* ``check_slicing`` (testSlicing, :543-571): 0 when every point is in a slice and no
  neighbours are more than one slice apart, else the first offending slice
  (1-based, as the reference returns it), -1 when a point is in no slice.
* ``graph_from_pairs``: the connectivity of an undirected graph from point
  pairs, self pairs and repeats dropped (AddEdge, :101-116); the reference's
  COO converter has an empty body (:587-595), so the graph of a matrix is
  taken from its nonzero off-diagonal entries.
"""
from __future__ import annotations

from collections import deque
from typing import List, Sequence

import numpy as np

UNREACHED = np.iinfo(np.int64).max


def graph_from_pairs(npts: int, pairs) -> List[List[int]]:
    nb = [[] for _ in range(npts)]
    seen = set()
    for i, j in pairs:
        i, j = int(i), int(j)
        if i == j:
            continue
        key = (min(i, j), max(i, j))
        if key in seen:
            continue
        seen.add(key)
        nb[i].append(j)
        nb[j].append(i)
    return nb


def graph_from_matrix(H) -> List[List[int]]:
    """points = rows of a square matrix, edges = nonzero off-diagonal entries"""
    H = np.asarray(H)
    r, c = np.nonzero(H)
    return graph_from_pairs(H.shape[0], zip(r, c))


def dist(g: Sequence[Sequence[int]], edge: Sequence[int]) -> np.ndarray:
    d = np.full(len(g), UNREACHED, dtype=np.int64)
    q = deque()
    for e in edge:
        d[e] = 0
        q.append(e)
    while q:
        p = q.popleft()
        for p2 in g[p]:
            if d[p2] == UNREACHED:
                d[p2] = d[p] + 1
                q.append(p2)
    return d


def slice_1contact(g, edge) -> List[List[int]]:
    d = dist(g, edge)
    if (d == UNREACHED).any():
        raise ValueError("a point is not connected to the contact")
    return [sorted(np.nonzero(d == k)[0].tolist()) for k in range(int(d.max()) + 1)]


def _subgraph(g, pts: List[int]):
    idx = {p: i for i, p in enumerate(pts)}
    return [[idx[q] for q in g[p] if q in idx] for p in pts]


def slice_2contacts(g, e1: Sequence[int], e2: Sequence[int], nmax: int) -> List[List[int]]:
    """The graph must be connected (the reference recurses without end when
    the contacts lie in different components)."""
    if (dist(g, list(e1)) == UNREACHED).any():
        raise ValueError("the graph is not connected")
    return _bisect(g, e1, e2, nmax)


def _bisect(g, e1, e2, nmax):
    n = len(g)
    if set(e1) & set(e2):
        return [list(range(n))]
    dl, dr = dist(g, e1), dist(g, e2)
    part = np.where(dl < dr, 1, 2)
    # edge points of each part: points with a neighbour in the other part
    edges = {1: [], 2: []}
    for i in range(n):
        if any(part[j] != part[i] for j in g[i]):
            edges[int(part[i])].append(i)
    out = []
    for p, (left, right) in ((1, (list(e1), None)), (2, (None, list(e2)))):
        pts = np.nonzero(part == p)[0].tolist()
        outer = left if p == 1 else right
        if len(pts) > nmax and len(pts) > len(outer):
            idx = {q: i for i, q in enumerate(pts)}
            a = [idx[q] for q in (left if p == 1 else edges[2])]
            b = [idx[q] for q in (edges[1] if p == 1 else right)]
            sub = _bisect(_subgraph(g, pts), a, b, nmax)
            out += [[pts[i] for i in s] for s in sub]
        else:
            out.append(pts)
    return out


def check_slicing(g, slices) -> int:
    where = np.full(len(g), -1)
    for k, s in enumerate(slices):
        where[list(s)] = k
    if (where == -1).any():
        return -1
    for k, s in enumerate(slices):
        for p in s:
            for q in g[p]:
                if abs(where[q] - k) > 1:
                    return k + 1
    return 0


def slice_of(slices, npts: int) -> np.ndarray:
    """slice number of every point (-1 for points in no slice)"""
    where = np.full(npts, -1, dtype=np.int32)
    for k, s in enumerate(slices):
        where[list(s)] = k
    return where
