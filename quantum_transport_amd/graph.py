"""Device slicing: a sparse device matrix cut into slabs that couple only to
their neighbours, through the library's qt_graph_slice (the role of
graph_partition::slice, graph_partition.f90:426-541), and the uneven block
lists the RGF sweeps take.

* ``slice_graph``: slice of every point of a graph given as point pairs;
* ``slice_matrix``: the same for the graph of a matrix's nonzero couplings;
* ``auto_slice``: the choice the reference's COO builder makes
  (deviceHam_mod.f90:145-167): breadth-first slices from the left contact,
  else from the right contact, else bisection between both contacts;
* ``blocks``: the slab-diagonal blocks Hii and couplings H_{s,s-1} of a
  matrix under a slicing (points of a slab in ascending order).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import rgf


def slice_graph(npts: int, pairs, left: Sequence[int], right: Optional[Sequence[int]] = None,
                nmax: int = 100) -> Tuple[np.ndarray, int]:
    """(slice of each point, number of slices); right = None slices by
    distance from the left contact, else bisects between the contacts."""
    e = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    lt = np.ascontiguousarray(left, dtype=np.int32)
    rt = np.ascontiguousarray([] if right is None else right, dtype=np.int32)
    out = np.zeros(npts, np.int32)
    ns = C.c_int32()
    rgf.check(rgf.lib().qt_graph_slice(npts, len(e), e.ctypes.data if len(e) else None, len(lt),
                                       lt.ctypes.data, len(rt), rt.ctypes.data if len(rt) else None,
                                       nmax, out.ctypes.data, C.byref(ns)))
    return out, ns.value


def matrix_pairs(H) -> np.ndarray:
    """point pairs (i < j) of the nonzero off-diagonal entries of a square matrix"""
    r, c = np.nonzero(np.asarray(H))
    off = r != c
    pairs = np.stack([np.minimum(r[off], c[off]), np.maximum(r[off], c[off])], 1)
    return np.unique(pairs, axis=0) if len(pairs) else np.zeros((0, 2), np.int64)


def slice_matrix(H, left: Sequence[int], right: Optional[Sequence[int]] = None,
                 nmax: int = 100) -> Tuple[np.ndarray, int]:
    return slice_graph(np.asarray(H).shape[0], matrix_pairs(H), left, right, nmax)


def _sizes(slice_of: np.ndarray, ns: int) -> np.ndarray:
    return np.bincount(slice_of, minlength=ns)


def auto_slice(H, nleft: int, nright: int, nmax: int = 100) -> Tuple[np.ndarray, int, str]:
    """Contacts are the first nleft and the last nright points.  Slices from
    the left contact unless the last slice is smaller than the right contact
    or the slices are unbalanced (largest - smallest > 2 x (smallest + 1),
    the reference's test on its slice tables, which count one more); then
    from the right contact (slices renumbered from the left), with the same
    test on the first slice; else two-contact bisection.  Returns the
    slicing and which of "left", "right", "both" was taken."""
    n = np.asarray(H).shape[0]
    pairs = matrix_pairs(H)
    left = np.arange(nleft)
    right = np.arange(n - nright, n)

    def unbalanced(sz):
        return sz.max() - sz.min() > 2 * (sz.min() + 1)

    s, ns = slice_graph(n, pairs, left, None, nmax)
    sz = _sizes(s, ns)
    if not (sz[-1] < nright or unbalanced(sz)):
        return s, ns, "left"
    s, ns = slice_graph(n, pairs, right, None, nmax)
    s = ns - 1 - s
    sz = _sizes(s, ns)
    if not (sz[0] < nleft or unbalanced(sz)):
        return s, ns, "right"
    s, ns = slice_graph(n, pairs, left, right, nmax)
    return s, ns, "both"


def blocks(H, slice_of: np.ndarray, ns: int) -> Tuple[List[int], List[np.ndarray], List[np.ndarray], np.ndarray]:
    """(sizes, Hii, couplings H_{s,s-1} for s = 1..ns-1, point order): the
    slab blocks of H under the slicing; raises if two slices that are not
    neighbours are coupled."""
    H = np.asarray(H)
    order = np.argsort(slice_of, kind="stable")
    sizes = _sizes(slice_of, ns).tolist()
    if min(sizes) == 0:
        raise ValueError("empty slice")
    off = np.concatenate([[0], np.cumsum(sizes)])
    P = H[np.ix_(order, order)]
    for s in range(ns):
        far = P[off[s]:off[s + 1], off[min(ns, s + 2)]:]
        if far.any():
            raise ValueError(f"slice {s} couples beyond its neighbour")
    Hii = [P[off[s]:off[s + 1], off[s]:off[s + 1]].copy() for s in range(ns)]
    H1i = [P[off[s]:off[s + 1], off[s - 1]:off[s]].copy() for s in range(1, ns)]
    return sizes, Hii, H1i, order


def sliced_device(H, slice_of: np.ndarray, ns: int, lead_hopping: float = -1.0):
    """hamiltonian.Device of the slabs of H under a slicing (one k point).
    The leads repeat the outer slabs, each copy coupled to the next by
    lead_hopping x identity (H1i[0] and H1i[nx])."""
    from .hamiltonian import Device

    sizes, Hii, H1i, order = blocks(H, slice_of, ns)
    Hl = [np.asfortranarray(lead_hopping * np.eye(sizes[0], dtype=complex))]
    Hl += [np.asfortranarray(h, dtype=complex) for h in H1i]
    Hl.append(np.asfortranarray(lead_hopping * np.eye(sizes[-1], dtype=complex)))
    return Device(sizes, [[np.asfortranarray(h, dtype=complex) for h in Hii]], [Hl], None, 1.0, 0, 1), order
