"""The library's device slicing (qt_graph_slice) against the oracle's
restatement of graph_partition::slice (graph_partition.f90:426-541), the
reference's slicing choice (deviceHam_mod.f90:145-167), and a sparse device
sliced into uneven slabs and solved on the GPU against the oracle's sweep."""
import numpy as np
import pytest

from oracle import graph as G
from quantum_transport_amd import graph as qg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf

    rgf.init(0)
    return rgf


def _random_connected(rng, n):
    perm = rng.permutation(n)
    pairs = [(perm[i], perm[rng.integers(0, i)]) for i in range(1, n)]
    pairs += [tuple(e) for e in rng.integers(0, n, (int(rng.integers(0, 2 * n)), 2))]
    return pairs


def test_slices_equal_oracle_on_random_graphs(qt):
    rng = np.random.default_rng(7)
    for trial in range(120):
        n = int(rng.integers(2, 1500 if trial % 10 == 0 else 300))
        pairs = _random_connected(rng, n)
        g = G.graph_from_pairs(n, pairs)
        left = sorted(rng.choice(n, int(rng.integers(1, max(2, n // 10))), replace=False).tolist())
        right = sorted(rng.choice(n, int(rng.integers(1, max(2, n // 10))), replace=False).tolist())
        nmax = int(rng.integers(1, 80))
        for r in (None, right):
            ref = G.slice_1contact(g, left) if r is None else G.slice_2contacts(g, left, r, nmax)
            got, ns = qg.slice_graph(n, pairs, left, r, nmax)
            assert ns == len(ref), (trial, r is None)
            assert np.array_equal(got, G.slice_of(ref, n)), (trial, r is None)


def test_errors(qt):
    pairs = [(0, 1), (1, 2), (3, 4)]
    with pytest.raises(qt.QTError):
        qg.slice_graph(5, pairs, [0], None)  # 3 and 4 unreachable
    with pytest.raises(qt.QTError):
        qg.slice_graph(5, pairs, [0], [4])
    with pytest.raises(qt.QTError):
        qg.slice_graph(3, [(0, 5)], [0], None)  # point id out of range
    with pytest.raises(qt.QTError):
        qg.slice_graph(3, [(0, 1), (1, 2)], [0], [2], nmax=0)
    s, ns = qg.slice_graph(3, [(0, 1), (1, 2)], [1], [1])  # shared contact: one slice
    assert ns == 1 and not s.any()


def _wire(nlayer, ny, nz, seed, diag=0.15, vacancy=0.08):
    """orbitals on an nlayer x ny x nz lattice (layer index slowest), hopping
    -1 between lattice neighbours with some bonds removed, plus random
    diagonal bonds between neighbouring layers; random on-site energies"""
    rng = np.random.default_rng(seed)
    n = nlayer * ny * nz
    pid = lambda x, y, z: (x * ny + y) * nz + z
    H = np.zeros((n, n))
    for x in range(nlayer):
        for y in range(ny):
            for z in range(nz):
                p = pid(x, y, z)
                H[p, p] = rng.uniform(-0.5, 0.5)
                for dx, dy, dz in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    X, Y, Z = x + dx, y + dy, z + dz
                    if X < nlayer and Y < ny and Z < nz and (dx == 1 or rng.random() > vacancy):
                        H[p, pid(X, Y, Z)] = H[pid(X, Y, Z), p] = -1.0
                if x + 1 < nlayer and rng.random() < diag:
                    q = pid(x + 1, int(rng.integers(ny)), int(rng.integers(nz)))
                    H[p, q] = H[q, p] = -0.3
    return H


def test_auto_slice_policy(qt):
    """A wire whose layers are breadth-first layers from the left face takes
    the left slicing; with a right contact larger than the last layer it
    moves on; every slicing it returns is valid."""
    H = _wire(10, 4, 4, 1, diag=0.0)
    s, ns, how = qg.auto_slice(H, 16, 16, nmax=40)
    assert how == "left" and ns == 10 and (np.bincount(s) == 16).all()
    s2, ns2, how2 = qg.auto_slice(H, 16, 20, nmax=40)
    assert how2 in ("right", "both")
    g = G.graph_from_matrix(H)
    for sl, nsl in ((s, ns), (s2, ns2)):
        assert G.check_slicing(g, [np.nonzero(sl == k)[0].tolist() for k in range(nsl)]) == 0


def test_sliced_device_solves_like_the_oracle(qt, oracle_mod):
    """A 14-layer 5 x 5 wire with vacancies and diagonal bonds, sliced by
    bisection between its end faces (nmax 60: slabs of one and two layers,
    25 and 50 orbitals): uneven slabs, solved on the
    device with the couplings of the slab sizes and compared with the
    oracle's sweep on the same blocks."""
    from tests.test_gpu_rgf import check_parity

    H = _wire(14, 5, 5, 3)
    n = H.shape[0]
    left, right = np.arange(25), np.arange(n - 25, n)
    s, ns = qg.slice_matrix(H, left, right, nmax=60)
    dev, order = qg.sliced_device(H, s, ns)
    print("slab sizes", dev.nm)
    assert len(set(dev.nm)) > 1  # uneven
    E = np.linspace(-2.5, 2.5, 7)
    Hd = qt.Hamiltonian.from_device(dev)
    gpu = qt.solve_batch(Hd, E, None, mul=0.0, mur=-0.1, coupling_fix=True)
    Hd.close()
    mkl = oracle_mod.rgf_batch(dev, E, None, 0.0, -0.1, coupling_fix=True, flavour="mkl")
    plain = oracle_mod.rgf_batch(dev, E, None, 0.0, -0.1, coupling_fix=True)
    check_parity(gpu, mkl, plain, dev.nm, "sliced wire")
