"""Pins the extended-precision sweep (tests/_xprec.py) that the GPU parity
tests use as the exact answer at ill-conditioned energies: its block
recursion equals the dense device inverse carried out in the same precision
(to ~1e-16, far below any FP64 error it is used to measure), its lead agrees
with the chain closed form and with the golden-pinned oracle, and the
committed fixtures are what the generator computes."""
import os

import numpy as np
import pytest

from quantum_transport_amd import hamiltonian as hm
from tests import _dense
from tests._xprec import X, dag, xinv, xrgf_forward, xsancho

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOLTZ = 8.61734e-05


def test_xinv_is_extended_precision():
    """A Hilbert-like matrix whose FP64 inverse loses ~10 digits: the
    extended inverse's residual sits at the 80-bit rounding level."""
    n = 9
    A = 1.0 / (np.arange(n)[:, None] + np.arange(n)[None, :] + 1.0) + 1e-3j * np.eye(n)
    Xi = xinv(A[None])[0]
    res = np.abs(A.astype(X) @ Xi - np.eye(n)).max()
    assert res < 1e-13, res
    assert res < 1e-2 * np.abs(A @ np.linalg.inv(A) - np.eye(n)).max()


@pytest.mark.parametrize("E", [-0.7, 1.25])
def test_xsancho_chain_and_oracle(oracle_mod, E):
    eps, t = 0.2, -1.0
    g = xsancho([E], 1e-5, np.array([[eps]]), np.array([[t]]))[0, 0, 0]
    ref = _dense.chain_surface_g(E + 1e-5j, eps, t)
    assert abs(complex(g) - ref) <= 1e-9 * abs(ref)
    _, dev = hm.synthetic_device(8, 3, seed=4, coupling_density=0.4)
    H00, H10 = dev.Hii[0][0], dev.H1i[0][0].conj().T
    gx = xsancho([E], 1e-5, H00, H10)[0].astype(np.complex128)
    go, _, _, _ = oracle_mod.sancho(None, H00, H10, E)
    # inside a band the FP64 decimation at eta = 1e-5 is itself only good to
    # ~1e-11 .. 1e-9 (the golden test_cnt values show the same spread)
    assert np.abs(gx - go).max() <= 1e-9 * np.abs(go).max()


@pytest.mark.parametrize("E", [-0.31, 0.17])
def test_xrgf_equals_dense_inverse(E):
    """Diagonal blocks of the recursion against the whole-device inverse,
    both in extended precision, with the same lead self-energies and the
    forward variant's occupations (first contact f(E - mur, TEMPr), last
    f(E - mul, TEMPl), rgf_mod.f90:100, :182); tr = -tre = Caroli
    transmission times the Fermi window."""
    n, nx = 6, 4
    _, dev = hm.synthetic_device(n, nx, seed=12, coupling_density=0.5)
    Hii, H1i = dev.Hii[0], dev.H1i[0]
    mul, mur, T = 0.0, -0.3, 300.0
    r = xrgf_forward(Hii, H1i, [E], mul, mur, T, T)
    gL = xsancho([E], 1e-5, Hii[0], H1i[0].conj().T)[0]
    gR = xsancho([E], 1e-5, Hii[-1], H1i[nx])[0]
    U0, UN = H1i[0].astype(X), H1i[nx].astype(X)
    sigL = U0 @ gL @ dag(U0)
    sigR = dag(UN) @ gR @ UN
    Hd, off = _dense.assemble(Hii, H1i)
    N = Hd.shape[0]
    Sig = np.zeros((N, N), X)
    Sig[:n, :n] += sigL
    Sig[-n:, -n:] += sigR
    G = xinv((np.longdouble(E) * np.eye(N, dtype=X) - Hd.astype(X) - Sig)[None])[0]
    fF = 1.0 / (1.0 + np.exp(np.longdouble((E - mur) / (BOLTZ * T))))
    fL = 1.0 / (1.0 + np.exp(np.longdouble((E - mul) / (BOLTZ * T))))
    GamL = np.zeros((N, N), X)
    GamR = np.zeros((N, N), X)
    GamL[:n, :n] = 1j * (sigL - dag(sigL))
    GamR[-n:, -n:] = 1j * (sigR - dag(sigR))
    Gl = G @ (1j * (GamL * fF + GamR * fL)) @ dag(G)
    Gg = Gl + G - dag(G)
    scale = np.abs(np.diagonal(G)).max()
    for key, M in (("gr_diag", G), ("gl_diag", Gl), ("gg_diag", Gg)):
        d = np.diagonal(M).astype(np.complex128)
        assert np.abs(r[key][0] - d).max() <= 1e-14 * scale, key
    T_caroli = np.trace(GamL[:n, :n] @ G[:n, -n:] @ GamR[-n:, -n:] @ dag(G[:n, -n:])).real
    Tw = float(T_caroli * (fF - fL))
    assert abs(Tw) > 1e-6
    assert abs(r["tr"][0] + r["tre"][0]) <= 1e-13 * abs(Tw)
    assert abs(abs(r["tr"][0]) - abs(Tw)) <= 1e-12 * abs(Tw)


def test_c1_fixture_reproduces_and_agrees_with_oracle(oracle_mod):
    """Two energies of tests/golden/c1_xprec.npz recomputed: equal to the
    stored spectra; at a well-conditioned one the oracle sits within 1e-11."""
    f = np.load(os.path.join(GOLDEN, "c1_xprec.npz"))
    _, dev = hm.synthetic_device(64, 32, seed=1001)
    E = hm.energy_grid(-3.0, 3.0, 256)
    assert np.array_equal(f["E"], E)
    idx = np.array([50, 177])
    r = xrgf_forward(dev.Hii[0], dev.H1i[0], E[idx], 0.0, -0.1)
    off = np.arange(33) * 64
    ldos = np.stack([-2 * r["gr_diag"][:, off[s]:off[s + 1]].sum(1).imag for s in range(32)], 1)
    ndos = np.stack([r["gl_diag"][:, off[s]:off[s + 1]].sum(1).imag for s in range(32)], 1)
    scale = np.abs(f["ldos"][idx]).max()
    assert np.abs(ldos - f["ldos"][idx]).max() <= 1e-14 * scale
    assert np.abs(ndos - f["ndos"][idx]).max() <= 1e-14 * scale
    assert np.abs(r["tr"] - f["tr"][idx]).max() <= 1e-14 * np.abs(f["tr"]).max()
    o = oracle_mod.rgf_batch(dev, E[idx[:1]], None, 0.0, -0.1)
    lo = -2 * np.stack([o["gr_diag"][:, off[s]:off[s + 1]].sum(1).imag for s in range(32)], 1)
    assert np.abs(lo - f["ldos"][idx[:1]]).max() <= 1e-11 * np.abs(f["ldos"][idx[0]]).max()
