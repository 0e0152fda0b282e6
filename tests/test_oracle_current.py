"""Pins of oracle/current.py (cell-bond currents and collision integrals of
the GW loop; the reference calls calc_block_current / calc_block_collision,
gw_mod.f90:384, :795, but does not define them):

* a ballistic chain of cells (ns cells of nb orbitals per slab, identical
  slabs, Fermi contacts at different potentials): the particle current
  (trace of the bond matrix) is the same through every bond, the intra-slab
  ones (from the cell blocks of H and G^<) and the inter-slab ones (from
  green_RGF_RS's cur), and equals the contact current tr that green_RGF_RS
  computes from the lead self-energy (gw_mod.f90:1230-1247), energy by
  energy; a transposed block, a wrong sign or a missing term breaks it;
* the collision integral vanishes at equilibrium (Sigma and G from one Fermi
  function) and equals (f_S - f_G) Gamma A out of equilibrium, which fixes
  the operand order and the signs;
* the energy integrals are the stated weighted sums.
"""
import numpy as np
import pytest

from oracle import current as cur_oracle


def chain(nb, ns, nx, seed):
    """Slabs of ns cells of nb orbitals on a 1D chain: on-site h0, hopping
    H_{c+1,c} = h1 between neighbouring cells.  Returns Hii, H1i
    (H1i[:, :, l] = H_{l,l-1}, the block coupling the last cell of slab l-1
    into the first cell of slab l)."""
    rng = np.random.default_rng(seed)
    h0 = rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))
    h0 = (h0 + h0.conj().T) / 4
    h1 = -1.0 * np.eye(nb) + 0.2 * (rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb)))
    nm = nb * ns
    H = np.zeros((nm, nm), complex)
    for j in range(ns):
        H[j * nb:(j + 1) * nb, j * nb:(j + 1) * nb] = h0
        if j + 1 < ns:
            H[(j + 1) * nb:(j + 2) * nb, j * nb:(j + 1) * nb] = h1
            H[j * nb:(j + 1) * nb, (j + 1) * nb:(j + 2) * nb] = h1.conj().T
    C = np.zeros((nm, nm), complex)
    C[0:nb, (ns - 1) * nb:ns * nb] = h1
    Hii = np.repeat(H[:, :, None], nx, axis=2)
    H1i = np.repeat(C[:, :, None], nx, axis=2)
    return np.asfortranarray(Hii), np.asfortranarray(H1i)


def chain_run(oracle_mod, nb, ns, nx, E, seed=5):
    Hii, H1i = chain(nb, ns, nx, seed)
    nm = nb * ns
    GL = np.zeros((nm, nm, nx, len(E)), complex, order="F")
    cu = np.zeros((nm, nm, nx, len(E)), order="F")
    tr = np.zeros(len(E))
    for e, en in enumerate(E):
        out = oracle_mod.green_rgf_rs(Hii, H1i, en, [0.4, -0.4], [300.0, 300.0])
        GL[..., e] = out["ndens"]
        cu[..., e] = out["cur"]
        tr[e] = out["tr"]
    return Hii, GL, cu, tr


@pytest.mark.parametrize("nb,ns", [(2, 3), (3, 1)])
def test_bond_current_conserved_and_equal_to_contact_current(oracle_mod, nb, ns):
    nx = 4
    E = np.linspace(-0.3, 0.3, 5)
    Hii, GL, cu, tr = chain_run(oracle_mod, nb, ns, nx, E)
    jd, tot, tote = cur_oracle.block_current(Hii, GL, cu, E, 2.0, nb, ns)
    flux = np.einsum("aace->ce", jd)  # trace per bond and energy
    bonds = nx * ns - 1              # the last cell has no bond inside the device
    assert np.all(flux[bonds:] == 0.0)
    scale = np.abs(tr).max()
    assert scale > 1e-3  # a real current flows in the bias window
    for e in range(len(E)):
        err = np.abs(flux[:bonds, e] - tr[e]).max() / scale
        print(f"nb={nb} ns={ns} E={E[e]:+.2f}: tr {tr[e]:.6e}, max |bond - tr| / max|tr| {err:.1e}")
        assert err < 1e-9, (e, flux[:bonds, e], tr[e])


def test_bond_current_breaks_for_wrong_definitions(oracle_mod):
    """the conservation check is sharp: a transposed G^< block or a dropped
    second term in the intra-slab formula violates it"""
    nb, ns, nx = 2, 3, 3
    E = np.array([0.1, 0.2])
    Hii, GL, cu, tr = chain_run(oracle_mod, nb, ns, nx, E)
    jd, _, _ = cur_oracle.block_current(Hii, GL, cu, E, 2.0, nb, ns)
    good = np.einsum("aace->ce", jd)[:nx * ns - 1]
    bad_t = cur_oracle.block_current(Hii, np.swapaxes(GL, 0, 1).copy(), cu, E, 2.0, nb, ns)[0]
    bad_h = cur_oracle.block_current(np.swapaxes(Hii, 0, 1).copy(), GL, cu, E, 2.0, nb, ns)[0]
    for bad in (bad_t, bad_h):
        flux = np.einsum("aace->ce", bad)[:nx * ns - 1]
        assert np.abs(flux - tr[None, :]).max() > 1e-3 * np.abs(tr).max()
    assert np.abs(good - tr[None, :]).max() < 1e-9 * np.abs(tr).max()


def test_current_energy_integrals():
    rng = np.random.default_rng(2)
    nb, ns, nx, nen = 2, 2, 3, 6
    nm = nb * ns
    Hii = rng.standard_normal((nm, nm, nx)) + 1j * rng.standard_normal((nm, nm, nx))
    GL = rng.standard_normal((nm, nm, nx, nen)) + 1j * rng.standard_normal((nm, nm, nx, nen))
    cu = rng.standard_normal((nm, nm, nx, nen))
    E = np.linspace(-1.0, 1.5, nen)
    jd, tot, tote = cur_oracle.block_current(Hii, GL, cu, E, 2.0, nb, ns)
    w = (E[1] - E[0]) * 1.6022e-19 / (2 * 3.14159265359) / 1.0546e-34 * 1.6022e-19 * 2.0
    np.testing.assert_allclose(tot, jd.sum(axis=3) * w, rtol=1e-13, atol=0)
    np.testing.assert_allclose(tote, (jd * E[None, None, None, :]).sum(axis=3) * w, rtol=1e-12,
                               atol=1e-30)
    # single energy: no integration step, zero totals
    _, t1, e1 = cur_oracle.block_current(Hii, GL[..., :1], cu[..., :1], E[:1], 2.0, nb, ns)
    assert not t1.any() and not e1.any()


def _herm_psd(rng, n):
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return X @ X.conj().T / n


def _fermi(E, mu, kT=0.025):
    return 1.0 / (1.0 + np.exp((E - mu) / kT))


def test_collision_equilibrium_and_closed_form():
    rng = np.random.default_rng(3)
    nm, nx, nen = 4, 2, 5
    E = np.linspace(-0.2, 0.2, nen)
    Gam = np.stack([np.stack([_herm_psd(rng, nm) for _ in range(nen)], -1) for _ in range(nx)], 2)
    A = np.stack([np.stack([_herm_psd(rng, nm) for _ in range(nen)], -1) for _ in range(nx)], 2)
    fS, fG = _fermi(E, 0.05), _fermi(E, -0.05)

    def lesser(X, f):
        return 1j * X * f

    def greater(X, f):
        return -1j * X * (1.0 - f)

    # equilibrium: one Fermi function for Sigma and G -> zero
    I0, S0 = cur_oracle.block_collision(lesser(Gam, fG), greater(Gam, fG), lesser(A, fG), greater(A, fG), E, 2.0)
    assert np.abs(S0).max() < 1e-13 * np.abs(np.einsum("ijxe,jkxe->ikxe", Gam, A)).max()
    # out of equilibrium: (f_S - f_G) Gamma A, not A Gamma
    I1, S1 = cur_oracle.block_collision(lesser(Gam, fS), greater(Gam, fS), lesser(A, fG), greater(A, fG), E, 2.0)
    ref = np.einsum("ijxe,jkxe->ikxe", Gam, A) * (fS - fG)[None, None, None, :]
    swapped = np.einsum("ijxe,jkxe->ikxe", A, Gam) * (fS - fG)[None, None, None, :]
    assert np.abs(S1 - ref).max() < 1e-13 * np.abs(ref).max()
    assert np.abs(S1 - swapped).max() > 1e-2 * np.abs(ref).max()
    f = (E[1] - E[0]) * 2.0 / (2 * 3.14159265359)
    np.testing.assert_allclose(I1, S1.sum(axis=3) * f, rtol=1e-13, atol=1e-30)
