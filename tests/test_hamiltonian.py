"""Host-side input preparation: deterministic generator, ham_dat round trip,
w90_MAT_DEF block convention (wannierHam3d.f90:273-301), energy grid
(negf_mod.f90:100-101)."""
import numpy as np
import pytest

from quantum_transport_amd import hamiltonian as hm


def test_splitmix_known_values():
    # SplitMix64 reference outputs for seed 0 (Steele et al. 2014 / Vigna's splitmix64.c)
    r = hm.SplitMix64(0).u64(3)
    assert [int(x) for x in r] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_generator_is_deterministic_and_hermitian():
    a, da = hm.synthetic_device(24, 3, seed=5, kz=(0.0, 0.9), tz_density=0.1)
    b, db = hm.synthetic_device(24, 3, seed=5, kz=(0.0, 0.9), tz_density=0.1)
    for ik in range(2):
        H = da.Hii[ik][0]
        np.testing.assert_array_equal(H, db.Hii[ik][0])
        np.testing.assert_allclose(H, H.conj().T, atol=1e-15)
    # H_ii(kz) = H_ii + T_z e^{ikz} + h.c.
    Tz = a.Hr[:, :, 1, 0, 0]
    H0 = a.Hr[:, :, 1, 0, 1]
    np.testing.assert_allclose(da.Hii[1][0], H0 + Tz * np.exp(0.9j) + Tz.conj().T * np.exp(-0.9j),
                               atol=1e-14)
    # the coupling is H_{i,i-1} = Hr(x=+1) = Hr(x=-1)^H, the same on every slab and lead
    np.testing.assert_array_equal(da.H1i[0][0], a.Hr[:, :, 2, 0, 1])
    np.testing.assert_allclose(a.Hr[:, :, 2, 0, 1], a.Hr[:, :, 0, 0, 1].conj().T)
    assert len(da.H1i[0]) == 4
    # about 8 in-block neighbours per site, hopping around -2.7 eV
    off = H0 - np.diag(np.diag(H0))
    deg = (np.abs(off) > 0).sum(1)
    assert 5 <= deg.mean() <= 9
    assert abs(np.real(off[np.abs(off) > 0]).mean() + 2.7) < 0.2


def test_ham_dat_round_trip(tmp_path):
    w = hm.synthetic_wannier(10, 3, tz_density=0.2)
    hm.write_ham_dat(tmp_path / "ham_dat", w)
    r = hm.read_ham_dat(str(tmp_path / "ham_dat"))
    np.testing.assert_array_equal(r.Hr, w.Hr)
    assert (r.xmin, r.ymin, r.zmin) == (w.xmin, w.ymin, w.zmin)
    np.testing.assert_array_equal(r.centres, w.centres)


def test_mat_def_blocks_of_ns_cells(golden_dir):
    """ns cells per slab: block (i,k) of Hii is Hr(x=i-k), of H1i Hr(x=i-k+ns)."""
    w = hm.read_ham_dat(f"{golden_dir}/ham_dat")
    Hii, H1i = w.mat_def(0.0, 0.0, 0.0, 3)
    nb = w.nb
    x0 = -w.xmin
    for i in range(3):
        for k in range(3):
            np.testing.assert_array_equal(Hii[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb],
                                          w.Hr[:, :, x0 + i - k, 0, 0])
            np.testing.assert_array_equal(H1i[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb],
                                          w.Hr[:, :, x0 + i - k + 3, 0, 0])
    np.testing.assert_allclose(Hii, Hii.conj().T, atol=1e-12)


def test_energy_grid():
    E = hm.energy_grid(-6.0, -1.0, 200)
    assert E[0] == -6.0 and abs(E[-1] + 1.0) < 1e-14 and len(E) == 200


def _dense_from_device(dev):
    n = sum(dev.nm)
    H = np.zeros((n, n), np.complex128)
    off = np.cumsum([0] + list(dev.nm))
    for s in range(dev.nx):
        H[off[s]:off[s + 1], off[s]:off[s + 1]] = dev.Hii[0][s]
        if s > 0:  # H1i[s] = H_{s,s-1}
            H[off[s]:off[s + 1], off[s - 1]:off[s]] = dev.H1i[0][s]
            H[off[s - 1]:off[s], off[s]:off[s + 1]] = dev.H1i[0][s].conj().T
    return H


def test_coo_route_rebuilds_the_device(tmp_path):
    """devH_build_fromCOOfile (deviceHam_mod.f90:91-219) with blocksize: the
    upper triangle of a block-tridiagonal Hermitian matrix written as a COO
    file (0-based, complex) comes back as the same blocks, H1i[s] = H_{s,s-1},
    the lead couplings repeating the outer interior couplings"""
    _, dev = hm.synthetic_device(6, 5, seed=11, coupling_density=0.4)
    H = _dense_from_device(dev)
    assert np.allclose(H, H.conj().T)
    path = tmp_path / "h_dat"
    hm.write_coo(str(path), H, use0index=True, iscomplex=True)
    got = hm.device_from_coo(str(path), 6, use0index=True, iscomplex=True)
    assert got.nm == [6] * 5 and got.nk == 1 and got.Sii is None
    for s in range(5):
        assert np.array_equal(got.Hii[0][s], dev.Hii[0][s])
    for s in range(1, 5):
        assert np.array_equal(got.H1i[0][s], dev.H1i[0][s])
    assert np.array_equal(got.H1i[0][0], got.H1i[0][1])
    assert np.array_equal(got.H1i[0][5], got.H1i[0][4])
    assert np.array_equal(_dense_from_device(got), H)


def test_coo_route_real_one_based_and_errors(tmp_path):
    """1-based real files (iscomplex=False keeps the real parts, as
    main.f90:113-114 reads h_dat); the last row of each slab stays in its slab;
    orders that are not a multiple of the block size and couplings beyond the
    next slab are rejected"""
    rng = np.random.default_rng(2)
    n, bs = 12, 4
    H = np.zeros((n, n))
    for i in range(n):
        for j in range(i, min(n, (i // bs + 2) * bs)):
            H[i, j] = H[j, i] = rng.standard_normal()
    path = tmp_path / "h1"
    hm.write_coo(str(path), H, use0index=False, iscomplex=False)
    r, c, v, norb = hm.read_coo(str(path), use0index=False)
    assert norb == n and r.min() == 1
    dev = hm.device_from_coo(str(path), bs, use0index=False)
    assert np.array_equal(_dense_from_device(dev).real, H)
    assert dev.Hii[0][0][bs - 1, bs - 1] == H[bs - 1, bs - 1]  # row bs is in slab 1
    with pytest.raises(ValueError):
        hm.device_from_coo(str(path), 5, use0index=False)
    H[0, n - 1] = H[n - 1, 0] = 1.0
    hm.write_coo(str(path), H, use0index=False)
    with pytest.raises(ValueError):
        hm.device_from_coo(str(path), bs, use0index=False)


def test_band_edges_two_band_chain_closed_form():
    """band_edges (wannierHam3d.f90:153-201) on a two-band chain along x with
    complex hoppings: E_b(k) = h_b + 2 Re(t_b e^{-i k Lx}), evaluated on the
    reference's grid k_i = i dk - pi/Lx, i = 1..nkx, nkx = floor(1/Lx/0.01) + 1.
    The sine term fixes the sign of the phase; one cell in y and z gives one
    k point there and zero transverse edge positions."""
    Lx = 4.0
    cell = np.diag([Lx, 3.0, 2.0])
    h, t = np.array([-1.0, 1.0]), np.array([0.2 + 0.3j, -0.25 + 0.1j])
    Hr = np.zeros((2, 2, 3, 1, 1), complex)
    Hr[:, :, 1, 0, 0] = np.diag(h)
    Hr[:, :, 2, 0, 0] = np.diag(t)
    Hr[:, :, 0, 0, 0] = np.diag(t.conj())
    w = hm.Wannier(nvb=1, spin_deg=2, cell=cell, Hr=Hr, xmin=-1, ymin=0, zmin=0,
                   centres=np.zeros((3, 2)))
    r = hm.band_edges(w)
    nkx = int(np.floor(1 / Lx / 0.01)) + 1
    assert r["nk"] == (nkx, 1, 1)
    k = np.arange(1, nkx + 1) * (2 * np.pi / Lx / (nkx - 1)) - np.pi / Lx
    E = h[:, None] + 2 * np.real(t[:, None] * np.exp(-1j * k[None, :] * Lx))
    np.testing.assert_allclose(r["EN"], E, atol=1e-13)
    assert r["VBM"] == pytest.approx(E[0].max(), abs=1e-13)
    assert r["CBM"] == pytest.approx(E[1].min(), abs=1e-13)
    assert r["Eg"] == pytest.approx(E[1].min() - E[0].max(), abs=1e-13)
    assert r["kt_cbm"] == 0.0 and r["kt_vbm"] == 0.0


def test_band_edges_golden_cnt_gap_is_positive():
    """The reference's CNT input (test_cnt/ham_dat): a semiconducting tube, so the
    scan finds CBM above VBM on the 24-point axial grid."""
    import os
    w = hm.read_ham_dat(os.path.join(os.path.dirname(__file__), "golden", "ham_dat"))
    r = hm.band_edges(w)
    assert r["nk"] == (24, 1, 1)
    assert 0.1 < r["Eg"] < 2.0
    assert np.all(np.diff(r["EN"], axis=0) >= -1e-12)


def test_bare_coulomb_closed_form():
    """Two orbitals 3 A apart along x in a cubic 6 A cell: V between them,
    their self-interaction, the next cell's pair at 3 A and 9 A (the
    coupling block), against e / (4 pi eps0 eps r) tanh(r / r0) by hand."""
    from quantum_transport_amd.hamiltonian import Wannier

    cell = np.diag([6.0, 6.0, 6.0])
    Hr = np.zeros((2, 2, 3, 1, 1), complex)
    centres = np.array([[0.0, 3.0], [0.0, 0.0], [0.0, 0.0]])
    w = Wannier(2, 1, cell, Hr, -1, 0, 0, centres)
    eps, r0 = 4.0, 0.3
    k = 1.6e-19 / (4 * 3.14159265359 * 8.85e-12 * eps * 1e-10)  # eV A
    Vii, V1i = w.coulomb_blocks(0.0, 0.0, 0.0, eps, r0)
    assert np.allclose(Vii, [[k / r0, k / 3 * np.tanh(10)], [k / 3 * np.tanh(10), k / r0]], rtol=1e-13)
    # V(I+1, I): orbital i in cell +1 (x + 6 A) against orbital j in cell 0
    want = np.array([[k / 6 * np.tanh(20), k / 3 * np.tanh(10)], [k / 9 * np.tanh(30), k / 6 * np.tanh(20)]])
    assert np.allclose(V1i, want, rtol=1e-13)
    assert np.allclose(w.bare_coulomb(0, 0, 0, eps, r0, False).diagonal(), 0.0)
