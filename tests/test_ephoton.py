"""e-photon self-energies in the monochromatic assumption
(calc_sigma_ephoton_monochromatic_ext, gw_mod.f90:1075-1122).

The oracle is pinned on the CPU by closed forms that a plausible mistake
breaks: M = I (the energy shift: G at ie and ie + 2 nop), a single-entry M
(operand order: M A M, not its transpose), per-slab scalar M (the slab index
of M), the prefactor pre_fact intensity / hw^2, and the anti-Hermitian
structure for Hermitian M and anti-Hermitian G^<>.  The GPU path is then
checked against the oracle."""
import numpy as np
import pytest


def _rand(rng, *sh):
    return rng.standard_normal(sh) + 1j * rng.standard_normal(sh)


def _inputs(rng, nm, length, nen, nop):
    M = _rand(rng, nm, nm, length)
    Gl = _rand(rng, nm, nm, length, nen + 2 * nop)
    Gg = _rand(rng, nm, nm, length, nen + 2 * nop)
    return M, Gl, Gg


@pytest.mark.parametrize("nop", [0, 1, 3])
def test_oracle_identity_coupling_is_shifted_sum(oracle_mod, nop):
    rng = np.random.default_rng(1 + nop)
    nm, length, nen = 3, 2, 4
    _, Gl, Gg = _inputs(rng, nm, length, nen, nop)
    M = np.repeat(np.eye(nm, dtype=complex)[:, :, None], length, axis=2)
    pre, inten, hw = 0.7, 2.5, 0.4
    f = pre * inten / hw**2
    Sr, Sl, Sg = oracle_mod.gw_sigma_ephoton(M, Gl, Gg, nop, pre, inten, hw)
    for ie in range(nen):
        np.testing.assert_allclose(Sl[..., ie], f * (Gl[..., ie] + Gl[..., ie + 2 * nop]), rtol=1e-14)
        np.testing.assert_allclose(Sg[..., ie], f * (Gg[..., ie] + Gg[..., ie + 2 * nop]), rtol=1e-14)
    assert np.all(Sr.real == 0.0)
    np.testing.assert_allclose(Sr.imag, (Sg - Sl).imag / 2, rtol=1e-14, atol=1e-14)


def test_oracle_single_entry_coupling_fixes_operand_order(oracle_mod):
    # M = e_0 e_1^T: M A M = A[1, 0] e_0 e_1^T; a transposed operand moves it
    rng = np.random.default_rng(7)
    nm, length, nen, nop = 3, 1, 2, 1
    _, Gl, Gg = _inputs(rng, nm, length, nen, nop)
    M = np.zeros((nm, nm, length), complex)
    M[0, 1, 0] = 1.0
    Sr, Sl, Sg = oracle_mod.gw_sigma_ephoton(M, Gl, Gg, nop, 1.0, 1.0, 1.0)
    for ie in range(nen):
        A = Gl[..., 0, ie] + Gl[..., 0, ie + 2]
        want = np.zeros((nm, nm), complex)
        want[0, 1] = A[1, 0]
        np.testing.assert_allclose(Sl[..., 0, ie], want, rtol=1e-15, atol=0)


def test_oracle_slab_coupling_and_prefactor(oracle_mod):
    # M_ix = c_ix I: Sigma(ix) = c_ix^2 f (G(ie) + G(ie + 2 nop)); doubling
    # intensity or halving hw scales by 2 and 4
    rng = np.random.default_rng(11)
    nm, length, nen, nop = 2, 4, 3, 2
    _, Gl, Gg = _inputs(rng, nm, length, nen, nop)
    c = np.array([1.0, -2.0, 0.5j, 3.0 + 1.0j])
    M = np.eye(nm, dtype=complex)[:, :, None] * c[None, None, :]
    _, Sl, _ = oracle_mod.gw_sigma_ephoton(M, Gl, Gg, nop, 1.0, 1.0, 1.0)
    for ix in range(length):
        for ie in range(nen):
            np.testing.assert_allclose(Sl[:, :, ix, ie],
                                       c[ix] ** 2 * (Gl[:, :, ix, ie] + Gl[:, :, ix, ie + 4]), rtol=1e-14)
    _, Sl2, _ = oracle_mod.gw_sigma_ephoton(M, Gl, Gg, nop, 1.0, 2.0, 1.0)
    _, Sl4, _ = oracle_mod.gw_sigma_ephoton(M, Gl, Gg, nop, 1.0, 1.0, 0.5)
    np.testing.assert_allclose(Sl2, 2 * Sl, rtol=1e-14)
    np.testing.assert_allclose(Sl4, 4 * Sl, rtol=1e-14)


def test_oracle_antihermitian_and_equal_inputs(oracle_mod):
    # Hermitian M, anti-Hermitian G^<>: Sigma^<> = M A M stays anti-Hermitian;
    # G^< = G^> gives Sigma^R = 0
    rng = np.random.default_rng(5)
    nm, length, nen, nop = 4, 2, 3, 1
    M, Gl, _ = _inputs(rng, nm, length, nen, nop)
    M = (M + M.conj().transpose(1, 0, 2)) / 2
    Gl = (Gl - Gl.conj().transpose(1, 0, 2, 3)) / 2
    Sr, Sl, Sg = oracle_mod.gw_sigma_ephoton(M, Gl, Gl, nop, 0.3, 1.7, 0.2)
    np.testing.assert_allclose(Sl, -Sl.conj().transpose(1, 0, 2, 3), atol=1e-12 * np.abs(Sl).max())
    np.testing.assert_array_equal(Sl, Sg)
    assert np.abs(Sr).max() == 0.0


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf
    rgf.init(0)
    return rgf


@pytest.mark.gpu
@pytest.mark.parametrize("nm,length,nen,nop", [(1, 1, 1, 0), (3, 2, 5, 1), (8, 3, 7, 2),
                                               (64, 4, 6, 3), (72, 2, 9, 4), (130, 3, 3, 1)])
def test_gpu_ephoton_matches_oracle(qt, oracle_mod, nm, length, nen, nop):
    print("ephoton", nm, length, nen, nop, flush=True)
    rng = np.random.default_rng(nm * 1000 + length * 10 + nop)
    M, Gl, Gg = _inputs(rng, nm, length, nen, nop)
    args = (nop, 0.8, 1.3, 0.25)
    got = qt.gw_sigma_ephoton(M, Gl, Gg, *args)
    ref = oracle_mod.gw_sigma_ephoton(M, Gl, Gg, *args)
    for g, r, name in zip(got, ref, "R<>"):
        scale = np.abs(r).max()
        assert np.abs(g - r).max() <= 1e-12 * scale, name
    assert np.all(got[0].real == 0.0)


@pytest.mark.gpu
def test_gpu_ephoton_device_pointers(qt, oracle_mod):
    """QT_F_DEVICE_OUTPUTS: every array resident on the device (the bench's
    route), two calls in a row with the second smaller, against the oracle."""
    import ctypes as C
    for nm, length, nen, nop in ((20, 3, 6, 2), (9, 2, 3, 1)):
        print("ephoton device", nm, length, nen, nop, flush=True)
        rng = np.random.default_rng(nm)
        M, Gl, Gg = (np.asfortranarray(a) for a in _inputs(rng, nm, length, nen, nop))
        ins = []
        for a in (M, Gl, Gg):
            b = qt.DeviceBuffer(a.nbytes)
            qt.check(qt.lib().qt_memcpy_h2d(b.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes)))
            ins.append(b)
        nbytes = 16 * nm * nm * length * nen
        outs = [qt.DeviceBuffer(nbytes) for _ in range(3)]
        qt.check(qt.lib().qt_gw_sigma_ephoton(nm, length, nen, nop, 0.5, 2.0, 0.7,
                                              *(b.ptr for b in ins), *(b.ptr for b in outs),
                                              qt.QT_F_DEVICE_OUTPUTS))
        got = [b.to_host(np.complex128, (nen, length, nm, nm)).transpose(3, 2, 1, 0) for b in outs]
        ref = oracle_mod.gw_sigma_ephoton(M, Gl, Gg, nop, 0.5, 2.0, 0.7)
        for g, r in zip(got, ref):
            assert np.abs(g - r).max() <= 1e-12 * np.abs(r).max()
        for b in ins + outs:
            b.free()


@pytest.mark.gpu
@pytest.mark.parametrize("budget", ["1", "420000", "3000000"])
def test_gpu_ephoton_energy_chunks(qt, oracle_mod, monkeypatch, budget):
    """The working set capped through QT_WORKSET_BYTES: 1 byte gives one
    energy per chunk, 420 kB about 2 energies on the host path (the last
    chunk shorter), 3 MB the whole axis.  Host and device paths give the
    same bits as the uncapped call and match the oracle."""
    import ctypes as C
    nm, length, nen, nop = 20, 3, 7, 2
    rng = np.random.default_rng(77)
    M, Gl, Gg = (np.asfortranarray(a) for a in _inputs(rng, nm, length, nen, nop))
    args = (nop, 0.6, 1.7, 0.45)
    free = qt.gw_sigma_ephoton(M, Gl, Gg, *args)
    monkeypatch.setenv("QT_WORKSET_BYTES", budget)
    print("ephoton chunked, budget", budget, flush=True)
    host = qt.gw_sigma_ephoton(M, Gl, Gg, *args)
    ins = []
    for a in (M, Gl, Gg):
        b = qt.DeviceBuffer(a.nbytes)
        qt.check(qt.lib().qt_memcpy_h2d(b.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes)))
        ins.append(b)
    outs = [qt.DeviceBuffer(16 * nm * nm * length * nen) for _ in range(3)]
    qt.check(qt.lib().qt_gw_sigma_ephoton(nm, length, nen, nop, *args[1:], *(b.ptr for b in ins),
                                          *(b.ptr for b in outs), qt.QT_F_DEVICE_OUTPUTS))
    dev = [b.to_host(np.complex128, (nen, length, nm, nm)).transpose(3, 2, 1, 0) for b in outs]
    for b in ins + outs:
        b.free()
    ref = oracle_mod.gw_sigma_ephoton(M, Gl, Gg, *args)
    for f, h, d, r in zip(free, host, dev, ref):
        assert np.array_equal(f, h) and np.array_equal(f, d)
        assert np.abs(h - r).max() <= 1e-12 * np.abs(r).max()


@pytest.mark.gpu
def test_gpu_ephoton_rejects_bad_arguments(qt):
    rng = np.random.default_rng(3)
    M, Gl, Gg = _inputs(rng, 2, 1, 2, 1)
    with pytest.raises(qt.QTError):
        qt.gw_sigma_ephoton(M, Gl, Gg, 1, 1.0, 1.0, 0.0)  # hw = 0
    with pytest.raises(ValueError):
        qt.gw_sigma_ephoton(M, Gl, Gg, 3, 1.0, 1.0, 1.0)  # axis shorter than 2 nop
    # no output energies (the axis holds only the 2 nop shifted points): empty results
    Sr, Sl, Sg = qt.gw_sigma_ephoton(M, Gl[..., :2], Gg[..., :2], 1, 1.0, 1.0, 1.0)
    assert Sr.shape == Sl.shape == Sg.shape == (2, 2, 1, 0)


def test_sigma_checkpoint_round_trip(tmp_path):
    """SCBA self-energy checkpoint: bit-identical round trip, memory order and
    iteration kept, a later save replaces the earlier one, shapes checked."""
    from quantum_transport_amd import rgf
    rng = np.random.default_rng(2)
    S = [np.asfortranarray(_rand(rng, 3, 3, 2, 5)) for _ in range(3)]
    path = str(tmp_path / "sigma.npz")
    rgf.save_sigma_checkpoint(path, 4, *S)
    it, *got = rgf.load_sigma_checkpoint(path)
    assert it == 4
    for g, s in zip(got, S):
        np.testing.assert_array_equal(g, s)
        assert g.flags.f_contiguous
    rgf.save_sigma_checkpoint(path, 5, *(2 * s for s in S))
    it, *got = rgf.load_sigma_checkpoint(path)
    assert it == 5
    np.testing.assert_array_equal(got[1], 2 * S[1])
    assert not (tmp_path / "sigma.npz.tmp").exists()
    with pytest.raises(ValueError):
        rgf.save_sigma_checkpoint(path, 6, S[0], S[1], S[2][..., :2])
