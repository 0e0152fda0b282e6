"""The device path at the shapes of the benchmark configurations that the
other GPU tests reach only in reduced form:

* uneven slabs drawn from {128, 192, ..., 768} (a short device with both
  extreme sizes against the oracle; the 80-slab device through
  size-independent properties, and the mixed mode against FP64 on it);
* the lead at nb = 256, eta = 1e-6, including energies at subband edges
  (Sancho-Rubio against the oracle with equal iteration counts; the
  surface-equation residual over a strided sample of the 4096 x 16 grid);
* the GW energy-axis assembly at nm = 256 and 2048 energies (P and Sigma of
  sampled elements against the oracle).
"""
import numpy as np
import pytest

from quantum_transport_amd import hamiltonian as hm
from tests.test_gpu_rgf import check_parity, spectra

pytestmark = pytest.mark.gpu

BOLTZ = 8.61734e-05


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf

    rgf.init(0)
    return rgf


# ----------------------------------------------------------------- uneven slabs


def test_uneven_extreme_sizes_against_oracle(qt, oracle_mod):
    """6 slabs including 128 and 768 orbitals, with jumps both ways
    (192 -> 768 -> 320 -> 128 -> 576), coupling H_{i,i+1} of the slab sizes
    in the second pass (QT_F_COUPLING_FIX)."""
    if not oracle_mod.available("mkl"):
        pytest.skip("needs the vendor-LAPACK oracle flavour")
    sizes = [192, 768, 320, 128, 576, 256]
    dev = hm.uneven_device(sizes, seed=3003)
    E = np.linspace(-1.8, 1.8, 6)
    H = qt.Hamiltonian.from_device(dev)
    gpu = qt.solve_batch(H, E, None, mul=0.0, mur=-0.1, coupling_fix=True)
    H.close()
    mkl = oracle_mod.rgf_batch(dev, E, None, 0.0, -0.1, coupling_fix=True, flavour="mkl")
    plain = oracle_mod.rgf_batch(dev, E, None, 0.0, -0.1, coupling_fix=True, nthreads=16)
    check_parity(gpu, mkl, plain, sizes, "uneven 6 slabs")


def test_uneven_80_slabs_properties_and_mixed(qt):
    """80 slabs of 128..768 orbitals (about 36k in all): the current leaving
    one contact enters the other, occupations lie inside the contacts' Fermi
    window, results do not depend on the batch they are solved in; then the
    mixed mode against FP64 on the same points, per observable."""
    rng = np.random.default_rng(0)
    sizes = [int(v) for v in rng.choice(np.arange(128, 769, 64), size=80)]
    dev = hm.uneven_device(sizes, 0)
    E = np.array([-1.4, -0.6, -0.05, 0.25, 0.9, 1.6])
    mul, mur, T = 0.0, -0.1, 300.0
    H = qt.Hamiltonian.from_device(dev)
    r = qt.solve_batch(H, E, None, mul=mul, mur=mur, coupling_fix=True)
    r2 = qt.solve_batch(H, E[::-1][:3], None, mul=mul, mur=mur, coupling_fix=True)
    mix = qt.solve_batch(H, E, None, mul=mul, mur=mur, coupling_fix=True, precision=qt.QT_PREC_MIXED)
    H.close()
    print(f"80 slabs, {sum(sizes)} orbitals; lead iterations {r['lead_iters'].tolist()}")
    assert (r["info"] == 0).all()
    for key in ("gr_diag", "gl_diag", "gg_diag", "tr", "tre"):
        assert np.array_equal(r[key][::-1][:3], r2[key]), key
    scale = np.abs(r["tr"]).max()
    assert np.abs(r["tr"] + r["tre"]).max() <= 1e-8 * max(scale, 1e-300)
    fa = 1 / (1 + np.exp((E - mul) / (BOLTZ * T)))
    fb = 1 / (1 + np.exp((E - mur) / (BOLTZ * T)))
    nocc, nemp = r["gl_diag"].imag, -r["gg_diag"].imag
    tot = nocc + nemp
    big = tot > 1e-8 * tot.max(axis=1, keepdims=True)
    occ = nocc / np.where(big, tot, 1.0)
    lo = np.broadcast_to(np.minimum(fa, fb)[:, None], occ.shape)
    hi = np.broadcast_to(np.maximum(fa, fb)[:, None], occ.shape)
    assert (occ[big] >= lo[big] - 1e-7).all() and (occ[big] <= hi[big] + 1e-7).all()
    sf, sm = spectra(r, sizes), spectra(mix, sizes)
    errs = {k: float(np.abs(sm[k] - sf[k]).max() / max(np.abs(sf[k]).max(), 1e-300)) for k in sf}
    print("mixed vs FP64 on 80 uneven slabs, max error / max |value|:",
          {k: f"{v:.1e}" for k, v in errs.items()})
    assert all(np.isfinite(v) and v < 1e-3 for v in errs.values())


# ----------------------------------------------------------------- lead at eta = 1e-6


def _lead256():
    kz = np.arange(16) * np.pi / 16
    w = hm.synthetic_wannier(256, 1003, tz_density=0.02)
    H00s, H10s = zip(*[w.mat_def(0.0, 0.0, q, 1) for q in kz])
    return kz, list(H00s), list(H10s)


def _subband_edges(H00, H10, lo, hi, count):
    """Extrema over kx of the lead's bulk bands E_n(kx) (eigenvalues of
    H00 + H10 e^{ikx} + H10^H e^{-ikx}) inside (lo, hi), `count` of them
    spread over that range."""
    ks = np.linspace(0, np.pi, 33)
    ev = np.array([np.linalg.eigvalsh(H00 + H10 * np.exp(1j * k) + H10.conj().T * np.exp(-1j * k))
                   for k in ks])
    edges = np.concatenate([ev.min(axis=0), ev.max(axis=0)])
    edges = np.sort(edges[(edges > lo) & (edges < hi)])
    return edges[np.linspace(0, len(edges) - 1, count).astype(int)]


def _surface_residual(H00, H10, E, eta, g):
    """|| (z - H00 - H10^H g H10)^-1 - g || / || (...)^-1 ||, z = E + i eta."""
    n = H00.shape[0]
    phi = np.linalg.inv((E + 1j * eta) * np.eye(n) - H00 - H10.conj().T @ g @ H10)
    return np.linalg.norm(phi - g) / np.linalg.norm(phi)


def test_lead256_against_oracle_with_band_edges(qt, oracle_mod):
    """nb = 256, eta = 1e-6: 10 energies of the grid and 6 subband edges,
    over several kz.  G00 against the oracle's decimation with the same
    iteration counts; where the two oracle flavours agree to 1e-11 the
    bar is 1e-10, elsewhere 8x their disagreement."""
    if not oracle_mod.available("mkl"):
        pytest.skip("needs the vendor-LAPACK oracle flavour")
    kz, H00s, H10s = _lead256()
    E = list(np.linspace(-2.7, 2.7, 10))
    k = list(np.arange(10) * 3 % 16)
    for ik in (0, 5, 11):
        for e in _subband_edges(H00s[ik], H10s[ik], -3.0, 3.0, 2):
            E.append(float(e))
            k.append(ik)
    E, k = np.array(E), np.array(k)
    G, _, it, cv = qt.sancho([H00s[i] for i in k], [H10s[i] for i in k], E, eta=1e-6, want_gbb=False)
    errs, spreads = [], []
    for p in range(len(E)):
        gm, _, itm, cvm = oracle_mod.sancho(None, H00s[k[p]], H10s[k[p]], E[p], eta=1e-6,
                                            want_gbb=False, flavour="mkl")
        gp, _, _, _ = oracle_mod.sancho(None, H00s[k[p]], H10s[k[p]], E[p], eta=1e-6, want_gbb=False)
        assert it[p] == itm and cv[p] == cvm, (p, it[p], itm)
        sc = np.abs(gm).max()
        errs.append(np.abs(G[p] - gm).max() / sc)
        spreads.append(np.abs(gp - gm).max() / sc)
    errs, spreads = np.array(errs), np.array(spreads)
    well = spreads <= 1e-11
    print(f"lead nb=256: G00 vs oracle max {errs.max():.1e} ({well.sum()} points where the flavours "
          f"agree: max {errs[well].max() if well.any() else 0:.1e}); flavour spread max {spreads.max():.1e}; "
          f"iterations {it.tolist()}")
    assert (errs[well] <= 1e-10).all()
    assert (errs[~well] <= 8.0 * spreads[~well]).all()


def test_lead256_surface_residual_strided_grid(qt, oracle_mod):
    """The decimation's G00 over 256 points strided through the 4096 x 16
    grid: retarded (-Im Tr G00 > 0) everywhere, and at every 8th point its
    surface-equation residual within 8x of the oracle decimation's own, with
    a median no more than twice the oracle's (the decimation at eta = 1e-6
    leaves residuals up to ~1e-6 at some points)."""
    if not oracle_mod.available("mkl"):
        pytest.skip("needs the vendor-LAPACK oracle flavour")
    kz, H00s, H10s = _lead256()
    E = hm.energy_grid(-3.0, 3.0, 4096)
    pts = (np.arange(256) * 2053) % (4096 * 16)
    Ep, kp = E[pts // 16], pts % 16
    G, _, it, cv = qt.sancho([H00s[i] for i in kp], [H10s[i] for i in kp], Ep, eta=1e-6, want_gbb=False)
    for g in G:
        assert -np.trace(g).imag > 0
    rg, ro = [], []
    for p in range(0, 256, 8):
        H00, H10 = H00s[kp[p]], H10s[kp[p]]
        go, _, ito, _ = oracle_mod.sancho(None, H00, H10, Ep[p], eta=1e-6, want_gbb=False, flavour="mkl")
        assert ito == it[p]
        rg.append(_surface_residual(H00, H10, Ep[p], 1e-6, G[p]))
        ro.append(_surface_residual(H00, H10, Ep[p], 1e-6, go))
    rg, ro = np.array(rg), np.array(ro)
    print(f"surface residual over 32 strided points: GPU median {np.median(rg):.1e} max {rg.max():.1e}; "
          f"oracle median {np.median(ro):.1e} max {ro.max():.1e}; iterations {it.min()}..{it.max()}, "
          f"converged {cv.sum()}/256")
    assert (rg <= np.maximum(1e-9, 8.0 * ro)).all()
    assert np.median(rg) <= 2.0 * max(np.median(ro), 1e-10)


# ----------------------------------------------------------------- GW assembly at nm = 256


def test_gw_assembly_nm256_nen2048_sampled(qt, oracle_mod):
    """P and Sigma (gw_mod.f90:444-743) at nm = 256, 2048 energies, one slab,
    one k point, ndiag = 16, on device arrays (2 GiB each, filled from the
    host in chunks).  With one k point every element ij of P and Sigma
    depends only on the series of ij and ji, so 8 in-band pairs (both
    orders) are checked against the oracle on the 2 x 2 problem of their two
    orbitals; 2 out-of-band pairs must be zero."""
    import ctypes as C

    nm, nen, ndiag = 256, 2048, 16
    dEn = 8.0 / (nen - 1)
    L = qt.lib()
    row_bytes = nen * 16
    nbytes = nm * nm * row_bytes
    rng = np.random.default_rng(56)
    pairs = []
    while len(pairs) < 8:
        i, j = (int(v) for v in rng.integers(0, nm, 2))
        if i != j and abs(i - j) <= ndiag:
            pairs.append((i, j))
    far = [(3, 3 + ndiag + 5), (200, 100)]
    # per pair the 2 x 2 problem's elements r + 2 c in order: ii, ji, ij, jj
    take = [x for i, j in pairs for x in (i + nm * i, j + nm * i, i + nm * j, j + nm * j)]
    take += [i + nm * j for i, j in far]
    bufs = []

    def filled():
        """device array of normal samples; returns it and the rows in `take`"""
        b = qt.DeviceBuffer(nbytes)
        bufs.append(b)
        rows = {}
        chunk = 4096
        for r0 in range(0, nm * nm, chunk):
            h = np.ascontiguousarray(rng.standard_normal((chunk, nen, 2)).view(np.complex128)[..., 0])
            qt.check(L.qt_memcpy_h2d(C.c_void_p(b.ptr.value + r0 * row_bytes), h.ctypes.data_as(C.c_void_p),
                                     C.c_size_t(h.nbytes)))
            for t in take:
                if r0 <= t < r0 + chunk:
                    rows[t] = h[t - r0].copy()
        return b, np.stack([rows[t] for t in take])

    def zeroed():
        b = qt.DeviceBuffer(nbytes)
        bufs.append(b)
        return b

    def gather(b):
        out = np.empty((len(take), nen), np.complex128)
        for q, t in enumerate(take):
            qt.check(L.qt_memcpy_d2h(out[q].ctypes.data_as(C.c_void_p), C.c_void_p(b.ptr.value + t * row_bytes),
                                     C.c_size_t(row_bytes)))
        return out

    def compare(name, got, ref_fn, inputs):
        worst = 0.0
        for q in range(len(pairs)):
            ref = ref_fn([a[4 * q:4 * q + 4][:, None, None, :] for a in inputs])
            for g, rr in zip(got, ref):
                for s in (1, 2):  # ji, ij
                    e = np.abs(g[4 * q + s] - rr[s, 0, 0]).max() / np.abs(rr[s, 0, 0]).max()
                    worst = max(worst, e)
        for g in got:
            assert not g[4 * len(pairs):].any()
        print(f"{name} at nm={nm}, nen={nen}: worst relative error over sampled elements {worst:.1e}")
        assert worst < 1e-11

    try:
        G, Gs = zip(*[filled() for _ in range(3)])
        P = [zeroed() for _ in range(3)]
        qt.check(L.qt_gw_polarization(nen, 1, 1, 1, nm, ndiag, dEn, 2.0, *(b.ptr for b in G),
                                      *(b.ptr for b in P), qt.QT_F_DEVICE_OUTPUTS))
        compare("P", [gather(b) for b in P],
                lambda sub: oracle_mod.gw_polarization(*sub, 1, 1, 1, dEn, 2.0), list(Gs))
        for b in P:
            b.free()
        W, Ws = zip(*[filled() for _ in range(3)])
        S = [zeroed() for _ in range(3)]
        qt.check(L.qt_gw_selfenergy(nen, 1, 1, 1, nm, ndiag, dEn, *(b.ptr for b in G),
                                    *(b.ptr for b in W), *(b.ptr for b in S), qt.QT_F_DEVICE_OUTPUTS))
        compare("Sigma", [gather(b) for b in S],
                lambda sub: oracle_mod.gw_selfenergy(*sub, 1, 1, 1, dEn), list(Gs) + list(Ws))
    finally:
        for b in bufs:
            b.free()
