"""Two GW-SCBA iterations end to end (the loop of gw_ephoton_3d_ijs,
gw_mod.f90:372-749): green_RGF_RS per energy with the current self-energy,
polarisation P, the W solve, self-energy Sigma and linear mixing, chained on
the device and, separately, on the oracle; the final G and Sigma of the two
chains agree.  Each stage is pinned on its own in test_gpu_gw.py,
test_gpu_gw_assembly.py and test_gpu_wsolve.py; this checks the hand-offs
(element-major layouts, energy grids, the Sigma that re-enters the RGF)."""
import numpy as np
import pytest

from quantum_transport_amd import hamiltonian as hm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf

    rgf.init(0)
    return rgf


def _system(nm, nx, seed):
    rng = np.random.default_rng(seed)
    _, dev = hm.synthetic_device(nm, nx, seed=seed, coupling_density=0.3)
    Hii = np.stack(dev.Hii[0], axis=2)
    H1i = np.stack(dev.H1i[0][:nx], axis=2)
    cr = lambda *sh, s=1.0: s * (rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    Vii = cr(nm, nm, nx, s=0.05)
    Vii = (Vii + Vii.conj().transpose(1, 0, 2)) / 2
    V1i = cr(nm, nm, nx, s=0.02)
    return Hii, H1i, Vii, V1i


def _em(blocks):
    """(nm, nm, nx, nen) -> element-major (nm*nm, nx, 1, nen), element i + j*nm"""
    nm, _, nx, nen = blocks.shape
    return blocks.reshape(nm * nm, nx, nen, order="F")[:, :, None, :].copy()


def _blocks(em, nm):
    """element-major (nm*nm, nx, 1, nen) -> (nm, nm, nx, nen)"""
    _, nx, _, nen = em.shape
    return em[:, :, 0, :].reshape(nm, nm, nx, nen, order="F")


def _scba(backend, kind, Hii, H1i, Vii, V1i, E, iters, ndiag, alpha):
    nm, _, nx = Hii.shape
    nen = len(E)
    dEn = E[1] - E[0]
    mu, T = [0.05, -0.05], [300.0, 300.0]
    S = [np.zeros((nm * nm, nx, 1, nen), complex) for _ in range(3)]  # R, <, >
    for _ in range(iters):
        SR, SL = _blocks(S[0], nm), _blocks(S[1], nm)
        G = [np.zeros((nm, nm, nx, nen), complex) for _ in range(3)]
        Cur = np.zeros((nm, nm, nx, nen))
        for e in range(nen):
            sr = np.asfortranarray(SR[..., e])
            sl = np.asfortranarray(SL[..., e])
            if kind == "device":
                out = backend.green_rgf_rs(T, nm, nx, E[e], mu, Hii, H1i, sigma_lesser_ph=sl,
                                           sigma_greater_ph=None, sigma_r_ph=sr)
            else:
                out = backend.green_rgf_rs(Hii, H1i, E[e], mu, T, sigma_lesser_ph=sl, sigma_r_ph=sr)
            for i, key in enumerate(("ldos", "ndens", "pdens")):
                G[i][..., e] = out[key]
            Cur[..., e] = out["cur"]
        Gr, Gl, Gg = (_em(g) for g in G)
        P = backend.gw_polarization(Gr, Gl, Gg, 1, 1, ndiag, dEn, 2.0)
        PR, PL, PG = (_blocks(p, nm) for p in P)
        if kind == "device":
            W = backend.gw_wsolve(Vii, V1i, PR, PL, PG)
        else:
            W = [np.zeros((nm, nm, nx, nen), complex) for _ in range(3)]
            for e in range(nen):
                for i, w in enumerate(backend.gw_wsolve(Vii, V1i, PR[..., e], PL[..., e], PG[..., e])):
                    W[i][..., e] = w
        Wr, Wl, Wg = (_em(np.asarray(w)) for w in W)
        Snew = backend.gw_selfenergy(Gr, Gl, Gg, Wr, Wl, Wg, 1, 1, ndiag, dEn)
        S = [np.asarray(backend.gw_mix(s.ravel(), n.ravel(), alpha)).reshape(s.shape)
             for s, n in zip(S, Snew)]
    return Gr, Gl, Gg, S, Cur


def test_two_scba_iterations_match_oracle(qt, oracle_mod):
    """Bar: 1e-9 of each quantity's largest value, widened to 20x the
    disagreement of the oracle's two LAPACK flavours run through the same
    chain (G^< at these real device energies moves by ~1e-7 between two
    correct FP64 operation orders)."""
    import functools
    import types

    nm, nx, nen, ndiag = 8, 4, 64, 2
    Hii, H1i, Vii, V1i = _system(nm, nx, seed=81)
    E = hm.energy_grid(-2.0, 2.0, nen)
    dev = _scba(qt, "device", Hii, H1i, Vii, V1i, E, 2, ndiag, 0.5)
    ref = _scba(oracle_mod, "oracle", Hii, H1i, Vii, V1i, E, 2, ndiag, 0.5)
    alt = None
    if oracle_mod.available("mkl"):
        mk = types.SimpleNamespace(**{n: functools.partial(getattr(oracle_mod, n), flavour="mkl") for n in
                                      ("green_rgf_rs", "gw_polarization", "gw_selfenergy", "gw_mix",
                                       "gw_wsolve")})
        alt = _scba(mk, "oracle", Hii, H1i, Vii, V1i, E, 2, ndiag, 0.5)
    names = ("G^R", "G^<", "G^>", "Sigma")
    pack = lambda r: r[:3] + (np.stack(r[3]),)
    for i, name in enumerate(names):
        a, b = pack(dev)[i], pack(ref)[i]
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / scale
        bar = 1e-9 if alt is None else max(1e-9, 20 * np.abs(pack(alt)[i] - b).max() / scale)
        print(f"scba {name}: max |dev - oracle| / max = {err:.2e} (bar {bar:.1e})")
        assert scale > 0 and err < bar, (name, err, bar)
    # the self-energy is not trivial: it changed G between the iterations
    assert np.abs(np.stack(ref[3])).max() > 1e-6


@pytest.mark.parametrize("nm,nx,nen,ndiag", [(8, 4, 64, 2), (16, 5, 128, 3)])
def test_device_scba_step_equals_host_chain(qt, nm, nx, nen, ndiag):
    """qt_gw_scba_step (G of all energies in batches with Sigma gathered from
    the element-major arrays and G scattered back, P, W, Sigma, mixing, all
    device resident) against the same two iterations chained from the host
    one energy at a time: the stages are the same kernels, so the results
    agree to rounding (the batch composition does not change a point's
    result)."""
    Hii, H1i, Vii, V1i = _system(nm, nx, seed=82 + nm)
    E = hm.energy_grid(-2.0, 2.0, nen)
    Gr, Gl, Gg, S, Cur = _scba(qt, "device", Hii, H1i, Vii, V1i, E, 2, ndiag, 0.5)
    hii = [Hii[:, :, s] for s in range(nx)]
    h1i = [H1i[:, :, s] for s in range(nx)] + [H1i[:, :, nx - 1]]  # right lead: H1i(nx)
    H = qt.Hamiltonian([hii], [h1i])
    st = qt.GWState(nm, nx, nen, Vii, V1i)
    cur = np.zeros((nm, nm, nx, nen), order="F")
    try:
        for it in range(2):
            print("device scba step", it, flush=True)
            tr, tre, status = qt.gw_scba_step(H, st, E, [0.05, -0.05], [300.0, 300.0], ndiag, 2.0, 0.5,
                                              cur=cur if it == 1 else None)
            assert status == 0
        G = st.host("G")
        Sd = st.host("S")
    finally:
        st.free()
        H.close()
    for name, a, b in (("G^R", G[0], Gr), ("G^<", G[1], Gl), ("G^>", G[2], Gg),
                       ("Sigma^R", Sd[0], S[0]), ("Sigma^<", Sd[1], S[1]), ("Sigma^>", Sd[2], S[2])):
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"scba step {name}: max |device loop - host chain| / max = {err:.1e}")
        assert err <= 1e-12, (name, err)
    # the slab-bond current matrices of the second iteration's G pass
    err = np.abs(cur - Cur).max() / np.abs(Cur).max()
    print(f"scba step cur: max |device loop - host chain| / max = {err:.1e}")
    assert np.abs(Cur).max() > 0 and err <= 1e-12
