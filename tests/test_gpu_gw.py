"""green_RGF_RS (gw_mod.f90:1131-1345) on the device against the oracle's
restatement: every output block, with scattering self-energies."""
import numpy as np
import pytest

from quantum_transport_amd import hamiltonian as hm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf

    rgf.init(0)
    return rgf


def _inputs(n, nx, seed):
    rng = np.random.default_rng(seed)
    _, dev = hm.synthetic_device(n, nx, seed=seed, coupling_density=0.2)
    Hii = np.stack(dev.Hii[0], axis=2)
    H1i = np.stack(dev.H1i[0][:nx], axis=2)
    sr = np.zeros((n, n, nx), complex)
    sl = np.zeros((n, n, nx), complex)
    for s in range(nx):
        a = 0.03 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        sr[:, :, s] = (a + a.T) / 2 - 0.04j * np.eye(n)
        b = 0.05 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        sl[:, :, s] = 1j * (b @ b.conj().T)
    return Hii, H1i, sr, sl


@pytest.mark.parametrize("n,nx,E", [(8, 5, -0.2), (8, 5, 0.45), (64, 6, 0.1), (96, 4, -1.3)])
def test_green_rgf_rs_parity(qt, oracle_mod, n, nx, E):
    Hii, H1i, sr, sl = _inputs(n, nx, seed=n + nx)
    mu, T = [0.05, -0.15], [300.0, 280.0]
    ref = oracle_mod.green_rgf_rs(Hii, H1i, E, mu, T, sigma_lesser_ph=sl, sigma_r_ph=sr)
    got = qt.green_rgf_rs(T, n, nx, E, mu, Hii, H1i, sigma_lesser_ph=sl, sigma_greater_ph=None,
                          sigma_r_ph=sr, offdiag=True)
    # the oracle in its other LAPACK flavour measures how sensitive each
    # output is to the operation order (the bar is widened to 20x that)
    alt = (oracle_mod.green_rgf_rs(Hii, H1i, E, mu, T, sigma_lesser_ph=sl, sigma_r_ph=sr, flavour="mkl")
           if oracle_mod.available("mkl") else None)
    for key in ("ldos", "ndens", "pdens", "cur", "GRi1", "GLi1", "GGi1"):
        a, b = got[key], ref[key]
        c = alt[key] if alt is not None else None
        if key in ("GRi1", "GLi1", "GGi1", "cur"):
            a, b = a[:, :, :-1], b[:, :, :-1]  # defined for l = 1..nx-1
            c = c[:, :, :-1] if c is not None else None
        scale = max(np.abs(ref["ldos"]).max(), np.abs(b).max())
        err = np.abs(a - b).max() / scale
        bar = 1e-10 if c is None else max(1e-10, 20 * np.abs(c - b).max() / scale)
        print(f"n={n} {key}: {err:.2e} (bar {bar:.1e})")
        assert err < bar, (key, err)
    for key in ("tr", "tre"):
        assert abs(got[key] - ref[key]) <= 1e-9 * max(1.0, abs(ref[key])), key
