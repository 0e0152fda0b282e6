"""QT_PREC_FP64_3M: complex products with three real matrix products,
    P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi),  Re = P1 - P2,
    Im = (P3 - P1) - P2,
on the FP64 matrix cores.  Rounding-error bound of a K-term product
(Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed., 23.2.4),
elementwise with |.| taken entrywise and u = 2^-53:
    |Re err| <= (K + 2) u (|Ar||Br| + |Ai||Bi|)
    |Im err| <= (K + 4) u (|Ar + Ai||Br + Bi| + |Ar||Br| + |Ai||Bi|)
(the conventional four-product form has |Ar||Bi| + |Ai||Br| in the second
line); the sweep in this mode against the FP64 one."""
import numpy as np
import pytest

from quantum_transport_amd import hamiltonian as hm

pytestmark = pytest.mark.gpu

OPS = [lambda x: x, lambda x: x.T, lambda x: x.conj().T]
U = 2.0 ** -53


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf

    rgf.init(0)
    return rgf


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("opa,opb", [(0, 0), (0, 2), (2, 0), (0, 1)])
@pytest.mark.parametrize("mnk", [(64, 64, 64), (130, 70, 200), (512, 512, 512)])
def test_3m_product_bound(qt, opa, opb, mnk):
    m, n, k = mnk
    rng = np.random.default_rng(m + 7 * n + 13 * k + opa * 3 + opb)
    batch = 2
    A = _rand(rng, batch, *((m, k) if opa == 0 else (k, m)))
    B = _rand(rng, batch, *((k, n) if opb == 0 else (n, k)))
    A[1] *= np.logspace(-6, 4, A.shape[1])[:, None]
    got = qt.zgemm_batched(A, B, opa, opb, precision=qt.QT_PREC_FP64_3M)
    four = qt.zgemm_batched(A, B, opa, opb)
    assert not np.array_equal(got, four)  # the three-product form ran
    for b in range(batch):
        oa, ob = OPS[opa](A[b]), OPS[opb](B[b])
        ar, ai, br, bi = np.abs(oa.real), np.abs(oa.imag), np.abs(ob.real), np.abs(ob.imag)
        # the reference is numpy's conventional complex128 product, itself
        # within (k + 2) u |A||B| of the exact one: that margin is added
        ref = oa @ ob
        conv = (k + 2) * U * (ar @ br + ai @ bi)
        im_b = (k + 4) * U * (np.abs(oa.real + oa.imag) @ np.abs(ob.real + ob.imag) + ar @ br + ai @ bi)
        ref_err = (k + 2) * U * (np.abs(oa) @ np.abs(ob))
        assert (np.abs(got[b].real - ref.real) <= conv + ref_err).all(), b
        assert (np.abs(got[b].imag - ref.imag) <= im_b + ref_err).all(), b


def _spectra_err(mix, ref, nm):
    off = np.cumsum([0] + list(nm))
    out = {}
    for name, key, coef in (("ldos", "gr_diag", -2.0), ("ndos", "gl_diag", 1.0), ("pdos", "gg_diag", -1.0)):
        a = np.stack([mix[key][:, off[i]:off[i + 1]].sum(axis=1).imag for i in range(len(nm))], 1) * coef
        b = np.stack([ref[key][:, off[i]:off[i + 1]].sum(axis=1).imag for i in range(len(nm))], 1) * coef
        out[name] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    for key in ("tr", "tre"):
        out[key] = float(np.abs(mix[key] - ref[key]).max() / max(np.abs(ref[key]).max(), 1e-300))
    return out


def test_3m_sweep_against_fp64(qt):
    """nb = 256 x 24 slabs, 2 kz, 16 energies (the mixed-mode case): every
    observable of the three-product sweep against the four-product one.
    Measured on MI355X: LDOS 4.5e-13, ndos 1.6e-12, pdos 8.9e-13, tr 2.9e-10,
    tre 2.3e-9 (the contact currents are differences of large terms: the
    40-bit mode has 2.4e-6 and 6.6e-5 there); bars 1e-10 and 1e-8."""
    kz = [0.0, np.pi / 2]
    _, dev = hm.synthetic_device(256, 24, seed=1002, kz=kz, tz_density=0.02)
    E = np.repeat(hm.energy_grid(-2.0, 2.0, 8), 2)
    k = np.tile(np.arange(2), 8)
    H = qt.Hamiltonian.from_device(dev)
    fp = qt.solve_batch(H, E, k, mul=0.0, mur=-0.1)
    g3 = qt.solve_batch(H, E, k, mul=0.0, mur=-0.1, precision=qt.QT_PREC_FP64_3M)
    H.close()
    assert (g3["info"] == 0).all()
    e = _spectra_err(g3, fp, dev.nm)
    print("3m vs fp64", e)
    assert max(e.values()) > 0.0
    for key, v in e.items():
        assert v <= (1e-8 if key in ("tr", "tre") else 1e-10), (key, v)


@pytest.mark.parametrize("case", ["C1", "n512"])
def test_3m_oracle_parity(qt, oracle_mod, case):
    """The three-product sweep held to the FP64 parity classification against
    the oracle: the nb = 64 x 32 slab grid of 256 energies and the nb = 512
    short device, as the four-product tests of test_gpu_rgf do."""
    import os

    from tests.test_gpu_rgf import GOLDEN, check_parity

    if not oracle_mod.available("mkl"):
        pytest.skip("needs the vendor-LAPACK oracle flavour")
    if case == "C1":
        _, dev = hm.synthetic_device(64, 32, seed=1001)
        E = hm.energy_grid(-3.0, 3.0, 256)
        k = None
        truth = np.load(os.path.join(GOLDEN, "c1_xprec.npz"))
        kw = {}
    else:
        _, dev = hm.synthetic_device(512, 4, seed=1002, kz=[0.0, np.pi / 2], tz_density=0.02)
        E = np.linspace(-1.9, 1.9, 16)
        k = np.arange(16) % 2
        path = os.path.join(GOLDEN, "n512_xprec.npz")
        truth = np.load(path) if os.path.exists(path) else None
        kw = {"nthreads": 16}
    mkl = oracle_mod.rgf_batch(dev, E, k, 0.0, -0.1, flavour="mkl")
    plain = oracle_mod.rgf_batch(dev, E, k, 0.0, -0.1, flavour="plain", **kw)
    H = qt.Hamiltonian.from_device(dev)
    gpu = qt.solve_batch(H, E, k, mul=0.0, mur=-0.1, precision=qt.QT_PREC_FP64_3M)
    H.close()
    assert (gpu["info"] == 0).all()
    check_parity(gpu, mkl, plain, dev.nm, f"{case} (3m)", truth=truth)
