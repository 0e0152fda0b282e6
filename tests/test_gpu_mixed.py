"""The int8 digit-slice complex GEMM (QT_PREC_MIXED: 6 digits, and
QT_PREC_MIXED47: 7 digits) through the C ABI.

Operands are rounded to Q = 40 (47) bits relative to their row (op(A)) or
column (op(B)) maximum, so for every element
    |C - C_exact| <= 16 K 2^-Q max_k|op(A)[r, k]| max_k|op(B)[k, c]|
(8 real rounding terms per complex product term, dropped digit pairs below
2^-(Q+2)); integer operands of up to 20 (22) bits are represented exactly,
and then the product is exact."""
import numpy as np
import pytest

from quantum_transport_amd import hamiltonian as hm

pytestmark = pytest.mark.gpu

OPS = [lambda x: x, lambda x: x.T, lambda x: x.conj().T]


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf

    rgf.init(0)
    return rgf


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _bound(opA, opB, k, bits=40):
    ra = np.maximum(np.abs(opA.real), np.abs(opA.imag)).max(axis=1)
    cb = np.maximum(np.abs(opB.real), np.abs(opB.imag)).max(axis=0)
    return 16.0 * k * 2.0 ** -bits * np.outer(ra, cb)


@pytest.mark.parametrize("bits", [40, 47])
@pytest.mark.parametrize("opa,opb", [(0, 0), (0, 2), (2, 0), (2, 2), (0, 1), (1, 0)])
@pytest.mark.parametrize("mnk", [(64, 64, 64), (70, 100, 130), (128, 96, 200), (200, 64, 513)])
def test_mixed_zgemm_bound(qt, opa, opb, mnk, bits):
    m, n, k = mnk
    rng = np.random.default_rng(m + 5 * n + 11 * k + opa * 3 + opb)
    batch = 3
    A = _rand(rng, batch, *((m, k) if opa == 0 else (k, m)))
    B = _rand(rng, batch, *((k, n) if opb == 0 else (n, k)))
    # rows and columns of very different size: the scales are per row / column
    A[1] *= np.logspace(-7, 3, A.shape[1])[:, None]
    B[2] *= np.logspace(5, -4, B.shape[2])[None, :]
    R0 = _rand(rng, batch, m, n)
    prec = qt.QT_PREC_MIXED if bits == 40 else qt.QT_PREC_MIXED47
    got = qt.zgemm_batched(A, B, opa, opb, alpha=-1.0, beta=1.0, R=R0, precision=prec)
    for b in range(batch):
        oa, ob = OPS[opa](A[b]), OPS[opb](B[b])
        ref = R0[b] - oa @ ob
        err = np.abs(got[b] - ref)
        # plus the FP64 rounding of the beta R additions here and on the device
        # (below the operand term at 40 bits, not at 47 on the rows scaled
        # down to 1e-7)
        bound = _bound(oa, ob, k, bits) + 2.0 ** -51 * (np.abs(R0[b]) + np.abs(ref))
        assert (err <= bound).all(), (b, float((err / bound).max()))
        # and the error is not far below the bound either: the operands are rounded
        assert np.abs(got[b] - ref).max() > 0.0


@pytest.mark.parametrize("digits", [6, 7])
@pytest.mark.parametrize("mnk", [(64, 64, 64), (96, 160, 300), (512, 512, 512)])
def test_mixed_zgemm_exact_on_integers(qt, mnk, digits):
    """20-bit (7 digits: 22-bit) integer operands are exact in 40 (47) bits,
    so every digit product is exact and the result equals the integer product
    bit for bit"""
    m, n, k = mnk
    rng = np.random.default_rng(k)
    batch = 2
    # 6 digits: |x| < 2^19 keeps every row maximum below 2^19, so the scaled
    # values are multiples of 2^21 and only the top three digits are nonzero
    # (x = -2^19 would need a fourth, whose pair with itself is one of the
    # dropped ones).  7 digits: |x| < 2^21, scaled values multiples of 2^26,
    # the top four digits nonzero, and every pair of them is kept.  Sums stay
    # below 2^53 at K <= 512, so the FP64 fold is exact too.
    lim = 2 ** 19 if digits == 6 else 2 ** 21
    lo, hi = -lim + 1, lim
    A = rng.integers(lo, hi, (batch, m, k)) + 1j * rng.integers(lo, hi, (batch, m, k))
    B = rng.integers(lo, hi, (batch, k, n)) + 1j * rng.integers(lo, hi, (batch, k, n))
    got = qt.zgemm_batched(A, B, 0, 0, precision=qt.QT_PREC_MIXED if digits == 6 else qt.QT_PREC_MIXED47)
    Ar, Ai = A.real.astype(object), A.imag.astype(object)
    for b in range(batch):
        # exact integer reference (object arithmetic), then rounded once to double
        Br, Bi = B[b].real.astype(object), B[b].imag.astype(object)
        re = Ar[b].dot(Br) - Ai[b].dot(Bi)
        im = Ar[b].dot(Bi) + Ai[b].dot(Br)
        ref = re.astype(np.float64) + 1j * im.astype(np.float64)
        assert np.array_equal(got[b], ref), b


def test_mixed_zgemm_zero_rows_and_small_shapes(qt):
    """zero rows/columns come out zero; shapes below 64 take the FP64 kernel"""
    rng = np.random.default_rng(3)
    A = _rand(rng, 2, 96, 80)
    B = _rand(rng, 2, 80, 72)
    A[0, 5] = 0.0
    B[1, :, 7] = 0.0
    got = qt.zgemm_batched(A, B, precision=qt.QT_PREC_MIXED)
    assert (got[0, 5] == 0).all() and (got[1, :, 7] == 0).all()
    A = _rand(rng, 2, 40, 33)
    B = _rand(rng, 2, 33, 50)
    got = qt.zgemm_batched(A, B, precision=qt.QT_PREC_MIXED)
    ref = A @ B
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13


# ----------------------------------------------------------------- mixed-precision sweep

def _spectra(res, nm):
    """per-slab LDOS, ndos, pdos (-2 Im Tr G^R, Im Tr G^<, -Im Tr G^>, the
    sums written by negf_mod.f90:143-149) and the contact currents."""
    off = np.cumsum([0] + list(nm))
    out = {}
    for name, key, coef in (("ldos", "gr_diag", -2.0), ("ndos", "gl_diag", 1.0), ("pdos", "gg_diag", -1.0)):
        d = np.asarray(res[key])
        out[name] = np.stack([d[:, off[i]:off[i + 1]].sum(axis=1).imag * coef for i in range(len(nm))], axis=1)
    out["tr"] = np.asarray(res["tr"])
    out["tre"] = np.asarray(res["tre"])
    return out


def _mixed_errors(mix, ref, nm):
    a, b = _spectra(mix, nm), _spectra(ref, nm)
    return {k: float(np.abs(a[k] - b[k]).max() / max(np.abs(b[k]).max(), 1e-300)) for k in a}


def test_mixed_solve_vs_oracle(qt, oracle_mod, flavour):
    """QT_PREC_MIXED against the oracle: nb = 64 x 32 identical slabs, 24
    energies in [-3, 3] (leads, inversions in FP64; every sweep product on the
    int8 path).  Errors relative to each quantity's largest value, measured on
    MI355X: LDOS 2.3e-7, ndos 6.9e-7, pdos 1.8e-6, tr 2.4e-7, tre 1.8e-7 (the
    FP64 device path: 4e-8 / 9e-12 at these energies, where z - H has no
    broadening) -- 40-bit operands do not reach a 1e-8 bar here.  The bar of
    1e-5 holds the measured accuracy with margin; a wrong digit, scale or
    dropped term moves these errors to O(1)."""
    _, dev = hm.synthetic_device(64, 32, seed=1001)
    E = hm.energy_grid(-3.0, 3.0, 256)[5::11]
    ref = oracle_mod.rgf_batch(dev, E, None, 0.0, -0.1, flavour=flavour)
    H = qt.Hamiltonian.from_device(dev)
    mix = qt.solve_batch(H, E, None, mul=0.0, mur=-0.1, precision=qt.QT_PREC_MIXED)
    fp = qt.solve_batch(H, E, None, mul=0.0, mur=-0.1)
    H.close()
    assert (mix["info"] == 0).all()
    assert np.array_equal(mix["lead_iters"], ref["sancho_iters"])
    e_mix = _mixed_errors(mix, ref, dev.nm)
    e_fp = _mixed_errors(fp, ref, dev.nm)
    print("mixed vs oracle", e_mix, "\nfp64 vs oracle", e_fp)
    for k, v in e_mix.items():
        assert v <= 1e-5, (k, v)


def test_mixed_solve_vs_fp64_kz(qt):
    """nb = 256 x 24 slabs, 2 kz, 16 energies across the band: mixed against
    the FP64 device path (itself held to the oracle).  Measured on MI355X:
    spectra 1.3e-8 - 5e-8, currents tr 3.8e-6, tre 6.9e-5 (the contact
    currents are differences of large terms); bars 1e-6 and 1e-3."""
    kz = [0.0, np.pi / 2]
    _, dev = hm.synthetic_device(256, 24, seed=1002, kz=kz, tz_density=0.02)
    E = np.repeat(hm.energy_grid(-2.0, 2.0, 8), 2)
    k = np.tile(np.arange(2), 8)
    H = qt.Hamiltonian.from_device(dev)
    fp = qt.solve_batch(H, E, k, mul=0.0, mur=-0.1)
    mix = qt.solve_batch(H, E, k, mul=0.0, mur=-0.1, precision=qt.QT_PREC_MIXED)
    H.close()
    e = _mixed_errors(mix, fp, dev.nm)
    print("mixed vs fp64", e)
    for key, v in e.items():
        assert v <= (1e-3 if key in ("tr", "tre") else 1e-6), (key, v)


@pytest.mark.parametrize("variant", ["backward", "uneven"])
def test_mixed_other_variants(qt, variant):
    """The mixed mode in the backward variant (nb = 64 x 8 slabs) and on
    uneven slabs of 64-192 orbitals (coupling fix): against the FP64 device
    path, with the bars of the nb = 256 case (spectra 1e-6, currents 1e-3)."""
    if variant == "backward":
        _, dev = hm.synthetic_device(64, 8, seed=71, coupling_density=0.1)
        kw = dict(variant="backward", mul=0.05, mur=-0.15)
    else:
        dev = hm.uneven_device([64, 128, 96, 192, 64], seed=72, coupling_density=0.1)
        kw = dict(mul=0.0, mur=-0.1, coupling_fix=True)
    E = hm.energy_grid(-2.5, 2.5, 12)
    H = qt.Hamiltonian.from_device(dev)
    fp = qt.solve_batch(H, E, None, **kw)
    mix = qt.solve_batch(H, E, None, precision=qt.QT_PREC_MIXED, **kw)
    H.close()
    assert (mix["info"] == 0).all()
    e = _mixed_errors(mix, fp, dev.nm)
    print(variant, "mixed vs fp64", e)
    assert max(e.values()) > 0.0  # the int8 products ran
    for key, v in e.items():
        assert v <= (1e-3 if key in ("tr", "tre") else 1e-6), (key, v)


def test_mixed_small_blocks_stay_fp64(qt):
    """Blocks below 64 orbitals take the FP64 product in the mixed mode, so
    the results equal the FP64 run bit for bit"""
    _, dev = hm.synthetic_device(32, 6, seed=73)
    E = hm.energy_grid(-2.0, 2.0, 6)
    H = qt.Hamiltonian.from_device(dev)
    fp = qt.solve_batch(H, E, None, mul=0.0, mur=-0.1)
    mix = qt.solve_batch(H, E, None, mul=0.0, mur=-0.1, precision=qt.QT_PREC_MIXED)
    H.close()
    for key in ("gr_diag", "gl_diag", "gg_diag", "tr", "tre"):
        assert np.array_equal(mix[key], fp[key]), key


def test_mixed47_more_accurate(qt):
    """QT_PREC_MIXED47 (47-bit operands) against the FP64 device path on the
    nb = 256 x 24 slab, 2 kz case above and on the uneven device: every
    observable at least 16x closer to FP64 than the 40-bit mode (7 more bits
    of every operand: 128x in the error bound), and the spectra within 1e-8.
    Measured on MI355X, 40 / 47 bits: LDOS 1.5e-8 / 1.4e-10, ndos 9.1e-8 /
    5.5e-10, pdos 4.3e-8 / 2.5e-10, tr 2.4e-6 / 4.2e-8, tre 6.6e-5 / 7.5e-7."""
    kz = [0.0, np.pi / 2]
    _, dev = hm.synthetic_device(256, 24, seed=1002, kz=kz, tz_density=0.02)
    E = np.repeat(hm.energy_grid(-2.0, 2.0, 8), 2)
    k = np.tile(np.arange(2), 8)
    H = qt.Hamiltonian.from_device(dev)
    fp = qt.solve_batch(H, E, k, mul=0.0, mur=-0.1)
    m6 = qt.solve_batch(H, E, k, mul=0.0, mur=-0.1, precision=qt.QT_PREC_MIXED)
    m7 = qt.solve_batch(H, E, k, mul=0.0, mur=-0.1, precision=qt.QT_PREC_MIXED47)
    H.close()
    assert (m7["info"] == 0).all()
    e6, e7 = _mixed_errors(m6, fp, dev.nm), _mixed_errors(m7, fp, dev.nm)
    print("40-bit", e6, "\n47-bit", e7)
    assert max(e7.values()) > 0.0  # the int8 products ran
    for key in e6:
        assert e7[key] <= e6[key] / 16, (key, e6[key], e7[key])
        if key not in ("tr", "tre"):
            assert e7[key] <= 1e-8, (key, e7[key])
