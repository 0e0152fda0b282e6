"""CPU checks of the digit-slice scheme behind QT_PREC_MIXED, restated in
numpy from the split and product kernels (quantum_transport_amd/csrc/
qt_gemm_i8.hip): per-row power-of-two scale putting the row maximum of
|re|, |im| in [2^39, 2^40), rounding to integers, 6 balanced base-128 digits
in [-64, 63], Karatsuba over digit pairs i + j <= 5 with int32 accumulators
per weight, FP64 fold.  No reference counterpart (the mode is new); the
checks pin the arithmetic the kernels implement."""
import numpy as np
import pytest

ND, QB = 6, 40


def split(Y):
    """rows of Y (R x K complex) -> (re digits, im digits, digit sums, scales)"""
    mx = np.maximum(np.abs(Y.real), np.abs(Y.imag)).max(axis=1)
    _, e = np.frexp(np.where(mx > 0, mx, 1.0))  # mx = f 2^e, f in [0.5, 1): ilogb = e - 1
    s = np.where(mx > 0, (QB - 1) - (e - 1), 0)
    xr = np.rint(np.ldexp(Y.real, s[:, None]))
    xi = np.rint(np.ldexp(Y.imag, s[:, None]))
    dr, di = [None] * ND, [None] * ND
    for d in range(ND - 1, -1, -1):
        hr, hi = np.floor((xr + 64) / 128), np.floor((xi + 64) / 128)
        dr[d] = (xr - 128 * hr).astype(np.int64)
        di[d] = (xi - 128 * hi).astype(np.int64)
        xr, xi = hr, hi
    assert not xr.any() and not xi.any()  # everything went into the ND digits
    return dr, di, [dr[d] + di[d] for d in range(ND)], s


def digits_value(ds):
    return sum(d.astype(np.float64) * 128.0 ** (ND - 1 - i) for i, d in enumerate(ds))


def product(A, B):
    """C = A B through the digit planes, as the product kernel computes it"""
    ar, ai, asum, s = split(A)
    br, bi, bsum, t = split(B.T)
    R = []
    for pa, pb in ((ar, br), (ai, bi), (asum, bsum)):
        acc = [sum(pa[i] @ pb[g - i].T for i in range(g + 1)) for g in range(ND)]
        assert max(np.abs(a).max() for a in acc) < 2 ** 31  # int32 accumulators
        r = np.zeros(acc[0].shape)
        for g in range(ND - 1, -1, -1):
            r = r + acc[g].astype(np.float64) * float(1 << (7 * (ND - 1 - g)))
        R.append(r)
    e = 7 * (ND - 1) - s[:, None] - t[None, :]
    return np.ldexp(R[0] - R[1], e) + 1j * np.ldexp(R[2] - R[0] - R[1], e)


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_digits_range_and_reconstruction():
    rng = np.random.default_rng(1)
    Y = _rand(rng, 40, 96) * np.logspace(-9, 6, 40)[:, None]
    Y[3] = 0.0
    dr, di, ds, s = split(Y)
    for d in dr + di:
        assert d.min() >= -64 and d.max() <= 63
    for d in ds:
        assert d.min() >= -128 and d.max() <= 126
    # the digits rebuild the rounded integers exactly, and those are the values to 40 bits
    xr = digits_value(dr)
    assert np.array_equal(xr, np.rint(np.ldexp(Y.real, s[:, None])))
    back = np.ldexp(xr, -s[:, None])
    rowmax = np.maximum(np.abs(Y.real), np.abs(Y.imag)).max(axis=1)
    assert (np.abs(back - Y.real) <= 2.0 ** -40 * rowmax[:, None]).all()
    assert not dr[0][3].any() and s[3] == 0  # a zero row: scale 0, zero digits


@pytest.mark.parametrize("mnk", [(16, 24, 64), (40, 33, 200), (20, 18, 512)])
def test_product_within_bound(mnk):
    m, n, k = mnk
    rng = np.random.default_rng(k)
    A = _rand(rng, m, k) * np.logspace(-7, 3, m)[:, None]
    B = _rand(rng, k, n) * np.logspace(5, -4, n)[None, :]
    C = product(A, B)
    ra = np.maximum(np.abs(A.real), np.abs(A.imag)).max(axis=1)
    cb = np.maximum(np.abs(B.real), np.abs(B.imag)).max(axis=0)
    err = np.abs(C - A @ B)
    bound = 16.0 * k * 2.0 ** -40 * np.outer(ra, cb)
    assert (err <= bound).all()
    # a dropped term or a wrong weight is far outside the bound
    assert (err > 0).any()


def test_product_exact_on_integers():
    rng = np.random.default_rng(5)
    lo, hi = -(2 ** 19) + 1, 2 ** 19
    A = rng.integers(lo, hi, (24, 256)) + 1j * rng.integers(lo, hi, (24, 256))
    B = rng.integers(lo, hi, (256, 20)) + 1j * rng.integers(lo, hi, (256, 20))
    C = product(A, B)
    Ar, Ai = A.real.astype(object), A.imag.astype(object)
    Br, Bi = B.real.astype(object), B.imag.astype(object)
    re = (Ar.dot(Br) - Ai.dot(Bi)).astype(np.float64)
    im = (Ar.dot(Bi) + Ai.dot(Br)).astype(np.float64)
    assert np.array_equal(C, re + 1j * im)


def test_minus_two_pow_19_needs_a_fourth_digit():
    """Integers below 2^19 in magnitude get scale 21, so bit 20 of the scaled
    values is clear and digits 3-5 vanish; a row holding -2^19 gets scale 20,
    and its odd entries then need digit 3, whose pair (3, 3) is dropped"""
    row = np.array([3.0, -5.0, 7.0, 2.0 ** 19 - 1])
    A = np.stack([row, np.append(row[:3], -(2.0 ** 19))]).astype(np.complex128)
    dr, _, _, s = split(A)
    assert s[0] == 21 and s[1] == 20
    assert not any(dr[d][0].any() for d in (3, 4, 5))
    assert dr[3][1].any()
