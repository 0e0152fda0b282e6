"""qt_invert_batched_noalloc (the no-allocation variant of qt_invert_batched,
device buffers in and out, pivots in the library's fixed scratch array)
against qt_invert_batched on the same blocks: bit-identical inverses and
info, on two calls in a row, the second one smaller."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf

    rgf.init(0)
    return rgf


def _blocks(rng, n, batch, singular=()):
    A = rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))
    A += 2.0 * np.sqrt(n) * np.eye(n)  # keep the bulk well conditioned
    for b in singular:
        A[b, 5, :] = 0.0  # an all-zero row: exactly singular
    return A


def _noalloc(qt, A, with_info):
    batch, n = A.shape[0], A.shape[1]
    Af = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))  # column-major blocks
    dA = qt.DeviceBuffer(Af.nbytes)
    dX = qt.DeviceBuffer(Af.nbytes)
    dI = qt.DeviceBuffer(4 * batch) if with_info else None
    try:
        qt.check(qt.lib().qt_memcpy_h2d(dA.ptr, Af.ctypes.data_as(C.c_void_p), C.c_size_t(Af.nbytes)))
        print(f"noalloc call: n={n} batch={batch} info={'device' if with_info else 'NULL'}", flush=True)
        st = qt.invert_batched_noalloc(n, batch, dA, dX, dI)
        X = np.transpose(dX.to_host(np.complex128, (batch, n, n)), (0, 2, 1))
        info = dI.to_host(np.int32, (batch,)) if with_info else None
    finally:
        for d in (dA, dX, dI):
            if d is not None:
                d.free()
    return st, X, info


def test_noalloc_matches_invert_batched_twice(qt):
    rng = np.random.default_rng(2024)
    # first call: 24 blocks of 96 with two singular ones, info on the device
    A1 = _blocks(rng, 96, 24, singular=(3, 17))
    st, X, info = _noalloc(qt, A1, True)
    Xr, inforef = qt.invert_batched(A1)
    assert st == qt.QT_ERR_SINGULAR
    assert np.array_equal(info, inforef) and info[3] != 0 and info[17] != 0
    assert np.array_equal(X, Xr)
    assert np.all(X[3] == 0) and np.all(X[17] == 0)
    ok = [b for b in range(24) if b not in (3, 17)]
    err = np.abs(np.einsum("bij,bjk->bik", A1[ok], X[ok]) - np.eye(96)).max()
    assert err < 1e-12, err
    # second call, smaller, after the first one left its pivots and flags in
    # the scratch array: no singular block, no info pointer
    A2 = _blocks(rng, 40, 5)
    st, X, _ = _noalloc(qt, A2, False)
    Xr, inforef = qt.invert_batched(A2)
    assert st == 0 and not inforef.any()
    assert np.array_equal(X, Xr)
    assert np.abs(np.einsum("bij,bjk->bik", A2, X) - np.eye(40)).max() < 1e-12


def test_noalloc_rejects_requests_beyond_scratch(qt):
    d = qt.DeviceBuffer(256)
    e = qt.DeviceBuffer(256)
    try:
        # 1024 * 1024 + 1024 + 256 int32 exceed the fixed scratch: rejected
        # before anything is launched (the buffers are never touched)
        st = qt.lib().qt_invert_batched_noalloc(C.c_int32(1024), C.c_int32(1024), d.ptr, e.ptr, None)
        assert st == qt.QT_ERR_TOO_LARGE
        assert b"scratch" in qt.lib().qt_status_string(st)
        # A and X overlapping: invalid
        st = qt.lib().qt_invert_batched_noalloc(C.c_int32(2), C.c_int32(1), d.ptr, d.ptr, None)
        assert st == 1
    finally:
        d.free()
        e.free()
