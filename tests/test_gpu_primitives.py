"""The two block primitives of the path against numpy: the batched complex
GEMM (MUL_C, matrix_c.f90:615-660) and the batched inverse (invert,
linalg.f90:42-62), through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf

    rgf.init(0)
    return rgf


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


OPS = [lambda x: x, lambda x: x.T, lambda x: x.conj().T]


@pytest.mark.parametrize("opa,opb", [(0, 0), (0, 2), (2, 0), (2, 2), (0, 1), (1, 0)])
@pytest.mark.parametrize("mnk", [(16, 16, 4), (64, 64, 16), (70, 33, 19), (128, 96, 200), (1, 1, 1)])
def test_zgemm_vs_numpy(qt, opa, opb, mnk):
    m, n, k = mnk
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    batch = 3
    A = _rand(rng, batch, *((m, k) if opa == 0 else (k, m)))
    B = _rand(rng, batch, *((k, n) if opb == 0 else (n, k)))
    R0 = _rand(rng, batch, m, n)
    got = qt.zgemm_batched(A, B, opa, opb, alpha=-1.0, beta=1.0, R=R0)
    ref = np.stack([R0[b] - OPS[opa](A[b]) @ OPS[opb](B[b]) for b in range(batch)])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    if not err < 1e-13:
        for b in range(min(batch, 3)):
            print("member", b, "got", got[b].ravel()[:4], "ref", ref[b].ravel()[:4],
                  "R0", R0[b].ravel()[:4])
    assert err < 1e-13, err


@pytest.mark.parametrize("n", [1, 3, 12, 16, 17, 40, 63, 64, 65, 96, 200, 513, 700, 1000])
def test_invert_vs_numpy(qt, n):
    rng = np.random.default_rng(n)
    A = _rand(rng, 3, n, n) + 2.0 * np.eye(n)
    X, info = qt.invert_batched(A)
    assert (info == 0).all()
    for b in range(3):
        ref = np.linalg.inv(A[b])
        err = np.abs(X[b] - ref).max() / np.abs(ref).max()
        assert err < 1e-11 * max(1.0, np.linalg.cond(A[b]) / 1e3), (n, b, err)


@pytest.mark.parametrize("n,batch", [(64, 8), (96, 11), (300, 10)])
def test_invert_split_batch(qt, n, batch):
    """batches of 8 and more are split over two streams: every member, the
    singular one included, must come back as with a single stream"""
    rng = np.random.default_rng(n + batch)
    A = _rand(rng, batch, n, n) + 2.0 * np.eye(n)
    A[batch // 2 + 1][:, 3] = 0.0  # singular member in the second half
    X, info = qt.invert_batched(A)
    for b in range(batch):
        if b == batch // 2 + 1:
            assert info[b] != 0 and not X[b].any()
            continue
        assert info[b] == 0
        ref = np.linalg.inv(A[b])
        err = np.abs(X[b] - ref).max() / np.abs(ref).max()
        assert err < 1e-11 * max(1.0, np.linalg.cond(A[b]) / 1e3), (n, b, err)


def test_invert_pivoting_and_singular(qt):
    A = np.array([[0, 1, 2], [1, 0, 3], [4, -3, 8]], dtype=np.complex128)
    S = A.copy()
    S[:, 1] = 0
    X, info = qt.invert_batched(np.stack([A, S]))
    np.testing.assert_allclose(X[0], np.linalg.inv(A), atol=1e-14)
    assert info[0] == 0 and info[1] == 2
    assert np.all(X[1] == 0)
