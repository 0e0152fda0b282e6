"""Pins the CPU restatement of the warm-start network (oracle/mlp.py) before
it checks the device kernels: BF16 rounding against its definition, the
output-row layout against a hand-built g0, and the whole gradient (network,
low-rank factors, loss) against central finite differences in float64."""
import numpy as np
import pytest

from oracle import mlp


def test_bf16_rounding():
    one = np.float32(1.0)
    assert mlp.bf16(np.array([one]))[0] == 1.0
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even one
    assert mlp.bf16(np.array([1 + 2.0 ** -8], np.float32))[0] == 1.0
    assert mlp.bf16(np.array([1 + 3 * 2.0 ** -8], np.float32))[0] == 1 + 2.0 ** -6
    assert mlp.bf16(np.array([1 + 2.0 ** -8 + 2.0 ** -12], np.float32))[0] == 1 + 2.0 ** -7
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(10000) * np.exp(rng.uniform(-20, 20, 10000))).astype(np.float32)
    r = mlp.bf16(x)
    assert np.all(r.view(np.uint32) & 0xFFFF == 0)  # 8 significant bits
    ulp = np.abs(r) * 2.0 ** -7
    assert np.all(np.abs(r - x) <= 0.5 * ulp * (1 + 1e-6))  # nearest


def test_output_layout_and_warm_start():
    n, R = 3, 2
    rng = np.random.default_rng(2)
    d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    U = rng.standard_normal((n, R)) + 1j * rng.standard_normal((n, R))
    V = rng.standard_normal((n, R)) + 1j * rng.standard_normal((n, R))
    row = np.zeros(2 * n + 4 * n * R + 2)  # padded row
    row[:n], row[n:2 * n] = d.real, d.imag
    o = 2 * n
    for i in range(n):
        for k in range(R):
            row[o + i * R + k] = U[i, k].real
            row[o + n * R + i * R + k] = U[i, k].imag
            row[o + 2 * n * R + i * R + k] = V[i, k].real
            row[o + 3 * n * R + i * R + k] = V[i, k].imag
    P = mlp.warm_start(row[None], n, R)[0]
    np.testing.assert_allclose(P, np.diag(d) + U @ V.conj().T, rtol=1e-14, atol=1e-14)
    loss, dY = mlp.loss_and_grad(row[None], P[None], n, R, np.float64)
    assert loss <= 1e-28 and np.abs(dY).max() <= 1e-15


def test_gradient_finite_differences():
    """float64 network: the analytic gradient of the loss with respect to
    every kind of parameter (weights and biases of all four layers) equals
    central differences."""
    n, R, nin, hid = 3, 2, 32, 8
    W = 2 * n + 4 * n * R + 2  # a padded output row, as the device keeps it
    dims = [nin, hid, hid, hid, W]
    rng = np.random.default_rng(3)
    cnt = sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(4))
    p = rng.standard_normal(cnt) * 0.4
    B = 4
    feat = mlp.features(rng.uniform(-2, 2, B), rng.uniform(0, np.pi, B), -2.0, 2.0)
    G = rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n))

    def loss_of(q):
        net = mlp.MLP(dims, q, bf16_products=False)
        return mlp.loss_and_grad(net.forward(feat), G, n, R, np.float64)[0]

    net = mlp.MLP(dims, p, bf16_products=False)
    _, dY = mlp.loss_and_grad(net.forward(feat), G, n, R, np.float64)
    g = net.backward(dY)
    # padded outputs carry no gradient
    assert not dY[:, 2 * n + 4 * n * R:].any()
    offs = np.cumsum([0] + [dims[i] * dims[i + 1] + dims[i + 1] for i in range(4)])
    picks = []
    for i in range(4):
        picks += list(rng.integers(offs[i], offs[i] + dims[i] * dims[i + 1], 4))  # weights
        picks += list(offs[i] + dims[i] * dims[i + 1] + rng.integers(0, dims[i + 1], 2))  # biases
    h = 1e-6
    for j in picks:
        e = np.zeros_like(p)
        e[j] = h
        fd = (loss_of(p + e) - loss_of(p - e)) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(abs(g).max(), 1e-8), (j, fd, g[j])
