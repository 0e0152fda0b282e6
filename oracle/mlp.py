"""CPU restatement of the learned lead warm start's network (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the device kernels in
quantum_transport_amd/csrc/qt_mlp.hip and qt_lead.hip (forward pass, loss
and its gradient, backward pass); imported by tests/ and bench.py's CPU leg,
never by the product.

The network has no reference counterpart.  It maps features of (E, kz) to
the output row (d re | d im | U re | U im | V re | V im, U(i, k) at i R + k)
of g0 = diag(d) + U V^H, through 3 GELU layers.  With ``bf16=True`` every
product operand is rounded to BF16 (round to nearest even) exactly where the
device rounds it, and products accumulate in float32; with ``bf16=False``
everything is float64, the form the finite-difference checks pin.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest-even BF16, returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32)


def features(E, kz, emin: float, emax: float, nin: int = 32) -> np.ndarray:
    """Scaled energy, cos kz, sin kz, sin / cos(pi j e) for j = 1..8, zero
    padded to nin (float32, before the BF16 rounding)."""
    E = np.asarray(E, np.float64)
    e = (2.0 * E - (emin + emax)) / (emax - emin)
    f = np.zeros((len(E), nin), np.float32)
    f[:, 0] = e
    f[:, 1] = np.cos(kz)
    f[:, 2] = np.sin(kz)
    for j in range(1, 9):
        f[:, 1 + 2 * j] = np.sin(np.pi * j * e)
        f[:, 2 + 2 * j] = np.cos(np.pi * j * e)
    return f


def unpack(params: np.ndarray, dims):
    """Flat parameters (W0, b0, ..., W3, b3; W row-major in x out) -> lists."""
    W, b, o = [], [], 0
    for i in range(4):
        cnt = dims[i] * dims[i + 1]
        W.append(params[o:o + cnt].reshape(dims[i], dims[i + 1]))
        o += cnt
        b.append(params[o:o + dims[i + 1]])
        o += dims[i + 1]
    assert o == len(params)
    return W, b


def _gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def _gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * 0.3989422804014327 * np.exp(-0.5 * x * x)


class MLP:
    def __init__(self, dims, params: np.ndarray, bf16_products: bool = True):
        self.dims = list(dims)
        self.rnd = bf16 if bf16_products else (lambda a: a)
        self.dt = np.float32 if bf16_products else np.float64
        self.W, self.b = unpack(np.asarray(params, self.dt), self.dims)

    def forward(self, feat: np.ndarray) -> np.ndarray:
        self.A, self.Z = [], []
        a = self.rnd(np.asarray(feat, self.dt))
        for i in range(4):
            self.A.append(a)
            z = (a @ self.rnd(self.W[i])).astype(self.dt) + self.b[i]
            self.Z.append(z)
            if i < 3:
                a = self.rnd(_gelu(z).astype(self.dt))
        return self.Z[3]

    def backward(self, dY: np.ndarray):
        gW, gb = [None] * 4, [None] * 4
        dz = np.asarray(dY, self.dt)
        for i in range(3, -1, -1):
            dzb = self.rnd(dz)
            gW[i] = (self.A[i].T @ dzb).astype(self.dt)
            gb[i] = dz.sum(axis=0).astype(self.dt)
            if i > 0:
                da = (dzb @ self.rnd(self.W[i]).T).astype(self.dt)
                dz = (da * _gelu_d(self.Z[i - 1])).astype(self.dt)
        return np.concatenate([np.concatenate([gW[i].ravel(), gb[i]]) for i in range(4)])


def lowrank(Y: np.ndarray, n: int, R: int):
    """Output rows -> complex d (B, n), U, V (B, n, R)."""
    Y = np.asarray(Y, np.float64)
    d = Y[:, :n] + 1j * Y[:, n:2 * n]
    o = 2 * n
    U = (Y[:, o:o + n * R] + 1j * Y[:, o + n * R:o + 2 * n * R]).reshape(-1, n, R)
    V = (Y[:, o + 2 * n * R:o + 3 * n * R] + 1j * Y[:, o + 3 * n * R:o + 4 * n * R]).reshape(-1, n, R)
    return d, U, V


def loss_and_grad(Y: np.ndarray, G: np.ndarray, n: int, R: int, out_dtype=np.float32):
    """Mean relative squared error of g0 = diag(d) + U V^H against G (B, n, n)
    and its gradient with respect to the output rows."""
    B = Y.shape[0]
    d, U, V = lowrank(Y, n, R)
    P = U @ np.conj(np.swapaxes(V, 1, 2))
    E = P - G
    E[:, np.arange(n), np.arange(n)] += d
    nE = (np.abs(E) ** 2).sum(axis=(1, 2))
    nG = (np.abs(G) ** 2).sum(axis=(1, 2))
    loss = float(np.mean(nE / nG))
    sc = (2.0 / (B * nG))[:, None, None]
    dU = sc * (E @ V)
    dV = sc * (np.conj(np.swapaxes(E, 1, 2)) @ U)
    dd = sc[:, :, 0] * E[:, np.arange(n), np.arange(n)]
    dY = np.zeros(Y.shape, np.float64)
    o = 2 * n
    dY[:, :n], dY[:, n:2 * n] = dd.real, dd.imag
    dY[:, o:o + n * R] = dU.real.reshape(B, -1)
    dY[:, o + n * R:o + 2 * n * R] = dU.imag.reshape(B, -1)
    dY[:, o + 2 * n * R:o + 3 * n * R] = dV.real.reshape(B, -1)
    dY[:, o + 3 * n * R:o + 4 * n * R] = dV.imag.reshape(B, -1)
    return loss, dY.astype(out_dtype)


def warm_start(Y: np.ndarray, n: int, R: int) -> np.ndarray:
    d, U, V = lowrank(Y, n, R)
    P = U @ np.conj(np.swapaxes(V, 1, 2))
    P[:, np.arange(n), np.arange(n)] += d
    return P
