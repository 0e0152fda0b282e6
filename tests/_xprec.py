"""Extended-precision (x87 80-bit, numpy clongdouble) restatement of the
forward RGF sweep, the reference against which the FP64 implementations
(the device path and both oracle flavours) are measured at energies where the
device problem is ill-conditioned (z = E + 0i in the device, rgf_mod.f90:62):
there two correct FP64 operation orders differ by far more than 1e-10, and
only a result carrying ~11 more bits tells which of them is closer to the
exact one.

Same physics and conventions as rgf_variableblock_forward
(rgf_mod.f90:46-269): Sancho-Rubio leads at z = E + i eta (:482-568), first
contact filled with f(E - mur, TEMPr), last with f(E - mul, TEMPl)
(:100, :182), spectral currents tr at the last contact and tre at the first
(:191-202, :256-266).  Identical slabs only; all points of a call are carried
as one batch.  Inputs are FP64; every operation after that is extended
precision.
"""
import numpy as np

X = np.clongdouble
BOLTZ = 8.61734e-05


def xinv(A):
    """Batched inverse (..., n, n) by Gauss-Jordan with partial pivoting."""
    A = np.array(A, dtype=X)
    bshape, n = A.shape[:-2], A.shape[-1]
    A = A.reshape(-1, n, n)
    P = A.shape[0]
    M = np.concatenate([A, np.broadcast_to(np.eye(n, dtype=X), (P, n, n))], axis=2)
    ar = np.arange(P)
    for k in range(n):
        col = M[:, k:, k]
        p = k + np.argmax(np.abs(col.real) + np.abs(col.imag), axis=1)
        rk = M[ar, k, :].copy()
        M[ar, k, :] = M[ar, p, :]
        M[ar, p, :] = rk
        M[:, k, :] /= M[:, k, k][:, None]
        f = M[:, :, k].copy()
        f[:, k] = 0
        M -= f[:, :, None] * M[:, k, None, :]
    return M[:, :, n:].reshape(*bshape, n, n)


def dag(A):
    return np.conj(np.swapaxes(A, -1, -2))


def xsancho(E, eta, H00, H10, tol=1e-100, nmax=200):
    """Surface Green's function of the semi-infinite lead for every E
    (rgf_mod.f90:482-568, same update order), iterated until sum |H01| < tol."""
    P, n = len(E), H00.shape[0]
    z = (np.asarray(E, dtype=np.longdouble) + X(1j) * np.longdouble(eta))[:, None, None]
    I = np.eye(n, dtype=X)
    HBB = np.broadcast_to(H00.astype(X), (P, n, n)).copy()
    HSS = HBB.copy()
    H10b = np.broadcast_to(H10.astype(X), (P, n, n)).copy()
    H01b = dag(H10b)
    active = np.ones(P, bool)
    for _ in range(nmax):
        A = xinv(z * I - HBB)
        B = A @ H10b
        C = H01b @ B
        HSS = np.where(active[:, None, None], HSS + C, HSS)
        HBB = np.where(active[:, None, None], HBB + C, HBB)
        Cp = H10b @ B
        B = A @ H01b
        Ap = H10b @ B
        HBB = np.where(active[:, None, None], HBB + Ap, HBB)
        H01n = H01b @ B
        H10b = np.where(active[:, None, None], Cp, H10b)
        H01b = np.where(active[:, None, None], H01n, H01b)
        err = np.abs(H01n).sum(axis=(1, 2))
        active &= ~(err < tol)
        if not active.any():
            break
    return xinv(z * I - HSS)


def ferm(x):
    return 1.0 / (1.0 + np.exp(np.asarray(x, dtype=np.longdouble)))


def xrgf_forward(Hii, H1i, E, mul, mur, templ=300.0, tempr=300.0, eta=1e-5):
    """Forward sweep of a device of identical slabs for all energies E.
    Hii: nx blocks, H1i: nx + 1 couplings (H1i[s] = H_{s,s-1}).  Returns
    dict(gr_diag, gl_diag, gg_diag) of shape (P, nx*n) and tr, tre (P,),
    rounded to FP64."""
    E = np.asarray(E, dtype=np.float64)
    P, nx, n = len(E), len(Hii), Hii[0].shape[0]
    z = E.astype(np.longdouble)[:, None, None].astype(X)
    I = np.eye(n, dtype=X)
    H = [h.astype(X) for h in Hii]
    U = [h.astype(X) for h in H1i]
    # leads: left surface with H10 = H1i[0]^H, right with H10 = H1i[nx]
    gL = xsancho(E, eta, Hii[0], H1i[0].conj().T)
    gR = xsancho(E, eta, Hii[-1], H1i[nx])
    sigL = U[0] @ gL @ dag(U[0])
    sigR = dag(U[nx]) @ gR @ U[nx]
    fF = ferm((E - mur) / (BOLTZ * tempr))[:, None, None]
    fL = ferm((E - mul) / (BOLTZ * templ))[:, None, None]
    injF = -(sigL - dag(sigL)) * fF
    injL = -(sigR - dag(sigR)) * fL
    Gl = [None] * nx
    Gln = [None] * nx
    Gl[0] = xinv(z * I - H[0] - sigL)
    Gln[0] = Gl[0] @ injF @ dag(Gl[0])
    for s in range(1, nx - 1):
        Gl[s] = xinv(z * I - H[s] - U[s] @ Gl[s - 1] @ dag(U[s]))
        Gln[s] = Gl[s] @ (U[s] @ Gln[s - 1] @ dag(U[s])) @ dag(Gl[s])
    s = nx - 1
    GR = [None] * nx
    GL = [None] * nx
    GR[s] = xinv(z * I - H[s] - U[s] @ Gl[s - 1] @ dag(U[s]) - sigR)
    GL[s] = GR[s] @ (injL + U[s] @ Gln[s - 1] @ dag(U[s])) @ dag(GR[s])
    GG = GL[s] + GR[s] - dag(GR[s])
    tr = np.trace(injL @ GG - (-(sigR - dag(sigR)) * (fL - 1)) @ GL[s], axis1=1, axis2=2).real
    GRd, GLd, GGd = [None] * nx, [None] * nx, [None] * nx
    GRd[s], GLd[s], GGd[s] = GR[s], GL[s], GG
    for s in range(nx - 2, -1, -1):
        W = U[s + 1]  # H_{s+1,s}; identical slabs make it H1i[s] as the reference reads it
        GN0 = Gl[s] @ dag(W) @ GRd[s + 1]
        GRd[s] = Gl[s] + GN0 @ W @ Gl[s]
        GLd[s] = (Gln[s] + Gl[s] @ dag(W) @ GLd[s + 1] @ W @ dag(Gl[s]) + GN0 @ W @ Gln[s]
                  + Gln[s] @ dag(W) @ dag(GN0))
        GGd[s] = GLd[s] + GRd[s] - dag(GRd[s])
    tre = np.trace(injF @ GGd[0] - (-(sigL - dag(sigL)) * (fF - 1)) @ GLd[0], axis1=1, axis2=2).real
    diag = lambda L: np.concatenate([np.diagonal(b, axis1=1, axis2=2) for b in L], axis=1)
    return dict(gr_diag=diag(GRd).astype(np.complex128), gl_diag=diag(GLd).astype(np.complex128),
                gg_diag=diag(GGd).astype(np.complex128), tr=tr.astype(np.float64),
                tre=tre.astype(np.float64))
