"""Dense brute-force NEGF on small devices (numpy), the independent check of
the block recursions: G = (E - H - Sigma_L - Sigma_R)^-1 over the whole
device, G^< = G Sigma^< G^H, G^> = G^< + G - G^H."""
import numpy as np


def assemble(Hii, H1i):
    """Full Hamiltonian from the diagonal blocks and H1i[s] = H_{s,s-1}."""
    nm = [h.shape[0] for h in Hii]
    off = np.concatenate([[0], np.cumsum(nm)])
    H = np.zeros((off[-1], off[-1]), dtype=np.complex128)
    for s, h in enumerate(Hii):
        H[off[s]:off[s + 1], off[s]:off[s + 1]] = h
    for s in range(1, len(Hii)):
        c = H1i[s]
        H[off[s]:off[s + 1], off[s - 1]:off[s]] = c
        H[off[s - 1]:off[s], off[s]:off[s + 1]] = c.conj().T
    return H, off


def ferm(x):
    return 1.0 / (1.0 + np.exp(x))


def dense_negf(Hii, H1i, E, sigL, sigR, fL, fR):
    """sigL/sigR: retarded contact self-energies on the first/last slab;
    fL/fR: their occupations.  Returns per-slab diagonal blocks and the
    Caroli transmission Tr[Gamma_L G_1N Gamma_R G_1N^H]."""
    H, off = assemble(Hii, H1i)
    n = H.shape[0]
    Sig = np.zeros_like(H)
    a, b = off[0], off[1]
    c, d = off[-2], off[-1]
    Sig[a:b, a:b] += sigL
    Sig[c:d, c:d] += sigR
    G = np.linalg.inv(E * np.eye(n) - H - Sig)
    GamL = np.zeros_like(H)
    GamR = np.zeros_like(H)
    GamL[a:b, a:b] = 1j * (sigL - sigL.conj().T)
    GamR[c:d, c:d] = 1j * (sigR - sigR.conj().T)
    Sl = 1j * (GamL * fL + GamR * fR)
    Gl = G @ Sl @ G.conj().T
    Gg = Gl + G - G.conj().T
    blocks = lambda M: [M[off[s]:off[s + 1], off[s]:off[s + 1]] for s in range(len(Hii))]
    T = np.trace(GamL[a:b, a:b] @ G[a:b, c:d] @ GamR[c:d, c:d] @ G[a:b, c:d].conj().T).real
    return dict(G_r=blocks(G), G_lesser=blocks(Gl), G_greater=blocks(Gg), T=T)


def chain_surface_g(z, eps, t):
    """Retarded surface Green's function of a semi-infinite 1-D chain with
    on-site eps and hopping t: g = (w - sqrt(w^2 - 4t^2)) / (2 t^2), w = z - eps,
    on the branch with Im g < 0."""
    w = z - eps
    r = np.sqrt(w * w - 4 * t * t + 0j)
    g1 = (w - r) / (2 * t * t)
    g2 = (w + r) / (2 * t * t)
    return g1 if g1.imag < 0 else g2
