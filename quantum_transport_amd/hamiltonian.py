"""Block-tridiagonal device Hamiltonians: Wannier ``ham_dat`` I/O, the
``w90_MAT_DEF`` block builder and the deterministic synthetic generator.

Host-side input preparation for the RGF path.  Block conventions follow the
reference:

* ``Hii[s]`` is the diagonal block of slab ``s`` (0-based here);
* ``H1i[s]`` is H_{s,s-1}, the coupling of slab ``s`` to slab ``s-1``;
  ``H1i[0]`` couples the left lead and ``H1i[nx]`` the right lead
  (deviceHam_mod.f90:278-286, rgf_mod.f90:124-125,157);
* ``Sii`` is the identity (deviceHam_mod.f90:281-284) and is represented
  by ``None``.

All blocks are complex128 numpy arrays in Fortran (column-major) order.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014), vectorised."""
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class SplitMix64:
    """Counter-based SplitMix64 stream: draw i is mix(seed + (i+1)*gamma).

    Sub-streams are derived with :meth:`stream` so that adding draws to one
    part of the generator never shifts another part."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def stream(self, tag: int) -> "SplitMix64":
        s = _mix64(np.array([self.state ^ np.uint64(tag & 0xFFFFFFFFFFFFFFFF)], dtype=np.uint64))[0]
        return SplitMix64(int(s))

    def u64(self, n: int) -> np.ndarray:
        idx = np.arange(self.count + 1, self.count + n + 1, dtype=np.uint64)
        self.count += n
        with np.errstate(over="ignore"):
            return _mix64(self.state + idx * _GAMMA)

    def uniform(self, n: int) -> np.ndarray:
        """Doubles in [0, 1) from the top 53 bits."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller, one normal per pair of uniforms (cosine branch)."""
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)

    def integers(self, n: int, hi: int) -> np.ndarray:
        return np.minimum((self.uniform(n) * hi).astype(np.int64), hi - 1)


# ---------------------------------------------------------------------------
# Wannier ham_dat
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class Wannier:
    """Parsed ``ham_dat`` (wannierHam3d.f90:72-233)."""

    nvb: int
    spin_deg: int
    cell: np.ndarray  # 3x3, columns alpha, beta, gamma
    Hr: np.ndarray  # (nb, nb, nx, ny, nz) complex128
    xmin: int
    ymin: int
    zmin: int
    centres: np.ndarray  # (3, nb)

    @property
    def nb(self) -> int:
        return self.Hr.shape[0]

    @property
    def xmax(self) -> int:
        return self.xmin + self.Hr.shape[2] - 1

    @property
    def ymax(self) -> int:
        return self.ymin + self.Hr.shape[3] - 1

    @property
    def zmax(self) -> int:
        return self.zmin + self.Hr.shape[4] - 1

    def _axes(self):
        alpha, beta, gamm = self.cell[:, 0], self.cell[:, 1], self.cell[:, 2]
        xhat = alpha / np.linalg.norm(alpha)
        yhat = -np.cross(xhat, gamm)
        yhat = yhat / np.linalg.norm(yhat)
        zhat = gamm / np.linalg.norm(gamm)
        return alpha, beta, gamm, xhat, yhat, zhat

    @property
    def Lx(self) -> float:
        alpha, _, _, xhat, _, _ = self._axes()
        return abs(float(np.dot(alpha, xhat)))

    def mat_def(self, kx: float, ky: float, kz: float, ns: int) -> Tuple[np.ndarray, np.ndarray]:
        """Diagonal and sub-diagonal slab blocks, wannierHam3d.f90:273-301.

        Block (i,k) of Hii sums Hr(x=i-k) e^{-i r.k}; block (i,k) of H1i sums
        Hr(x=i-k+ns) with the same phase r = (i-k) alpha + j beta + l gamma."""
        alpha, beta, gamm, xhat, yhat, zhat = self._axes()
        nb = self.nb
        kv = kx * xhat + ky * yhat + kz * zhat
        Hii = np.zeros((nb * ns, nb * ns), dtype=np.complex128, order="F")
        H1i = np.zeros((nb * ns, nb * ns), dtype=np.complex128, order="F")
        for i in range(ns):
            for k in range(ns):
                for j in range(self.ymin, self.ymax + 1):
                    for l in range(self.zmin, self.zmax + 1):
                        r = (i - k) * alpha + j * beta + l * gamm
                        ph = np.exp(-1j * np.dot(r, kv))
                        x0 = i - k
                        if self.xmin <= x0 <= self.xmax:
                            Hii[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb] += (
                                self.Hr[:, :, x0 - self.xmin, j - self.ymin, l - self.zmin] * ph)
                        x1 = i - k + ns
                        if self.xmin <= x1 <= self.xmax:
                            H1i[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb] += (
                                self.Hr[:, :, x1 - self.xmin, j - self.ymin, l - self.zmin] * ph)
        return Hii, H1i

    def bare_coulomb(self, a1: int, a2: int, a3: int, eps: float, r0: float, ldiag: bool) -> np.ndarray:
        """V_ij between orbital i of cell (a1, a2, a3) and orbital j of cell 0
        (wannierHam3d.f90, bare_coulomb): e / (4 pi eps0 eps |r|) tanh(|r| / r0)
        in eV with |r| in Angstrom, r = a1 alpha + a2 beta + a3 gamma + c_i -
        c_j; at r = 0 the self-interaction e / (4 pi eps0 eps r0) when ldiag,
        else 0 (the constants e = 1.6e-19 C, eps0 = 8.85e-12 F/m as there)."""
        alpha, beta, gamm = self.cell[:, 0], self.cell[:, 1], self.cell[:, 2]
        pref = 1.6e-19 / (4.0 * 3.14159265359 * 8.85e-12 * eps * 1e-10)
        r = (a1 * alpha + a2 * beta + a3 * gamm)[:, None, None] + self.centres[:, :, None] - self.centres[:, None, :]
        d = np.linalg.norm(r, axis=0)
        with np.errstate(divide="ignore", invalid="ignore"):
            V = np.where(d > 0, pref / d * np.tanh(d / r0), pref / r0 if ldiag else 0.0)
        return V

    def coulomb_blocks(self, kx: float, ky: float, kz: float, eps: float, r0: float, ns: int = 1,
                       ldiag: bool = True) -> Tuple[np.ndarray, np.ndarray]:
        """Diagonal and sub-diagonal blocks V(I,I), V(I+1,I) of the bare
        Coulomb operator (w90_bare_coulomb_blocks, wannierHam3d.f90:363-390):
        block (i,k) sums bare_coulomb(i-k, j, l) (and i-k+ns for the
        coupling) over the transverse cells with the phase of mat_def."""
        alpha, beta, gamm, xhat, yhat, zhat = self._axes()
        nb = self.nb
        kv = kx * xhat + ky * yhat + kz * zhat
        Vii = np.zeros((nb * ns, nb * ns), dtype=np.complex128, order="F")
        V1i = np.zeros((nb * ns, nb * ns), dtype=np.complex128, order="F")
        for i in range(ns):
            for k in range(ns):
                for j in range(self.ymin, self.ymax + 1):
                    for l in range(self.zmin, self.zmax + 1):
                        ph = np.exp(-1j * np.dot((i - k) * alpha + j * beta + l * gamm, kv))
                        Vii[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb] += self.bare_coulomb(i - k, j, l, eps, r0,
                                                                                            ldiag) * ph
                        V1i[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb] += self.bare_coulomb(i - k + ns, j, l, eps,
                                                                                            r0, ldiag) * ph
        return Vii, V1i


def band_edges(w: Wannier, chunk: int = 512) -> dict:
    """Band-structure scan for the band edges, wannierHam3d.f90:153-201: the bulk
    H(k) of w90_MAT_DEF_3D (:304-323, the sum of Hr over every cell with
    e^{-i r.k}) diagonalised (upper triangle, as zheev 'U') on the k grid of
    w90_PLOT_BZ (:236-270): nk = floor(1/L / 0.01) + 1 points per axis (1 where
    Hr spans one cell), k_i = i dk - pi/L for i = 1..nk, dk = 2 pi/L/(nk - 1) or
    pi/L when nk = 1.  VBM = max over k of band nvb, CBM = min of band nvb + 1;
    kt_vbm / kt_cbm follow :184-191, including their use of nkx with the
    transverse step 2 pi/Ly/(nky - 1) (or 2 pi/Ly) and 0 when Hr spans one
    cell in y.  Returns CBM, VBM, Eg, kt_cbm, kt_vbm (in 2 pi/Ly, as printed)
    and the eigenvalues EN (nb, nkx nky nkz)."""
    alpha, beta, gamm, xhat, yhat, zhat = w._axes()
    Lx = abs(float(np.dot(alpha, xhat)))
    Ly = abs(float(np.dot(beta, yhat)))
    Lz = float(np.linalg.norm(gamm))
    nb = w.nb
    ncx, ncy, ncz = w.Hr.shape[2:]
    nk = [1 if nc == 1 else int(np.floor((1.0 / L) / 0.01)) + 1
          for nc, L in ((ncx, Lx), (ncy, Ly), (ncz, Lz))]
    axes = []
    for n, L in zip(nk, (Lx, Ly, Lz)):
        dk = 2 * np.pi / L / (n - 1) if n > 1 else np.pi / L
        axes.append(np.arange(1, n + 1) * dk - np.pi / L)
    # column i + j nkx + l nkx nky (0-based) holds (kx_i, ky_j, kz_l)
    KZ, KY, KX = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    kv = (KX.reshape(-1)[:, None] * xhat + KY.reshape(-1)[:, None] * yhat
          + KZ.reshape(-1)[:, None] * zhat)
    cells = [(i, j, l) for i in range(w.xmin, w.xmax + 1) for j in range(w.ymin, w.ymax + 1)
             for l in range(w.zmin, w.zmax + 1)]
    R = np.array([i * alpha + j * beta + l * gamm for i, j, l in cells])
    Hc = np.stack([w.Hr[:, :, i - w.xmin, j - w.ymin, l - w.zmin] for i, j, l in cells])
    EN = np.empty((nb, kv.shape[0]))
    for k0 in range(0, kv.shape[0], chunk):
        ph = np.exp(-1j * (kv[k0:k0 + chunk] @ R.T))          # (k, cells)
        H = np.einsum("kc,cab->kab", ph, Hc)
        EN[:, k0:k0 + chunk] = np.linalg.eigvalsh(H, UPLO="U").T
    nvb = w.nvb
    iv = int(np.argmax(EN[nvb - 1])) + 1   # 1-based, first maximum (maxloc)
    ic = int(np.argmin(EN[nvb])) + 1
    dky = 2 * np.pi / Ly / (nk[1] - 1) if nk[1] > 1 else 2 * np.pi / Ly
    if ncy == 1:
        kt_vbm = kt_cbm = 0.0
    else:
        kt_vbm = (iv % nk[0]) * dky - np.pi / Ly
        kt_cbm = (ic % nk[0]) * dky - np.pi / Ly
    VBM, CBM = float(EN[nvb - 1, iv - 1]), float(EN[nvb, ic - 1])
    return {"CBM": CBM, "VBM": VBM, "Eg": CBM - VBM, "kt_cbm": kt_cbm / (2 * np.pi / Ly),
            "kt_vbm": kt_vbm / (2 * np.pi / Ly), "nk": tuple(nk), "EN": EN}


def read_ham_dat(path: str) -> Wannier:
    """Free-format reader with the record order of wannierHam3d.f90:89-111,
    147-152 and 204-208."""
    with open(path, "r") as fh:
        lines = [ln for ln in fh.read().splitlines()]
    pos = 0

    def nxt() -> str:
        nonlocal pos
        while not lines[pos].strip():
            pos += 1
        ln = lines[pos]
        pos += 1
        return ln

    nvb, spin = (int(float(t)) for t in nxt().split()[:2])
    nxt()  # comment
    cell = np.zeros((3, 3))
    for c in range(3):
        cell[:, c] = [float(t) for t in nxt().split()[:3]]
    nxt()  # comment
    n = int(nxt().split()[0])
    rows = np.array([[float(t) for t in nxt().split()[:7]] for _ in range(n)], dtype=np.float64)
    xs, ys, zs = (np.floor(rows[:, c]).astype(np.int64) for c in range(3))
    ms, ns_ = np.floor(rows[:, 3]).astype(np.int64), np.floor(rows[:, 4]).astype(np.int64)
    nb = int(ms.max())
    xmin, ymin, zmin = int(xs.min()), int(ys.min()), int(zs.min())
    Hr = np.zeros((nb, nb, xs.max() - xmin + 1, ys.max() - ymin + 1, zs.max() - zmin + 1),
                  dtype=np.complex128)
    Hr[ms - 1, ns_ - 1, xs - xmin, ys - ymin, zs - zmin] = rows[:, 5] + 1j * rows[:, 6]
    nxt()  # comment
    centres = np.array([[float(t) for t in nxt().split()[:3]] for _ in range(nb)]).T
    return Wannier(nvb, spin, cell, Hr, xmin, ymin, zmin, centres)


def write_ham_dat(path: str, w: Wannier, fmt: str = "%.17g") -> None:
    """Writes the ``ham_dat`` record layout; values at full precision."""
    nb = w.nb
    out = [f"{w.nvb} {w.spin_deg}", "cell"]
    for c in range(3):
        out.append(" ".join(fmt % v for v in w.cell[:, c]))
    out.append("hamiltonian")
    recs = []
    Hr = w.Hr
    for ix in range(Hr.shape[2]):
        for iy in range(Hr.shape[3]):
            for iz in range(Hr.shape[4]):
                blk = Hr[:, :, ix, iy, iz]
                nzm, nzn = np.nonzero(blk.T)  # column-major walk: n outer, m inner
                for n_, m_ in zip(nzm, nzn):
                    v = blk[m_, n_]
                    recs.append(f"{ix + w.xmin} {iy + w.ymin} {iz + w.zmin} {m_ + 1} {n_ + 1} "
                                f"{fmt % v.real} {fmt % v.imag}")
    out.append(str(len(recs)))
    out.extend(recs)
    out.append("orbital_location")
    for i in range(nb):
        out.append(" ".join(fmt % v for v in w.centres[:, i]))
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


# ---------------------------------------------------------------------------
# devH_build_fromWannierFile equivalent
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class Device:
    """Block lists for nk k-points: Hii[ik][s], H1i[ik][s] (nx+1 blocks)."""

    nm: List[int]
    Hii: List[List[np.ndarray]]
    H1i: List[List[np.ndarray]]
    Sii: Optional[List[List[np.ndarray]]] = None
    Lx: float = 1.0
    nb: int = 0
    ns: int = 1

    @property
    def nx(self) -> int:
        return len(self.nm)

    @property
    def nk(self) -> int:
        return len(self.Hii)


def device_from_wannier(w: Wannier, nx: int, ns: int, kpts: Sequence[Tuple[float, float]]) -> Device:
    """deviceHam_mod.f90:234-290: every slab gets the same H00/H10 at kx=0,
    k = (ky, kz) per k-point; Sii is the identity."""
    Hii, H1i = [], []
    for ky, kz in kpts:
        H00, H10 = w.mat_def(0.0, ky, kz, ns)
        Hii.append([H00] * nx)
        H1i.append([H10] * (nx + 1))
    return Device([w.nb * ns] * nx, Hii, H1i, None, w.Lx, w.nb, ns)


# ---------------------------------------------------------------------------
# devH_build_fromCOOfile equivalent
# ---------------------------------------------------------------------------


def read_coo(path: str, use0index: bool = True, iscomplex: bool = False):
    """load_COOmatrix (deviceHam_mod.f90:40-89): one entry per line,
    ``i j re`` or (iscomplex) ``i j re im``; with use0index the indices are
    0-based.  Returns 1-based row and column arrays, the complex values and
    the matrix order.  The reference sets the order to the largest index read
    before the 0-based shift (one less than the order for 0-based files);
    here it is the largest 1-based index."""
    rows, cols, vals = [], [], []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if len(tok) < 3:
                continue
            i, j = int(tok[0]), int(tok[1])
            re = float(tok[2])
            im = float(tok[3]) if iscomplex and len(tok) > 3 else 0.0
            if use0index:
                i, j = i + 1, j + 1
            rows.append(i)
            cols.append(j)
            vals.append(complex(re, im))
    r = np.array(rows, dtype=np.int64)
    c = np.array(cols, dtype=np.int64)
    norb = int(max(r.max(), c.max())) if len(r) else 0
    return r, c, np.array(vals, dtype=np.complex128), norb


def write_coo(path: str, H: np.ndarray, use0index: bool = True, iscomplex: bool = False,
              upper_only: bool = True) -> None:
    """Writes the nonzeros of H in the format read_coo reads (for tests and
    for converting devices built elsewhere)."""
    off = 0 if use0index else 1
    with open(path, "w") as fh:
        for i, j in zip(*np.nonzero(H)):
            if upper_only and j < i:
                continue
            v = H[i, j]
            if iscomplex:
                fh.write(f"{i + off} {j + off} {v.real:.17g} {v.imag:.17g}\n")
            else:
                fh.write(f"{i + off} {j + off} {v.real:.17g}\n")


def device_from_coo(hpath: str, blocksize: int, use0index: bool = True, iscomplex: bool = False,
                    spath: Optional[str] = None) -> Device:
    """devH_build_fromCOOfile with a fixed block size (deviceHam_mod.f90:91-219,
    the route main.f90:113-114 takes for calc_type /= 'w90'): nx = order /
    blocksize uniform slabs of one Hermitian matrix given by the entries of
    its upper triangle (an entry inside a diagonal block is mirrored, as the
    reference does; an entry of block row ix and column ix+1 is H_{ix,ix+1}).

    Kept from the reference: the file format, the mirroring, S = identity
    without an overlap file, the lead couplings repeating the adjacent
    interior coupling.  Different, because the reference's block does not run
    as written: the slab of 1-based row r is (r-1)/blocksize (the reference's
    r/blocksize puts the last row of each slab into the next one, at local
    row 0); H1i[s] holds H_{s,s-1} = H_{s-1,s}^H, the convention the solver
    reads (the reference stores H_{s,s+1} there and leaves the last coupling
    zero); entries below the diagonal are accepted (a lower block entry is
    H_{s,s-1} itself); the order counts 0-based indices correctly.  The
    reference's threshold filter and graph slicing are commented out there
    and not provided."""
    r, c, v, norb = read_coo(hpath, use0index, iscomplex)
    if blocksize <= 0 or norb % blocksize:
        raise ValueError(f"matrix order {norb} is not a multiple of blocksize {blocksize}")
    nx = norb // blocksize
    if nx < 2:
        raise ValueError("the COO route needs at least two slabs")
    bs = blocksize
    Hii = [np.zeros((bs, bs), np.complex128, order="F") for _ in range(nx)]
    H1i = [np.zeros((bs, bs), np.complex128, order="F") for _ in range(nx + 1)]
    for i, j, x in zip(r - 1, c - 1, v):
        si, sj = i // bs, j // bs
        mi, mj = i - si * bs, j - sj * bs
        if si == sj:
            Hii[si][mi, mj] = x
            Hii[si][mj, mi] = np.conj(x)
        elif sj == si + 1:        # H_{si, si+1}: H1i[si+1] = its conjugate transpose
            H1i[sj][mj, mi] = np.conj(x)
        elif sj == si - 1:        # H_{si, si-1} itself
            H1i[si][mi, mj] = x
        else:
            raise ValueError(f"entry ({i + 1}, {j + 1}) couples slabs {si + 1} and {sj + 1}: "
                             "the device is not block-tridiagonal at this block size")
    H1i[0] = H1i[1].copy(order="F")
    H1i[nx] = H1i[nx - 1].copy(order="F")
    Sii = None
    if spath is not None:
        sr, sc, sv, sn = read_coo(spath, use0index, iscomplex)
        if sn > norb:
            raise ValueError("overlap matrix larger than the Hamiltonian")
        Sii = [np.zeros((bs, bs), np.complex128, order="F") for _ in range(nx)]
        for i, j, x in zip(sr - 1, sc - 1, sv):
            si, sj = i // bs, j // bs
            if si != sj:
                raise ValueError("the overlap must be block diagonal")
            Sii[si][i - si * bs, j - sj * bs] = x
            Sii[si][j - sj * bs, i - si * bs] = np.conj(x)
        Sii = [Sii]
    return Device([bs] * nx, [Hii], [H1i], Sii, 1.0, bs, 1)


def energy_grid(emin: float, emax: float, nen: int) -> np.ndarray:
    """negf_mod.f90:100-101."""
    if nen == 1:
        return np.array([emin])
    dE = (emax - emin) / float(nen - 1)
    return emin + np.arange(nen, dtype=np.float64) * dE


# ---------------------------------------------------------------------------
# synthetic generator
# ---------------------------------------------------------------------------

HOP_T = -2.7  # eV


def _hop(rng: SplitMix64, n: int) -> np.ndarray:
    return HOP_T * (1.0 + 0.1 * rng.normal(n))


def _intra_block(rng: SplitMix64, nb: int, partners: int = 4) -> np.ndarray:
    """Real symmetric block: diagonal U[-1,1] eV; each site draws
    ``partners`` random partners (about 2*partners bonds per site), hopping
    t = -2.7 (1 + 0.1 N(0,1)) eV."""
    H = np.zeros((nb, nb), dtype=np.complex128)
    H[np.arange(nb), np.arange(nb)] = 2.0 * rng.stream(1).uniform(nb) - 1.0
    pr = rng.stream(2)
    hv = rng.stream(3)
    for i in range(nb):
        js = pr.integers(partners, nb - 1)
        js = js + (js >= i)  # skip self
        ts = _hop(hv, partners)
        for j, t in zip(js, ts):
            H[i, j] = t
            H[j, i] = t
    return H


def _sparse_block(rng: SplitMix64, nr: int, nc: int, density: float) -> np.ndarray:
    mask = rng.stream(1).uniform(nr * nc).reshape(nc, nr).T < density
    vals = _hop(rng.stream(2), nr * nc).reshape(nc, nr).T
    return np.where(mask, vals, 0.0).astype(np.complex128)


def synthetic_wannier(nb: int, seed: int, tz_density: float = 0.0,
                      coupling_density: float = 0.05, a: float = 3.0) -> Wannier:
    """Wannier model of the synthetic device (one cell per slab).

    x in {-1,0,1}: Hr(x=0) = H_ii, Hr(x=-1) = H_{i,i+1} (random sparse at
    ``coupling_density``), Hr(x=+1) = H_{i,i+1}^H.  With ``tz_density`` > 0 a
    transverse z neighbour is added: Hr(z=-1) = T_z, Hr(z=+1) = T_z^H, so that
    H_ii(kz) = H_ii + T_z e^{i kz} + h.c. (|gamma| = 1).  Orbital centres lie
    on a 3 Angstrom grid."""
    rng = SplitMix64(seed)
    H0 = _intra_block(rng.stream(10), nb)
    Hup = _sparse_block(rng.stream(20), nb, nb, coupling_density)
    nz = 3 if tz_density > 0 else 1
    Hr = np.zeros((nb, nb, 3, 1, nz), dtype=np.complex128)
    zc = 1 if nz == 3 else 0
    Hr[:, :, 1, 0, zc] = H0
    Hr[:, :, 0, 0, zc] = Hup
    Hr[:, :, 2, 0, zc] = Hup.conj().T
    if nz == 3:
        Tz = _sparse_block(rng.stream(30), nb, nb, tz_density)
        Hr[:, :, 1, 0, 0] = Tz
        Hr[:, :, 1, 0, 2] = Tz.conj().T
    cell = np.diag([a, 40.0, 1.0])
    side = int(math.ceil(nb ** (1.0 / 3.0)))
    grid = np.array([(i % side, (i // side) % side, i // (side * side)) for i in range(nb)],
                    dtype=np.float64).T * 3.0
    return Wannier(0, 1, cell, Hr, -1, 0, -(nz // 2), grid)


def synthetic_device(nb: int, nx: int, seed: int, kz: Sequence[float] = (0.0,),
                     tz_density: float = 0.0, coupling_density: float = 0.05) -> Tuple[Wannier, Device]:
    w = synthetic_wannier(nb, seed, tz_density=tz_density, coupling_density=coupling_density)
    return w, device_from_wannier(w, nx, 1, [(0.0, k) for k in kz])


def uneven_device(sizes: Sequence[int], seed: int, coupling_density: float = 0.05) -> Device:
    """Device with slab sizes ``sizes`` built block by block (the direct /
    COO route: the Wannier format cannot express uneven slabs).  The leads
    repeat the first and last slab; H1i[s] is nm[s] x nm[s-1]."""
    rng = SplitMix64(seed)
    nx = len(sizes)
    Hii = [_intra_block(rng.stream(100 + s), n) for s, n in enumerate(sizes)]
    H1i = [_sparse_block(rng.stream(1000), sizes[0], sizes[0], coupling_density).conj().T]
    for s in range(1, nx):
        H1i.append(_sparse_block(rng.stream(1000 + s), sizes[s], sizes[s - 1], coupling_density))
    H1i.append(_sparse_block(rng.stream(1000 + nx), sizes[-1], sizes[-1], coupling_density).conj().T)
    Hii = [np.asfortranarray(h) for h in Hii]
    H1i = [np.asfortranarray(h) for h in H1i]
    return Device(list(sizes), [Hii], [H1i], None, 1.0, 0, 1)
