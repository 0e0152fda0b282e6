"""ctypes bindings of the CPU oracle (libqtoracle*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package never
imports this module.

Functions mirror the reference routines they restate:

* :func:`sancho`       rgf_mod.f90:482-568 / gw_mod.f90:1378-1464
* :func:`invert`       linalg.f90:42-62
* :func:`rgf`          rgf_mod.f90:46-269 (forward) and :272-472 (backward)
* :func:`rgf_batch`    the E x k loop of negf_mod.f90:115-127
* :func:`corr1d2`, :func:`conv1d2`  fft_mod.f90:58-91, 129-168
* :func:`green_rgf_rs` gw_mod.f90:1131-1345
* :func:`gw_polarization`, :func:`gw_selfenergy`, :func:`gw_mix`, :func:`gw_sigma_ephoton`
                       gw_mod.f90:444-519, 641-743, 747-749
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

_libs: Dict[str, C.CDLL] = {}


class _Point(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("nm", C.POINTER(C.c_int)), ("En", C.c_double),
        ("mul", C.c_double), ("mur", C.c_double), ("templ", C.c_double), ("tempr", C.c_double),
        ("Hii", C.POINTER(C.c_void_p)), ("H1i", C.POINTER(C.c_void_p)), ("Sii", C.POINTER(C.c_void_p)),
        ("sigma_lesser_ph", C.POINTER(C.c_void_p)), ("sigma_r_ph", C.POINTER(C.c_void_p)),
        ("lead_eta", C.c_double), ("lead_nmax", C.c_int), ("lead_tol", C.c_double),
        ("coupling_fix", C.c_int),
    ]


class _Result(C.Structure):
    _fields_ = [
        ("G_r", C.POINTER(C.c_void_p)), ("G_lesser", C.POINTER(C.c_void_p)),
        ("G_greater", C.POINTER(C.c_void_p)), ("Jdens", C.POINTER(C.c_void_p)),
        ("Gl", C.POINTER(C.c_void_p)), ("Gln", C.POINTER(C.c_void_p)),
        ("tr", C.c_double), ("tre", C.c_double), ("sancho_iters", C.c_int * 2), ("info", C.c_int),
    ]


def build() -> None:
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def available(flavour: str = "plain") -> bool:
    """True when that flavour is built and loads (the vendor flavour needs
    its BLAS/LAPACK library at run time)."""
    name = "libqtoracle.so" if flavour == "plain" else "libqtoracle_mkl.so"
    if not os.path.exists(os.path.join(BUILD, name)):
        return False
    try:
        lib(flavour)
        return True
    except OSError:
        return False


def lib(flavour: str = "plain") -> C.CDLL:
    if flavour not in _libs:
        name = "libqtoracle.so" if flavour == "plain" else "libqtoracle_mkl.so"
        path = os.path.join(BUILD, name)
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.qto_sancho.restype = C.c_int
        L.qto_invert.restype = C.c_int
        L.qto_ferm.restype = C.c_double
        L.qto_ferm.argtypes = [C.c_double]
        L.qto_rgf_forward.restype = C.c_int
        L.qto_rgf_backward.restype = C.c_int
        L.qto_rgf_batch.restype = C.c_int
        L.qto_blas_flavour.restype = C.c_char_p
        _libs[flavour] = L
    return _libs[flavour]


def _f(a: np.ndarray) -> np.ndarray:
    return np.asfortranarray(a, dtype=np.complex128)


def _ptrs(arrs: Sequence[Optional[np.ndarray]]):
    return (C.c_void_p * len(arrs))(*[None if a is None else a.ctypes.data for a in arrs])


def ferm(a: float, flavour: str = "plain") -> float:
    return lib(flavour).qto_ferm(a)


def invert(A: np.ndarray, flavour: str = "plain"):
    A = _f(A).copy(order="F")
    info = lib(flavour).qto_invert(A.ctypes.data_as(C.c_void_p), C.c_int(A.shape[0]))
    return A, info


def zgemm(opa: int, opb: int, A: np.ndarray, B: np.ndarray, flavour: str = "plain") -> np.ndarray:
    A, B = _f(A), _f(B)
    m = A.shape[0] if opa == 0 else A.shape[1]
    k = A.shape[1] if opa == 0 else A.shape[0]
    n = B.shape[1] if opb == 0 else B.shape[0]
    Cm = np.zeros((m, n), dtype=np.complex128, order="F")
    one, zero = np.complex128(1.0), np.complex128(0.0)
    lib(flavour).qto_zgemm(C.c_int(opa), C.c_int(opb), C.c_int(m), C.c_int(n), C.c_int(k),
                           _cplx(one), A.ctypes.data_as(C.c_void_p),
                           C.c_int(A.shape[0]), B.ctypes.data_as(C.c_void_p), C.c_int(B.shape[0]),
                           _cplx(zero), Cm.ctypes.data_as(C.c_void_p), C.c_int(m))
    return Cm


class _Cplx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def _cplx(z) -> _Cplx:
    return _Cplx(float(np.real(z)), float(np.imag(z)))


def sancho(S00, H00, H10, E: float, eta: float = 1e-5, nmax: int = 50, tol: float = 1e-100,
           reset_on_fail: bool = False, want_gbb: bool = True, flavour: str = "plain"):
    nm = H00.shape[0]
    S00 = _f(np.eye(nm) if S00 is None else S00)
    H00, H10 = _f(H00), _f(H10)
    G00 = np.zeros((nm, nm), dtype=np.complex128, order="F")
    GBB = np.zeros((nm, nm), dtype=np.complex128, order="F") if want_gbb else None
    conv = C.c_int(0)
    it = lib(flavour).qto_sancho(
        C.c_int(nm), C.c_double(E), C.c_double(eta), C.c_int(nmax), C.c_double(tol),
        C.c_int(1 if reset_on_fail else 0), S00.ctypes.data_as(C.c_void_p),
        H00.ctypes.data_as(C.c_void_p), H10.ctypes.data_as(C.c_void_p),
        G00.ctypes.data_as(C.c_void_p), None if GBB is None else GBB.ctypes.data_as(C.c_void_p),
        C.byref(conv))
    return G00, GBB, it, bool(conv.value)


def rgf(Hii: List[np.ndarray], H1i: List[np.ndarray], E: float, mul: float, mur: float,
        templ: float = 300.0, tempr: float = 300.0, Sii=None, sigma_lesser_ph=None,
        sigma_r_ph=None, variant: str = "forward", coupling_fix: bool = False,
        lead_eta: float = 1e-5, lead_nmax: int = 50, lead_tol: float = 1e-100,
        flavour: str = "plain") -> dict:
    """One energy point; returns the full block lists like the reference."""
    nx = len(Hii)
    nm = [h.shape[0] for h in Hii]
    Hii = [_f(h) for h in Hii]
    H1i = [_f(h) for h in H1i]
    Sii = None if Sii is None else [_f(s) for s in Sii]
    sl = None if sigma_lesser_ph is None else [_f(s) for s in sigma_lesser_ph]
    sr = None if sigma_r_ph is None else [_f(s) for s in sigma_r_ph]
    out = {k: [np.zeros((n, n), dtype=np.complex128, order="F") for n in nm]
           for k in ("G_r", "G_lesser", "G_greater", "Jdens", "Gl", "Gln")}
    nm_c = (C.c_int * nx)(*nm)
    keep = []

    def pp(lst):
        if lst is None:
            return None
        a = _ptrs(lst)
        keep.append(a)
        return C.cast(a, C.POINTER(C.c_void_p))

    p = _Point(nx, C.cast(nm_c, C.POINTER(C.c_int)), E, mul, mur, templ, tempr,
               pp(Hii), pp(H1i), pp(Sii), pp(sl), pp(sr), lead_eta, lead_nmax, lead_tol,
               1 if coupling_fix else 0)
    r = _Result(pp(out["G_r"]), pp(out["G_lesser"]), pp(out["G_greater"]), pp(out["Jdens"]),
                pp(out["Gl"]), pp(out["Gln"]), 0.0, 0.0, (C.c_int * 2)(0, 0), 0)
    L = lib(flavour)
    fn = L.qto_rgf_forward if variant == "forward" else L.qto_rgf_backward
    st = fn(C.byref(p), C.byref(r))
    if st != 0:
        raise ValueError(f"oracle rgf returned {st}")
    out.update(tr=r.tr, tre=r.tre, sancho_iters=(r.sancho_iters[0], r.sancho_iters[1]), info=r.info)
    return out


def rgf_batch(device, En: np.ndarray, kidx: Optional[np.ndarray], mul: float, mur: float,
              templ: float = 300.0, tempr: float = 300.0, variant: str = "forward",
              coupling_fix: bool = False, nthreads: int = 0, flavour: str = "plain") -> dict:
    """E x k batch (one point per OpenMP thread); returns tr, tre and the
    diagonals of G_r, G_lesser, G_greater per point."""
    nx, nk = device.nx, device.nk
    Hii = [_f(h) for ik in range(nk) for h in device.Hii[ik]]
    H1i = [_f(h) for ik in range(nk) for h in device.H1i[ik]]
    npts = len(En)
    En = np.ascontiguousarray(En, dtype=np.float64)
    kidx = np.zeros(npts, dtype=np.int32) if kidx is None else np.ascontiguousarray(kidx, dtype=np.int32)
    nd = sum(device.nm)
    tr = np.zeros(npts)
    tre = np.zeros(npts)
    gr = np.zeros((npts, nd), dtype=np.complex128)
    gl = np.zeros((npts, nd), dtype=np.complex128)
    gg = np.zeros((npts, nd), dtype=np.complex128)
    it = np.zeros((npts, 2), dtype=np.int32)
    nm_c = (C.c_int * nx)(*device.nm)
    hp, h1p = _ptrs(Hii), _ptrs(H1i)
    st = lib(flavour).qto_rgf_batch(
        C.c_int(0 if variant == "forward" else 1), C.c_int(nx), nm_c, C.c_int(nk), hp, h1p, None,
        C.c_int(npts), En.ctypes.data_as(C.c_void_p), kidx.ctypes.data_as(C.c_void_p),
        C.c_double(mul), C.c_double(mur), C.c_double(templ), C.c_double(tempr),
        C.c_int(1 if coupling_fix else 0), C.c_int(nthreads), tr.ctypes.data_as(C.c_void_p),
        tre.ctypes.data_as(C.c_void_p), gr.ctypes.data_as(C.c_void_p),
        gl.ctypes.data_as(C.c_void_p), gg.ctypes.data_as(C.c_void_p), it.ctypes.data_as(C.c_void_p))
    if st != 0:
        raise ValueError(f"oracle rgf_batch returned {st}")
    return dict(tr=tr, tre=tre, gr_diag=gr, gl_diag=gl, gg_diag=gg, sancho_iters=it)


def corr1d2(X: np.ndarray, Y: np.ndarray, flavour: str = "plain") -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.complex128)
    Y = np.ascontiguousarray(Y, dtype=np.complex128)
    Z = np.zeros_like(X)
    lib(flavour).qto_corr1d2(C.c_int(len(X)), X.ctypes.data_as(C.c_void_p),
                             Y.ctypes.data_as(C.c_void_p), Z.ctypes.data_as(C.c_void_p))
    return Z


def conv1d2(X: np.ndarray, Y: np.ndarray, flavour: str = "plain") -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.complex128)
    Y = np.ascontiguousarray(Y, dtype=np.complex128)
    Z = np.zeros_like(X)
    lib(flavour).qto_conv1d2(C.c_int(len(X)), X.ctypes.data_as(C.c_void_p),
                             Y.ctypes.data_as(C.c_void_p), Z.ctypes.data_as(C.c_void_p))
    return Z


def green_rgf_rs(Hii, H1i, E: float, mu, TEMP, sigma_lesser_ph=None, sigma_r_ph=None,
                 offdiag: bool = True, flavour: str = "plain") -> dict:
    """gw_mod.f90:1131-1345 on (nm, nm, nx) stacks (Fortran order)."""
    Hii = np.asfortranarray(Hii, dtype=np.complex128)
    H1i = np.asfortranarray(H1i, dtype=np.complex128)
    nm, _, nx = Hii.shape
    sl = None if sigma_lesser_ph is None else np.asfortranarray(sigma_lesser_ph, dtype=np.complex128)
    sr = None if sigma_r_ph is None else np.asfortranarray(sigma_r_ph, dtype=np.complex128)
    out = {k: np.zeros((nm, nm, nx), np.complex128, order="F") for k in ("ndens", "pdens", "ldos")}
    out["cur"] = np.zeros((nm, nm, nx), np.float64, order="F")
    for k in ("GRi1", "GLi1", "GGi1"):
        out[k] = np.zeros((nm, nm, nx), np.complex128, order="F") if offdiag else None
    tr = C.c_double(0.0)
    tre = C.c_double(0.0)
    T = np.ascontiguousarray(TEMP, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    st = lib(flavour).qto_green_rgf_rs(p(T), C.c_int(nm), C.c_int(nx), C.c_double(E), p(mu), p(Hii),
                                       p(H1i), p(sl), p(sr), p(out["ndens"]), p(out["pdens"]),
                                       p(out["ldos"]), C.byref(tr), C.byref(tre), p(out["cur"]),
                                       p(out["GRi1"]), p(out["GLi1"]), p(out["GGi1"]))
    if st != 0:
        raise ValueError(f"oracle green_rgf_rs returned {st}")
    out.update(tr=tr.value, tre=tre.value)
    return out


def _gw_arrays(*arrs):
    return [np.ascontiguousarray(a, dtype=np.complex128) for a in arrs]


def gw_polarization(Gr, Gl, Gg, nky: int, nkz: int, ndiag: int, dEn: float, spindeg: float,
                    flavour: str = "plain"):
    """P^R, P^<, P^> (gw_mod.f90:444-519) of element-major arrays of shape
    (nm*nm, nx, nk, nen), element ij = row + col*nm."""
    Gr, Gl, Gg = _gw_arrays(Gr, Gl, Gg)
    nij, nx, nk, nen = Gr.shape
    nm = int(round(nij ** 0.5))
    assert nm * nm == nij and nk == nky * nkz
    P = [np.zeros_like(Gr) for _ in range(3)]
    lib(flavour).qto_gw_polarization(
        C.c_int(nen), C.c_int(nky), C.c_int(nkz), C.c_int(nx), C.c_int(nm), C.c_int(ndiag),
        C.c_double(dEn), C.c_double(spindeg), *(a.ctypes.data_as(C.c_void_p) for a in (Gr, Gl, Gg)),
        *(p.ctypes.data_as(C.c_void_p) for p in P))
    return tuple(P)


def gw_selfenergy(Gr, Gl, Gg, Wr, Wl, Wg, nky: int, nkz: int, ndiag: int, dEn: float,
                  flavour: str = "plain"):
    """new Sigma^R, Sigma^<, Sigma^> (gw_mod.f90:641-743), same layout."""
    Gr, Gl, Gg, Wr, Wl, Wg = _gw_arrays(Gr, Gl, Gg, Wr, Wl, Wg)
    nij, nx, nk, nen = Gr.shape
    nm = int(round(nij ** 0.5))
    S = [np.zeros_like(Gr) for _ in range(3)]
    lib(flavour).qto_gw_selfenergy(
        C.c_int(nen), C.c_int(nky), C.c_int(nkz), C.c_int(nx), C.c_int(nm), C.c_int(ndiag),
        C.c_double(dEn), *(a.ctypes.data_as(C.c_void_p) for a in (Gr, Gl, Gg, Wr, Wl, Wg)),
        *(s.ctypes.data_as(C.c_void_p) for s in S))
    return tuple(S)


def gw_mix(S, Snew, alpha: float, flavour: str = "plain"):
    S = np.ascontiguousarray(S, dtype=np.complex128).copy()
    Snew = np.ascontiguousarray(Snew, dtype=np.complex128)
    lib(flavour).qto_gw_mix(C.c_int64(S.size), C.c_double(alpha), S.ctypes.data_as(C.c_void_p),
                            Snew.ctypes.data_as(C.c_void_p))
    return S


def gw_sigma_ephoton(Mii, Gl, Gg, nop: int, pre_fact: float, intensity: float, hw: float,
                     flavour: str = "plain"):
    """e-photon Sigma^R, Sigma^<, Sigma^> (gw_mod.f90:1075-1122) on Fortran-ordered
    Mii (nm, nm, length) and G (nm, nm, length, nen + 2 nop)."""
    M = np.asfortranarray(Mii, dtype=np.complex128)
    G = [np.asfortranarray(a, dtype=np.complex128) for a in (Gl, Gg)]
    nm, _, length = M.shape
    nen = G[0].shape[3] - 2 * nop
    S = [np.zeros((nm, nm, length, nen), np.complex128, order="F") for _ in range(3)]
    lib(flavour).qto_gw_sigma_ephoton(C.c_int(nm), C.c_int(length), C.c_int(nen), C.c_int(nop),
                                      C.c_double(pre_fact), C.c_double(intensity), C.c_double(hw),
                                      *(a.ctypes.data_as(C.c_void_p) for a in [M] + G),
                                      *(x.ctypes.data_as(C.c_void_p) for x in S))
    return tuple(S)


def gw_wsolve(Vii, V1i, PR, PL, PG, flavour: str = "plain"):
    """W^R, W^<, W^> diagonal blocks (the W solve called at gw_mod.f90:597) on
    (nm, nm, nx) Fortran-ordered stacks; V1i[:, :, i] = V_{i,i-1}."""
    arrs = [np.asfortranarray(a, dtype=np.complex128) for a in (Vii, V1i, PR, PL, PG)]
    nm, _, nx = arrs[0].shape
    W = [np.zeros((nm, nm, nx), np.complex128, order="F") for _ in range(3)]
    st = lib(flavour).qto_gw_wsolve(C.c_int(nm), C.c_int(nx),
                                    *(a.ctypes.data_as(C.c_void_p) for a in arrs),
                                    *(x.ctypes.data_as(C.c_void_p) for x in W))
    if st != 0:
        raise ValueError("singular block in the W solve")
    return tuple(W)
