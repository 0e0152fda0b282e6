"""Python host layer over the C ABI of libqtrgf.so (include/qt_rgf.h).

Mirrors the reference's RGF entry points for Python callers and tests:

* :func:`rgf_variableblock_forward` / :func:`rgf_variableblock_backward`
  -- rgf_mod.f90:46-269 / 272-472, one energy point, full block lists;
* :func:`solve_batch` -- the E x k loop of negf_mod.f90:115-127 as one
  batched device call returning the diagonal blocks and contact currents;
* :func:`sancho` -- the lead decimation of rgf_mod.f90:482-568.

There is no CPU fallback: every call goes to the HIP library and raises
:class:`QTError` when it is missing or no device is visible.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libqtrgf.so")

QT_RGF_FORWARD = 0
QT_RGF_BACKWARD = 1
QT_RGF_GW = 2
QT_PREC_FP64 = 0
QT_PREC_MIXED = 1
QT_PREC_MIXED47 = 2
QT_PREC_FP64_3M = 3
QT_F_COUPLING_FIX = 0x1
QT_F_DEVICE_OUTPUTS = 0x2
QT_F_LEAD_RESET = 0x4
QT_INFO_SINGULAR = 0x1
QT_INFO_LEAD_NMAX = 0x2
QT_ERR_SINGULAR = 5
QT_ERR_TOO_LARGE = 8
QT_NOALLOC_SCRATCH_BYTES = 64 << 20

EXPORTED = ["qt_init", "qt_finish", "qt_status_string", "qt_ham_create", "qt_ham_destroy",
            "qt_rgf_solve", "qt_sancho", "qt_device_alloc", "qt_device_free", "qt_memcpy_h2d",
            "qt_memcpy_d2h", "qt_synchronize", "qt_set_kernel_timing", "qt_last_timing",
            "qt_zgemm_batched", "qt_zgemm_batched_prec", "qt_invert_batched", "qt_corr1d2_batched",
            "qt_conv1d2_batched", "qt_gw_polarization", "qt_gw_selfenergy", "qt_gw_mix",
            "qt_gw_wsolve_batched", "qt_gw_sigma_ephoton", "qt_invert_batched_noalloc", "qt_lead_create",
            "qt_lead_destroy", "qt_lead_train", "qt_lead_surface", "qt_lead_param_count", "qt_lead_get_params",
            "qt_lead_set_params", "qt_lead_eval", "qt_graph_slice", "qt_gw_scba_step", "qt_gw_block_current",
            "qt_gw_block_collision"]


class QTError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"qt status {status}: {msg}")
        self.status = status


class Params(C.Structure):
    _fields_ = [("variant", C.c_int32), ("precision", C.c_int32), ("flags", C.c_int32),
                ("lead_nmax", C.c_int32), ("mul", C.c_double), ("mur", C.c_double),
                ("templ", C.c_double), ("tempr", C.c_double), ("lead_eta", C.c_double),
                ("lead_tol", C.c_double)]


class Outputs(C.Structure):
    _fields_ = [("tr", C.POINTER(C.c_double)), ("tre", C.POINTER(C.c_double)),
                ("lead_iters", C.POINTER(C.c_int32)), ("info", C.POINTER(C.c_int32)),
                ("gr_diag", C.c_void_p), ("gl_diag", C.c_void_p), ("gg_diag", C.c_void_p),
                ("G_r", C.POINTER(C.c_void_p)), ("G_lesser", C.POINTER(C.c_void_p)),
                ("G_greater", C.POINTER(C.c_void_p)), ("Jdens", C.POINTER(C.c_void_p)),
                ("Gl", C.POINTER(C.c_void_p)), ("Gln", C.POINTER(C.c_void_p)),
                ("cur", C.POINTER(C.c_void_p)), ("GRi1", C.POINTER(C.c_void_p)),
                ("GLi1", C.POINTER(C.c_void_p)), ("GGi1", C.POINTER(C.c_void_p))]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("lead_ms", C.c_double), ("forward_ms", C.c_double),
                ("backward_ms", C.c_double), ("gemm_ms", C.c_double), ("gemm_flops", C.c_double),
                ("inv_ms", C.c_double), ("inv_flops", C.c_double), ("gemm_launches", C.c_int64),
                ("inv_launches", C.c_int64), ("lead_iterations", C.c_int64), ("batch", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class LeadOpts(C.Structure):
    _fields_ = [("rank", C.c_int32), ("hidden", C.c_int32), ("batch", C.c_int32), ("epochs", C.c_int32),
                ("lr", C.c_float), ("newton_max", C.c_int32), ("newton_tol", C.c_double),
                ("smith_max", C.c_int32), ("sancho_nmax", C.c_int32), ("seed", C.c_uint64)]


class LeadStat(C.Structure):
    _fields_ = [("newton_steps", C.c_int32), ("smith_doublings", C.c_int32), ("fallback", C.c_int32),
                ("sancho_iters", C.c_int32), ("residual", C.c_double), ("warm_err", C.c_double),
                ("sancho_err", C.c_double)]


class LeadReport(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("network_ms", C.c_double), ("newton_ms", C.c_double),
                ("fallback_ms", C.c_double), ("compare_ms", C.c_double), ("points", C.c_int64),
                ("fallbacks", C.c_int64), ("newton_steps", C.c_int64), ("smith_doublings", C.c_int64),
                ("sancho_iterations", C.c_int64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


QT_LEAD_SANCHO = 0
QT_LEAD_NEWTON_MLP = 1
QT_LEAD_NEWTON_BARE = 2
QT_LEAD_NEWTON_GIVEN = 3
QT_LEAD_F_COMPARE = 0x1

_lib: Optional[C.CDLL] = None
_inited = False


def lib() -> C.CDLL:
    """Loads libqtrgf.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise QTError(-1, f"{LIBPATH} not built (python -m quantum_transport_amd.build)")
        L = C.CDLL(LIBPATH)
        L.qt_status_string.restype = C.c_char_p
        L.qt_status_string.argtypes = [C.c_int]
        for name in EXPORTED:
            if name != "qt_status_string":
                getattr(L, name).restype = C.c_int
        L.qt_device_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
        L.qt_device_free.argtypes = [C.c_void_p]
        L.qt_memcpy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.qt_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        for name in ("qt_corr1d2_batched", "qt_conv1d2_batched"):
            getattr(L, name).argtypes = [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_int32]
        L.qt_gw_polarization.argtypes = [C.c_int32] * 6 + [C.c_double] * 2 + [C.c_void_p] * 6 \
            + [C.c_int32]
        L.qt_gw_selfenergy.argtypes = [C.c_int32] * 6 + [C.c_double] + [C.c_void_p] * 9 \
            + [C.c_int32]
        L.qt_gw_mix.argtypes = [C.c_int64, C.c_double, C.c_void_p, C.c_void_p, C.c_int32]
        L.qt_invert_batched_noalloc.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                                C.c_void_p]
        L.qt_lead_create.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                     C.c_double, C.c_double, C.POINTER(LeadOpts), C.POINTER(C.c_void_p)]
        L.qt_lead_destroy.argtypes = [C.c_void_p]
        L.qt_lead_train.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.qt_lead_surface.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
        L.qt_lead_param_count.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.qt_lead_get_params.argtypes = [C.c_void_p, C.c_void_p]
        L.qt_lead_set_params.argtypes = [C.c_void_p, C.c_void_p]
        L.qt_lead_eval.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]
        L.qt_gw_wsolve_batched.argtypes = [C.c_int32] * 3 + [C.c_void_p] * 8 + [C.c_int32]
        L.qt_gw_sigma_ephoton.argtypes = [C.c_int32] * 4 + [C.c_double] * 3 + [C.c_void_p] * 6 \
            + [C.c_int32]
        L.qt_gw_block_current.argtypes = [C.c_int32] * 4 + [C.c_void_p, C.c_double] + [C.c_void_p] * 6
        L.qt_gw_block_collision.argtypes = [C.c_int32] * 3 + [C.c_void_p, C.c_double] + [C.c_void_p] * 6
        L.qt_graph_slice.argtypes = [C.c_int32, C.c_int64, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                     C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_int32)]
        L.qt_gw_scba_step.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int32, C.c_void_p, C.c_int32,
                                      C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def check(st: int) -> int:
    if st != 0:
        raise QTError(st, lib().qt_status_string(st).decode())
    return st


def init(device: int = 0) -> None:
    global _inited
    if not _inited:
        check(lib().qt_init(C.c_int(device)))
        _inited = True


def finish() -> None:
    global _inited
    if _inited:
        check(lib().qt_finish())
        _inited = False


def _f(a) -> np.ndarray:
    return np.asfortranarray(a, dtype=np.complex128)


def _ptrs(arrs: Sequence[Optional[np.ndarray]]):
    return (C.c_void_p * len(arrs))(*[None if a is None else a.ctypes.data for a in arrs])


class Hamiltonian:
    """Device-resident block-tridiagonal Hamiltonian (qt_ham_create).

    Hii[ik][s] (nx blocks per k-point), H1i[ik][s] (nx+1 blocks,
    H1i[s] = H_{s,s-1}), optional Sii[ik][s] (None = identity)."""

    def __init__(self, Hii: List[List[np.ndarray]], H1i: List[List[np.ndarray]],
                 Sii: Optional[List[List[np.ndarray]]] = None):
        init()
        self.nk = len(Hii)
        self.nx = len(Hii[0])
        self.nm = [int(h.shape[0]) for h in Hii[0]]
        self._keep = [_f(h) for ik in range(self.nk) for h in Hii[ik]]
        self._keep1 = [_f(h) for ik in range(self.nk) for h in H1i[ik]]
        self._keepS = None if Sii is None else [_f(h) for ik in range(self.nk) for h in Sii[ik]]
        nm = (C.c_int32 * self.nx)(*self.nm)
        h = C.c_void_p()
        check(lib().qt_ham_create(C.c_int32(self.nx), nm, C.c_int32(self.nk), _ptrs(self._keep),
                                  _ptrs(self._keep1),
                                  None if self._keepS is None else _ptrs(self._keepS), C.byref(h)))
        self.handle = h
        self._keep = self._keep1 = self._keepS = None  # uploaded

    @classmethod
    def from_device(cls, dev) -> "Hamiltonian":
        return cls(dev.Hii, dev.H1i, dev.Sii)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().qt_ham_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_params(variant: str = "forward", mul: float = 0.0, mur: float = 0.0,
                templ: float = 300.0, tempr: float = 300.0, coupling_fix: bool = False,
                lead_eta: float = 1e-5, lead_nmax: int = 50, lead_tol: float = 1e-100,
                lead_reset: bool = False, device_outputs: bool = False,
                precision: int = QT_PREC_FP64) -> Params:
    flags = (QT_F_COUPLING_FIX if coupling_fix else 0) | (QT_F_DEVICE_OUTPUTS if device_outputs else 0) \
        | (QT_F_LEAD_RESET if lead_reset else 0)
    code = {"forward": QT_RGF_FORWARD, "backward": QT_RGF_BACKWARD, "gw": QT_RGF_GW}[variant]
    return Params(code, precision, flags, lead_nmax, mul, mur, templ, tempr, lead_eta, lead_tol)


def solve_batch(H: Hamiltonian, En, kidx=None, *, variant: str = "forward", mul: float = 0.0,
                mur: float = 0.0, templ: float = 300.0, tempr: float = 300.0,
                coupling_fix: bool = False, sigma_lesser_ph=None, sigma_r_ph=None,
                diag: bool = True, full: bool = False, lead_eta: float = 1e-5,
                lead_nmax: int = 50, raise_singular: bool = False,
                precision: int = QT_PREC_FP64) -> dict:
    """Batched RGF over points (En[ip], kidx[ip]).  Returns tr, tre,
    lead_iters, info and (diag=True) the diagonals of G^R, G^<, G^> per
    slab, [npts, sum(nm)]; with full=True also the full block lists
    G_r/G_lesser/G_greater/Jdens/Gl/Gln as [npts][nx] arrays."""
    En = np.ascontiguousarray(En, dtype=np.float64)
    npts = len(En)
    kidx = np.zeros(npts, np.int32) if kidx is None else np.ascontiguousarray(kidx, dtype=np.int32)
    nd = sum(H.nm)
    res = dict(tr=np.zeros(npts), tre=np.zeros(npts), lead_iters=np.zeros((npts, 2), np.int32),
               info=np.zeros(npts, np.int32))
    o = Outputs()
    o.tr = res["tr"].ctypes.data_as(C.POINTER(C.c_double))
    o.tre = res["tre"].ctypes.data_as(C.POINTER(C.c_double))
    o.lead_iters = res["lead_iters"].ctypes.data_as(C.POINTER(C.c_int32))
    o.info = res["info"].ctypes.data_as(C.POINTER(C.c_int32))
    if diag:
        for k in ("gr_diag", "gl_diag", "gg_diag"):
            res[k] = np.zeros((npts, nd), np.complex128)
            setattr(o, k, res[k].ctypes.data)
    keep = []
    if full:
        for k in ("G_r", "G_lesser", "G_greater", "Jdens", "Gl", "Gln"):
            blocks = [[np.zeros((n, n), np.complex128, order="F") for n in H.nm] for _ in range(npts)]
            res[k] = blocks
            arr = _ptrs([b for pt in blocks for b in pt])
            keep.append(arr)
            setattr(o, k, C.cast(arr, C.POINTER(C.c_void_p)))
    sl = sr = None
    if sigma_r_ph is not None:
        sls = [_f(b) for pt in sigma_lesser_ph for b in pt]
        srs = [_f(b) for pt in sigma_r_ph for b in pt]
        keep += [sls, srs]
        sl, sr = _ptrs(sls), _ptrs(srs)
    p = make_params(variant, mul, mur, templ, tempr, coupling_fix, lead_eta, lead_nmax,
                    precision=precision)
    st = lib().qt_rgf_solve(H.handle, C.byref(p), C.c_int32(npts), En.ctypes.data_as(C.c_void_p),
                            kidx.ctypes.data_as(C.c_void_p), sl, sr, C.byref(o))
    if st == QT_ERR_SINGULAR and not raise_singular:
        res["status"] = st
    else:
        check(st)
        res["status"] = 0
    return res


def _single(variant, Hii, H1i, Sii, En, mul, mur, TEMPl, TEMPr, sigma_lesser_ph, sigma_r_ph,
            coupling_fix):
    H = Hamiltonian([Hii], [H1i], None if Sii is None else [Sii])
    try:
        sl = None if sigma_lesser_ph is None else [sigma_lesser_ph]
        sr = None if sigma_r_ph is None else [sigma_r_ph]
        r = solve_batch(H, [En], None, variant=variant, mul=mul, mur=mur, templ=TEMPl,
                        tempr=TEMPr, coupling_fix=coupling_fix, sigma_lesser_ph=sl,
                        sigma_r_ph=sr, diag=False, full=True)
    finally:
        H.close()
    out = {k: r[k][0] for k in ("G_r", "G_lesser", "G_greater", "Jdens", "Gl", "Gln")}
    out.update(tr=float(r["tr"][0]), tre=float(r["tre"][0]), lead_iters=tuple(r["lead_iters"][0]),
               info=int(r["info"][0]))
    return out


def rgf_variableblock_forward(Hii, H1i, En, mul, mur, TEMPl=300.0, TEMPr=300.0, Sii=None,
                              sigma_lesser_ph=None, sigma_r_ph=None, coupling_fix=False) -> dict:
    """rgf_mod.f90:46-269 on the device (one point, full block lists)."""
    return _single("forward", Hii, H1i, Sii, En, mul, mur, TEMPl, TEMPr, sigma_lesser_ph,
                   sigma_r_ph, coupling_fix)


def rgf_variableblock_backward(Hii, H1i, En, mul, mur, TEMPl=300.0, TEMPr=300.0, Sii=None,
                               sigma_lesser_ph=None, sigma_r_ph=None) -> dict:
    """rgf_mod.f90:272-472 on the device (one point, full block lists)."""
    return _single("backward", Hii, H1i, Sii, En, mul, mur, TEMPl, TEMPr, sigma_lesser_ph,
                   sigma_r_ph, False)


def sancho(H00s, H10s, En, S00=None, eta: float = 1e-5, nmax: int = 50, tol: float = 1e-100,
           reset_on_fail: bool = False, want_gbb: bool = True):
    """Batched Sancho-Rubio (rgf_mod.f90:482-568): lists of H00/H10 per
    point; returns G00, GBB lists, iterations and convergence flags."""
    init()
    En = np.ascontiguousarray(En, dtype=np.float64)
    npts = len(En)
    nm = H00s[0].shape[0]
    H00s = [_f(h) for h in H00s]
    H10s = [_f(h) for h in H10s]
    G00 = [np.zeros((nm, nm), np.complex128, order="F") for _ in range(npts)]
    GBB = [np.zeros((nm, nm), np.complex128, order="F") for _ in range(npts)] if want_gbb else None
    it = np.zeros(npts, np.int32)
    cv = np.zeros(npts, np.int32)
    S = None if S00 is None else _f(S00)
    check(lib().qt_sancho(C.c_int32(nm), C.c_int32(npts), En.ctypes.data_as(C.c_void_p),
                          C.c_double(eta), C.c_int32(nmax), C.c_double(tol),
                          C.c_int32(1 if reset_on_fail else 0),
                          None if S is None else S.ctypes.data_as(C.c_void_p), _ptrs(H00s),
                          _ptrs(H10s), _ptrs(G00), None if GBB is None else _ptrs(GBB),
                          it.ctypes.data_as(C.c_void_p), cv.ctypes.data_as(C.c_void_p)))
    return G00, GBB, it, cv.astype(bool)


def set_kernel_timing(enable: bool) -> None:
    init()
    check(lib().qt_set_kernel_timing(C.c_int(1 if enable else 0)))


def last_timing() -> dict:
    t = Timing()
    check(lib().qt_last_timing(C.byref(t)))
    return t.as_dict()


class DeviceBuffer:
    """Raw device allocation through the library (outputs kept resident)."""

    def __init__(self, nbytes: int):
        init()
        self.nbytes = nbytes
        p = C.c_void_p()
        check(lib().qt_device_alloc(C.c_size_t(nbytes), C.byref(p)))
        self.ptr = p

    def to_host(self, dtype, shape) -> np.ndarray:
        a = np.empty(shape, dtype=dtype)
        check(lib().qt_memcpy_d2h(a.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(a.nbytes)))
        return a

    def free(self) -> None:
        if self.ptr:
            lib().qt_device_free(self.ptr)
            self.ptr = None


def solve_batch_device(H: Hamiltonian, En, kidx, bufs: dict, *, variant: str = "forward",
                       mul: float = 0.0, mur: float = 0.0, templ: float = 300.0,
                       tempr: float = 300.0, coupling_fix: bool = False,
                       precision: int = QT_PREC_FP64) -> None:
    """Batched solve with outputs left in device buffers (bufs: tr, tre,
    gr_diag, gl_diag, gg_diag -> DeviceBuffer)."""
    En = np.ascontiguousarray(En, dtype=np.float64)
    npts = len(En)
    kidx = np.zeros(npts, np.int32) if kidx is None else np.ascontiguousarray(kidx, dtype=np.int32)
    o = Outputs()
    if "tr" in bufs:
        o.tr = C.cast(bufs["tr"].ptr, C.POINTER(C.c_double))
    if "tre" in bufs:
        o.tre = C.cast(bufs["tre"].ptr, C.POINTER(C.c_double))
    for k in ("gr_diag", "gl_diag", "gg_diag"):
        if k in bufs:
            setattr(o, k, bufs[k].ptr)
    p = make_params(variant, mul, mur, templ, tempr, coupling_fix, device_outputs=True,
                    precision=precision)
    st = lib().qt_rgf_solve(H.handle, C.byref(p), C.c_int32(npts), En.ctypes.data_as(C.c_void_p),
                            kidx.ctypes.data_as(C.c_void_p), None, None, C.byref(o))
    if st != QT_ERR_SINGULAR:
        check(st)


def zgemm_batched(A: np.ndarray, B: np.ndarray, opa: int = 0, opb: int = 0, alpha: float = 1.0,
                  beta: float = 0.0, R: Optional[np.ndarray] = None,
                  precision: int = QT_PREC_FP64) -> np.ndarray:
    """R[b] = alpha op(A[b]) op(B[b]) + beta R[b] on the device (MUL_C,
    matrix_c.f90:615-660).  A, B: (batch, rows, cols) stacks of column-major
    blocks (each block stored Fortran-ordered).  precision=QT_PREC_MIXED runs
    the product on the int8 matrix cores (operands rounded to 40 bits),
    QT_PREC_MIXED47 likewise with 47 bits; QT_PREC_FP64_3M in FP64 with three
    real products per complex one."""
    init()
    A = np.asarray(A, dtype=np.complex128)
    B = np.asarray(B, dtype=np.complex128)
    batch = A.shape[0]
    ar, ac = A.shape[1:]
    br, bc = B.shape[1:]
    m, k = (ar, ac) if opa == 0 else (ac, ar)
    n = bc if opb == 0 else br
    Af = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))  # column-major blocks
    Bf = np.ascontiguousarray(np.transpose(B, (0, 2, 1)))
    if R is None:
        Rf = np.zeros((batch, n, m), np.complex128)
    else:
        # always a copy: the library writes the result into this buffer
        Rf = np.array(np.transpose(np.asarray(R, np.complex128), (0, 2, 1)), order="C", copy=True)
    check(lib().qt_zgemm_batched_prec(C.c_int32(opa), C.c_int32(opb), C.c_int32(m), C.c_int32(n),
                                      C.c_int32(k), C.c_int32(batch), C.c_double(alpha),
                                      Af.ctypes.data_as(C.c_void_p), C.c_int64(ar), C.c_int64(ar * ac),
                                      Bf.ctypes.data_as(C.c_void_p), C.c_int64(br), C.c_int64(br * bc),
                                      C.c_double(beta), Rf.ctypes.data_as(C.c_void_p), C.c_int64(m),
                                      C.c_int64(m * n), C.c_int32(precision)))
    return np.transpose(Rf, (0, 2, 1))


def invert_batched(A: np.ndarray):
    """Inverse of a (batch, n, n) stack on the device (invert,
    linalg.f90:42-62); returns (inverses, info)."""
    init()
    A = np.asarray(A, dtype=np.complex128)
    batch, n = A.shape[0], A.shape[1]
    # always a copy: the inverse is written in place
    Af = np.array(np.transpose(A, (0, 2, 1)), order="C", copy=True)
    info = np.zeros(batch, np.int32)
    st = lib().qt_invert_batched(C.c_int32(n), C.c_int32(batch), Af.ctypes.data_as(C.c_void_p),
                                 info.ctypes.data_as(C.c_void_p))
    if st not in (0, QT_ERR_SINGULAR):
        check(st)
    return np.transpose(Af, (0, 2, 1)), info


def invert_batched_noalloc(n: int, batch: int, A: "DeviceBuffer", X: "DeviceBuffer",
                           info: Optional["DeviceBuffer"] = None) -> int:
    """qt_invert_batched_noalloc on device buffers: X[b] = inv(A[b]) for
    batch column-major n x n blocks, A destroyed, no device allocation.
    Returns the status (QT_OK or QT_ERR_SINGULAR); raises on other errors."""
    init()
    st = lib().qt_invert_batched_noalloc(C.c_int32(n), C.c_int32(batch), A.ptr, X.ptr,
                                         info.ptr if info is not None else None)
    if st not in (0, QT_ERR_SINGULAR):
        check(st)
    return st


def _corr_conv(name: str, X, Y) -> np.ndarray:
    init()
    X = np.ascontiguousarray(X, dtype=np.complex128)
    Y = np.ascontiguousarray(Y, dtype=np.complex128)
    if X.shape != Y.shape or X.ndim < 1:
        raise ValueError("X and Y must have the same shape (..., n)")
    n = X.shape[-1]
    Z = np.empty_like(X)
    check(getattr(lib(), name)(n, X.size // n if n else 0, X.ctypes.data_as(C.c_void_p),
                               Y.ctypes.data_as(C.c_void_p), Z.ctypes.data_as(C.c_void_p), 0))
    return Z


def corr1d2(X, Y) -> np.ndarray:
    """Energy-axis correlation of every series pair along the last axis
    (corr1d2, fft_mod.f90:58-91): Z[..., i] = sum_e X[..., e] Y[..., e - i + n//2]."""
    return _corr_conv("qt_corr1d2_batched", X, Y)


def conv1d2(X, Y) -> np.ndarray:
    """Energy-axis convolution along the last axis (conv1d2,
    fft_mod.f90:129-168): Z[..., e] = sum_l X[..., e - l] Y[..., l + n//2]."""
    return _corr_conv("qt_conv1d2_batched", X, Y)


def corr_conv_device(mode: str, n: int, nseries: int, X: "DeviceBuffer", Y: "DeviceBuffer",
                     Z: "DeviceBuffer") -> None:
    """corr1d2 / conv1d2 on device-resident series (mode 'corr' or 'conv')."""
    name = {"corr": "qt_corr1d2_batched", "conv": "qt_conv1d2_batched"}[mode]
    check(getattr(lib(), name)(n, nseries, X.ptr, Y.ptr, Z.ptr, QT_F_DEVICE_OUTPUTS))


def _gw_host(arrs):
    out = [np.ascontiguousarray(a, dtype=np.complex128) for a in arrs]
    if any(a.shape != out[0].shape or a.ndim != 4 for a in out):
        raise ValueError("GW arrays must share the shape (nm*nm, nx, nk, nen)")
    nij = out[0].shape[0]
    nm = int(round(nij ** 0.5))
    if nm * nm != nij:
        raise ValueError("first axis must be nm*nm")
    return out, nm


def gw_polarization(Gr, Gl, Gg, nky: int, nkz: int, ndiag: int, dEn: float, spindeg: float):
    """P^R, P^<, P^> of gw_ephoton_3d_ijs (gw_mod.f90:444-550) on the device.
    Arrays: element-major (nm*nm, nx, nky*nkz, nen), element ij = row + col*nm."""
    init()
    (Gr, Gl, Gg), nm = _gw_host((Gr, Gl, Gg))
    _, nx, nk, nen = Gr.shape
    if nk != nky * nkz:
        raise ValueError("third axis must be nky*nkz")
    P = [np.empty_like(Gr) for _ in range(3)]
    check(lib().qt_gw_polarization(nen, nky, nkz, nx, nm, ndiag, dEn, spindeg,
                                   *(a.ctypes.data for a in (Gr, Gl, Gg)),
                                   *(p.ctypes.data for p in P), 0))
    return tuple(P)


def gw_selfenergy(Gr, Gl, Gg, Wr, Wl, Wg, nky: int, nkz: int, ndiag: int, dEn: float):
    """new Sigma^R, Sigma^<, Sigma^> (gw_mod.f90:641-743) on the device."""
    init()
    arrs, nm = _gw_host((Gr, Gl, Gg, Wr, Wl, Wg))
    _, nx, nk, nen = arrs[0].shape
    if nk != nky * nkz:
        raise ValueError("third axis must be nky*nkz")
    S = [np.empty_like(arrs[0]) for _ in range(3)]
    check(lib().qt_gw_selfenergy(nen, nky, nkz, nx, nm, ndiag, dEn, *(a.ctypes.data for a in arrs),
                                 *(x.ctypes.data for x in S), 0))
    return tuple(S)


def gw_sigma_ephoton(Mii, G_lesser, G_greater, nop: int, pre_fact: float, intensity: float,
                     hw: float):
    """e-photon self-energies Sigma^R, Sigma^<, Sigma^> in the monochromatic
    assumption (calc_sigma_ephoton_monochromatic_ext, gw_mod.f90:1075-1122) on
    the device.  Mii: (nm, nm, length); G_lesser, G_greater: (nm, nm, length,
    nen + 2 nop), Fortran-ordered like the reference's arrays.  Returns three
    (nm, nm, length, nen) arrays."""
    init()
    M = np.asfortranarray(Mii, dtype=np.complex128)
    Gl = np.asfortranarray(G_lesser, dtype=np.complex128)
    Gg = np.asfortranarray(G_greater, dtype=np.complex128)
    if M.ndim != 3 or Gl.ndim != 4 or Gl.shape != Gg.shape:
        raise ValueError("Mii must be (nm, nm, length), G (nm, nm, length, nen + 2 nop)")
    nm, nm2, length = M.shape
    if nm != nm2 or Gl.shape[:3] != (nm, nm, length):
        raise ValueError("block shapes disagree")
    nen = Gl.shape[3] - 2 * nop
    if nop < 0 or nen < 0:
        raise ValueError("energy axis shorter than 2 nop")
    S = [np.zeros((nm, nm, length, nen), dtype=np.complex128, order="F") for _ in range(3)]
    check(lib().qt_gw_sigma_ephoton(nm, length, nen, nop, pre_fact, intensity, hw, M.ctypes.data,
                                    Gl.ctypes.data, Gg.ctypes.data, *(x.ctypes.data for x in S), 0))
    return tuple(S)


def gw_block_current(Hii, G_lesser, cur, En, spindeg: float, nb: int, ns: int):
    """Cell-bond current spectra and their energy integrals (the role of
    calc_block_current, gw_mod.f90:384; definition in qt_gw_block_current)
    on the device.  Hii: (nm, nm, nx); G_lesser: (nm, nm, nx, nen) diagonal
    blocks; cur: (nm, nm, nx, nen) slab-bond matrices of green_RGF_RS;
    nm = nb ns.  Returns jdens (nb, nb, nx ns, nen), tot_cur and tot_ecur
    (nb, nb, nx ns), Fortran-ordered."""
    init()
    H = np.asfortranarray(Hii, dtype=np.complex128)
    GL = np.asfortranarray(G_lesser, dtype=np.complex128)
    cu = np.asfortranarray(cur, dtype=np.float64)
    E = np.ascontiguousarray(En, dtype=np.float64)
    if H.ndim != 3 or GL.ndim != 4 or cu.shape != GL.shape or GL.shape[:3] != H.shape:
        raise ValueError("Hii (nm, nm, nx), G_lesser and cur (nm, nm, nx, nen)")
    nm, _, nx, nen = GL.shape
    if nm != nb * ns or len(E) != nen:
        raise ValueError("nm must be nb * ns and En hold nen energies")
    jd = np.zeros((nb, nb, nx * ns, nen), order="F")
    tc = np.zeros((nb, nb, nx * ns), order="F")
    te = np.zeros((nb, nb, nx * ns), order="F")
    check(lib().qt_gw_block_current(nen, nx, nb, ns, E.ctypes.data, spindeg, H.ctypes.data, GL.ctypes.data,
                                    cu.ctypes.data, jd.ctypes.data, tc.ctypes.data, te.ctypes.data))
    return jd, tc, te


def gw_block_collision(Sigma_lesser, Sigma_greater, G_lesser, G_greater, En, spindeg: float):
    """Collision integrals (the role of calc_block_collision, gw_mod.f90:795;
    definition in qt_gw_block_collision) on the device.  All four inputs
    (nm, nm, nx, nen) diagonal blocks.  Returns Itot (nm, nm, nx) and Ispec
    (nm, nm, nx, nen), Fortran-ordered."""
    init()
    A = [np.asfortranarray(x, dtype=np.complex128) for x in (Sigma_lesser, Sigma_greater, G_lesser, G_greater)]
    E = np.ascontiguousarray(En, dtype=np.float64)
    if A[0].ndim != 4 or any(x.shape != A[0].shape for x in A) or A[0].shape[0] != A[0].shape[1]:
        raise ValueError("inputs must be (nm, nm, nx, nen)")
    nm, _, nx, nen = A[0].shape
    if len(E) != nen:
        raise ValueError("En must hold nen energies")
    Ispec = np.zeros((nm, nm, nx, nen), dtype=np.complex128, order="F")
    Itot = np.zeros((nm, nm, nx), dtype=np.complex128, order="F")
    check(lib().qt_gw_block_collision(nen, nx, nm, E.ctypes.data, spindeg, *(x.ctypes.data for x in A),
                                      Itot.ctypes.data, Ispec.ctypes.data))
    return Itot, Ispec


def gw_mix(S, Snew, alpha: float) -> np.ndarray:
    """S + alpha (Snew - S) (gw_mod.f90:747-749) on the device."""
    init()
    S = np.ascontiguousarray(S, dtype=np.complex128).copy()
    Snew = np.ascontiguousarray(Snew, dtype=np.complex128)
    if S.shape != Snew.shape:
        raise ValueError("shape mismatch")
    check(lib().qt_gw_mix(S.size, alpha, S.ctypes.data, Snew.ctypes.data, 0))
    return S


def save_sigma_checkpoint(path: str, iteration: int, Sigma_r, Sigma_lesser, Sigma_greater) -> None:
    """Write the SCBA self-energies of one iteration to ``path`` (.npz, no
    pickled objects).  The reference keeps no SCBA state between runs (it only
    appends Id_iteration.dat per iteration, gw_mod.f90:426-428); this lets a
    GW-SCBA loop resume at ``iteration + 1``.  The file is written under a
    temporary name and renamed, so an interrupted write leaves the previous
    checkpoint intact."""
    arrs = {name: np.asarray(a, dtype=np.complex128) for name, a in
            (("sigma_r", Sigma_r), ("sigma_lesser", Sigma_lesser), ("sigma_greater", Sigma_greater))}
    if len({a.shape for a in arrs.values()}) != 1:
        raise ValueError("the three self-energies must have one shape")
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        np.savez(fh, iteration=np.int64(iteration),
                 fortran_order=np.bool_(all(a.flags.f_contiguous and not a.flags.c_contiguous
                                            for a in arrs.values())), **arrs)
    os.replace(tmp, path)


def load_sigma_checkpoint(path: str):
    """(iteration, Sigma_r, Sigma_lesser, Sigma_greater) from a file written by
    :func:`save_sigma_checkpoint`; arrays come back in the memory order saved."""
    with np.load(path, allow_pickle=False) as z:
        it = int(z["iteration"])
        order = "F" if bool(z["fortran_order"]) else "C"
        out = tuple(np.array(z[n], order=order) for n in ("sigma_r", "sigma_lesser", "sigma_greater"))
    return (it,) + out


def gw_wsolve(Vii, V1i, PR, PL, PG):
    """W^R, W^<, W^> diagonal blocks for a batch of frequencies on the device.
    Vii, V1i: (nm, nm, nx); PR/PL/PG: (nm, nm, nx, nw) (Fortran order, the
    by_energies layout); returns (WR, WL, WG) of the same shape."""
    init()
    Vii = np.asfortranarray(Vii, dtype=np.complex128)
    V1i = np.asfortranarray(V1i, dtype=np.complex128)
    P = [np.asfortranarray(p, dtype=np.complex128) for p in (PR, PL, PG)]
    nm, _, nx = Vii.shape
    if P[0].ndim == 3:
        P = [p[..., None] for p in P]
    nw = P[0].shape[3]
    W = [np.zeros_like(P[0]) for _ in range(3)]
    st = lib().qt_gw_wsolve_batched(nm, nx, nw, Vii.ctypes.data, V1i.ctypes.data,
                                    *(p.ctypes.data for p in P), *(x.ctypes.data for x in W), 0)
    if st != QT_ERR_SINGULAR:
        check(st)
    return tuple(W)


def green_rgf_rs(TEMP, nm: int, nx: int, E: float, mu, Hii, H1i, sigma_lesser_ph=None,
                 sigma_greater_ph=None, sigma_r_ph=None, offdiag: bool = False) -> dict:
    """green_RGF_RS (gw_mod.f90:1131-1345) on the device for one energy.
    Hii, H1i, sigma_*: (nm, nm, nx) arrays (Fortran order, as the reference
    passes them); sigma_greater_ph is accepted and unused, as in the
    reference.  Returns ndens, pdens, ldos (nm, nm, nx), cur (real), tr, tre
    and, with offdiag, GRi1, GLi1, GGi1."""
    Hii = np.asarray(Hii, dtype=np.complex128)
    H1i = np.asarray(H1i, dtype=np.complex128)
    hii = [Hii[:, :, s] for s in range(nx)]
    h1i = [H1i[:, :, s] for s in range(nx)] + [H1i[:, :, nx - 1]]  # right lead: H1i(nx)
    H = Hamiltonian([hii], [h1i])
    init()
    keep = []

    def table(nblk, dtype):
        blocks = [np.zeros((nm, nm), dtype, order="F") for _ in range(nblk)]
        arr = _ptrs(blocks)
        keep.append(arr)
        return blocks, C.cast(arr, C.POINTER(C.c_void_p))

    o = Outputs()
    tr = np.zeros(1)
    tre = np.zeros(1)
    info = np.zeros(1, np.int32)
    o.tr = tr.ctypes.data_as(C.POINTER(C.c_double))
    o.tre = tre.ctypes.data_as(C.POINTER(C.c_double))
    o.info = info.ctypes.data_as(C.POINTER(C.c_int32))
    res = {}
    for key, field, dt in (("ldos", "G_r", np.complex128), ("ndens", "G_lesser", np.complex128),
                           ("pdens", "G_greater", np.complex128), ("cur", "cur", np.float64)):
        blocks, ptr = table(nx, dt)
        res[key] = blocks
        setattr(o, field, ptr)
    if offdiag:
        for key in ("GRi1", "GLi1", "GGi1"):
            blocks, ptr = table(nx, np.complex128)
            res[key] = blocks
            setattr(o, key, ptr)
    sl = sr = None
    if sigma_r_ph is not None:
        sls = [_f(np.asarray(sigma_lesser_ph)[:, :, s]) for s in range(nx)]
        srs = [_f(np.asarray(sigma_r_ph)[:, :, s]) for s in range(nx)]
        keep += [sls, srs]
        sl, sr = _ptrs(sls), _ptrs(srs)
    p = make_params("gw", mu[0], mu[1], TEMP[0], TEMP[1], lead_eta=1e-3, lead_nmax=100,
                    lead_reset=True)
    En = np.array([E], dtype=np.float64)
    k = np.zeros(1, np.int32)
    try:
        st = lib().qt_rgf_solve(H.handle, C.byref(p), C.c_int32(1), En.ctypes.data_as(C.c_void_p),
                                k.ctypes.data_as(C.c_void_p), sl, sr, C.byref(o))
        if st != QT_ERR_SINGULAR:
            check(st)
    finally:
        H.close()
    out = {key: np.stack(v, axis=2) for key, v in res.items()}
    out.update(tr=float(tr[0]), tre=float(tre[0]), info=int(info[0]))
    return out


class GWState:
    """Device-resident arrays of a GW-SCBA run (qt_gw_scba_step): S (Sigma^R,
    Sigma^<, Sigma^>, updated by every step), G and P (written), element-major
    (nm*nm, nx, 1, nen) complex128, and the Coulomb blocks Vii, V1i."""

    def __init__(self, nm: int, nx: int, nen: int, Vii, V1i, S0=None):
        init()
        self.nm, self.nx, self.nen = nm, nx, nen
        self.nbytes = 16 * nm * nm * nx * nen
        self.S = [DeviceBuffer(self.nbytes) for _ in range(3)]
        self.G = [DeviceBuffer(self.nbytes) for _ in range(3)]
        self.P = [DeviceBuffer(self.nbytes) for _ in range(3)]
        V = [np.asfortranarray(v, dtype=np.complex128) for v in (Vii, V1i)]
        self.V = [DeviceBuffer(v.nbytes) for v in V]
        for b, v in zip(self.V, V):
            check(lib().qt_memcpy_h2d(b.ptr, v.ctypes.data_as(C.c_void_p), C.c_size_t(v.nbytes)))
        for i, b in enumerate(self.S):
            a = np.zeros(nm * nm * nx * nen, np.complex128) if S0 is None else \
                np.ascontiguousarray(S0[i], dtype=np.complex128).ravel()
            check(lib().qt_memcpy_h2d(b.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes)))

    def host(self, which: str):
        shape = (self.nm * self.nm, self.nx, 1, self.nen)
        return [b.to_host(np.complex128, shape) for b in getattr(self, which)]

    def free(self):
        for b in self.S + self.G + self.P + self.V:
            b.free()


def gw_scba_step(H: "Hamiltonian", state: GWState, En, mu, TEMP, ndiag: int, spindeg: float = 2.0,
                 alpha: float = 0.5, lead_eta: float = 1e-3, lead_nmax: int = 100, cur=None):
    """One GW-SCBA iteration on the device (qt_gw_scba_step): G by the
    green_RGF_RS recursion with the state's Sigma, P, W, Sigma_new, mixing.
    cur: None or a Fortran-ordered float64 (nm, nm, nx, nen) array that
    receives the slab-bond current matrices of the G pass.  Returns (tr, tre,
    status)."""
    En = np.ascontiguousarray(En, dtype=np.float64)
    p = make_params("gw", mu[0], mu[1], TEMP[0], TEMP[1], lead_eta=lead_eta, lead_nmax=lead_nmax,
                    lead_reset=True)
    tr = np.zeros(len(En))
    tre = np.zeros(len(En))
    ptrs = lambda bufs: (C.c_void_p * 3)(*(b.ptr.value for b in bufs))
    S, G, P = ptrs(state.S), ptrs(state.G), ptrs(state.P)
    if cur is not None and (cur.dtype != np.float64 or not cur.flags.f_contiguous
                            or cur.shape != (state.nm, state.nm, state.nx, len(En))):
        raise ValueError("cur must be a Fortran-ordered float64 (nm, nm, nx, nen) array")
    st = lib().qt_gw_scba_step(H.handle, C.byref(p), len(En), En.ctypes.data_as(C.c_void_p), ndiag, spindeg,
                               alpha, state.V[0].ptr, state.V[1].ptr, S, G, P, tr.ctypes.data_as(C.c_void_p),
                               tre.ctypes.data_as(C.c_void_p), None if cur is None else cur.ctypes.data)
    if st != QT_ERR_SINGULAR:
        check(st)
    return tr, tre, st


# ----------------------------------------------------------------- learned lead warm start


class Lead:
    """A lead with the learned warm start of its surface Green's function
    (qt_lead_*): H00[ik], H10[ik] per kz (the principal layer and its
    coupling), z = E + i eta.  Modes: "sancho" (plain decimation,
    rgf_mod.f90:482-568), "newton" (Newton from the trained network's g0) and
    "bare" (Newton from (z - H00)^-1, the method without the network); both
    Newton modes fall back to decimation."""

    MODES = {"sancho": QT_LEAD_SANCHO, "newton": QT_LEAD_NEWTON_MLP, "bare": QT_LEAD_NEWTON_BARE,
             "given": QT_LEAD_NEWTON_GIVEN}

    def __init__(self, H00s, H10s, kz, eta: float, emin: float, emax: float, *, rank: int = 32,
                 hidden: int = 256, batch: int = 128, epochs: int = 30, lr: float = 1e-3,
                 newton_max: int = 4, newton_tol: float = 1e-10, smith_max: int = 30,
                 sancho_nmax: int = 50, seed: int = 7):
        init()
        self.n = int(H00s[0].shape[0])
        self.nk = len(H00s)
        keep = [_f(h) for h in H00s] + [_f(h) for h in H10s]
        kz = np.ascontiguousarray(kz, dtype=np.float64)
        self.opts = LeadOpts(rank, hidden, batch, epochs, lr, newton_max, newton_tol, smith_max, sancho_nmax, seed)
        h = C.c_void_p()
        check(lib().qt_lead_create(C.c_int32(self.n), C.c_int32(self.nk), _ptrs(keep[:self.nk]),
                                   _ptrs(keep[self.nk:]), kz.ctypes.data_as(C.c_void_p), C.c_double(eta),
                                   C.c_double(emin), C.c_double(emax), C.byref(self.opts), C.byref(h)))
        self.handle = h
        self.rank, self.hidden = rank, hidden
        self.width = (2 * self.n + 4 * self.n * rank + 31) // 32 * 32

    def train(self, E, kidx) -> np.ndarray:
        E = np.ascontiguousarray(E, dtype=np.float64)
        k = np.ascontiguousarray(kidx, dtype=np.int32)
        loss = np.zeros(self.opts.epochs)
        check(lib().qt_lead_train(self.handle, C.c_int32(len(E)), E.ctypes.data_as(C.c_void_p),
                                  k.ctypes.data_as(C.c_void_p), loss.ctypes.data_as(C.c_void_p)))
        return loss

    def surface(self, mode: str, E, kidx, compare: bool = False, blocks: bool = False, start=None) -> dict:
        """start: warm-start blocks for mode "given" (returned updated)."""
        E = np.ascontiguousarray(E, dtype=np.float64)
        k = np.ascontiguousarray(kidx, dtype=np.int32)
        npts = len(E)
        stats = (LeadStat * max(npts, 1))()
        rep = LeadReport()
        if start is not None:
            blocks = True
            G = [np.array(b, np.complex128, order="F") for b in start]
        else:
            G = [np.zeros((self.n, self.n), np.complex128, order="F") for _ in range(npts)] if blocks else None
        check(lib().qt_lead_surface(self.handle, C.c_int32(self.MODES[mode]),
                                    C.c_int32(QT_LEAD_F_COMPARE if compare else 0), C.c_int32(npts),
                                    E.ctypes.data_as(C.c_void_p), k.ctypes.data_as(C.c_void_p),
                                    _ptrs(G) if blocks else None, stats, C.byref(rep)))
        out = {f: np.array([getattr(stats[i], f) for i in range(npts)]) for f, _ in LeadStat._fields_}
        out["report"] = rep.as_dict()
        if blocks:
            out["G00"] = G
        return out

    def params(self) -> np.ndarray:
        cnt = C.c_int64()
        check(lib().qt_lead_param_count(self.handle, C.byref(cnt)))
        p = np.zeros(cnt.value, np.float32)
        check(lib().qt_lead_get_params(self.handle, p.ctypes.data_as(C.c_void_p)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().qt_lead_set_params(self.handle, p.ctypes.data_as(C.c_void_p)))

    def eval(self, E, kidx, G):
        """One forward + backward pass: (output rows, loss, parameter gradient)."""
        E = np.ascontiguousarray(E, dtype=np.float64)
        k = np.ascontiguousarray(kidx, dtype=np.int32)
        B = len(E)
        Gf = np.ascontiguousarray(np.transpose(np.asarray(G, np.complex128), (0, 2, 1)))  # column-major blocks
        Y = np.zeros((B, self.width), np.float32)
        loss = C.c_double()
        grad = np.zeros_like(self.params())
        check(lib().qt_lead_eval(self.handle, C.c_int32(B), E.ctypes.data_as(C.c_void_p),
                                 k.ctypes.data_as(C.c_void_p), Gf.ctypes.data_as(C.c_void_p),
                                 Y.ctypes.data_as(C.c_void_p), C.byref(loss), grad.ctypes.data_as(C.c_void_p)))
        return Y, loss.value, grad

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().qt_lead_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
