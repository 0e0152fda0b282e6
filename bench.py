#!/usr/bin/env python3
"""RGF throughput at the production shape on one MI355X.

Workload: block-tridiagonal device of 100 identical slabs with nb=512
orbitals each, 4 kz points (H_ii(kz) = H_ii + T_z e^{ikz} + h.c.), energies
on the 1024-point grid in [-2, 2] eV, Fermi contacts mu_L=0, mu_R=-0.1 eV,
T=300 K; both Sancho-Rubio leads and the full forward + backward sweep
(G^R, G^<, G^> diagonals, contact currents) per energy point, complex128.

A step is one device batch of --batch energy points.  The (E, kz) pairs are
taken in the order of a fixed stride-2053 permutation of the 1024 x 4 grid,
so every batch (warm-up and timed) spreads over the whole energy window and
the mix of lead decimation depths is the grid's.  Inputs (Hamiltonian
blocks) are resident in HBM before the timed region; outputs stay on the
device.

The JSON line carries the roofline of the dominant kernel (the batched
complex GEMM on the FP64 matrix cores: algorithmic 8mnk flops per launch over
the launch durations measured with HIP events on the library stream) and a
CPU baseline (the FP64 oracle on the vendor LAPACK, one energy point per
thread).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RGF energy-points/s @nb=512×100 slabs: FP64 vs mixed+warm-start @≤1e-8 err"
# complex products and inversions run per slab of a ballistic sweep, 8 n^3 flops
# each: pass 1 five products and the inversion, pass 2 seven products between
# slabs of equal size (eight otherwise)
SLAB_FLOPS_EQUAL = 13 * 8
SLAB_FLOPS_UNEVEN = 14 * 8
PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 matrix
# HBM bytes moved by one 512^3 complex product without an addend (A and B read
# once, C written once), measured with rocprofv3 FETCH_SIZE / WRITE_SIZE and a
# copy-kernel calibration; a launch moves this per batch member
GEMM_BYTES_PER_MEMBER_512 = 12 * 2 ** 20
STRIDE = 2053  # odd stride of the permutation that orders the (E, kz) points


def host_facts() -> dict:
    """nproc, CPU model and the BLAS/LAPACK library of the CPU baseline."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = "plain C restatement (no vendor BLAS)"
    try:
        import oracle

        if oracle.available("mkl"):
            mk = C.CDLL(os.path.join("/opt/conda/lib", "libmkl_rt.so"))
            buf = C.create_string_buffer(256)
            mk.MKL_Get_Version_String(buf, C.c_int(256))
            blas = buf.value.decode(errors="replace").strip()
    except (OSError, AttributeError):
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "blas": blas}


def cpu_baseline(nb: int, slabs: int, threads: int, seed: int, validate: bool = False) -> dict:
    """Times the oracle (vendor LAPACK flavour, one point per OpenMP thread)
    on a bounded sample and extrapolates to the full slab count: the cost of
    a point is lead + slabs * slab, with slab and lead separated from runs at
    2 and 6 slabs.  With ``validate`` the same points are also timed at the
    full slab count (about 50 s on 16 threads)."""
    import oracle
    from quantum_transport_amd import hamiltonian as hm

    flavour = "mkl" if oracle.available("mkl") else "plain"
    kz = [0.0, np.pi / 4, np.pi / 2, 3 * np.pi / 4]
    w = hm.synthetic_wannier(nb, seed, tz_density=0.02)
    E = hm.energy_grid(-2.0, 2.0, 1024)[:: max(1, 1024 // threads)][:threads]
    k = np.arange(threads) % 4
    times = {}
    for nx in (2, 6):
        dev = hm.device_from_wannier(w, nx, 1, [(0.0, q) for q in kz])
        t0 = time.perf_counter()
        oracle.rgf_batch(dev, E, k, 0.0, -0.1, nthreads=threads, flavour=flavour)
        times[nx] = time.perf_counter() - t0
    slab = max(times[6] - times[2], 0.0) / 4.0
    lead = max(times[2] - 2 * slab, 0.0)
    t_full = lead + slabs * slab  # wall time of `threads` points in parallel
    out = {
        "value": threads / t_full,
        "unit": "energy points/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{threads} energy points (one per thread, {flavour} ZGEMM/ZGETRF/ZGETRI) at "
                   f"nb={nb}, timed at 2 and 6 slabs ({times[2]:.1f}s, {times[6]:.1f}s) and "
                   f"extrapolated to {slabs} slabs (lead {lead:.1f}s + {slabs} x {slab:.2f}s)"),
    }
    out.update(host_facts())
    if validate:
        # the same points at the full slab count, timed directly
        dev = hm.device_from_wannier(w, slabs, 1, [(0.0, q) for q in kz])
        t0 = time.perf_counter()
        oracle.rgf_batch(dev, E, k, 0.0, -0.1, nthreads=threads, flavour=flavour)
        t_direct = time.perf_counter() - t0
        out["validation"] = {"direct_s": round(t_direct, 2), "extrapolated_s": round(t_full, 2),
                             "ratio": round(t_direct / t_full, 3)}
    return out


def run_gw(args) -> dict:
    """GW-SCBA energy-axis stage (P and Sigma assembly, gw_mod.f90:444-743) at
    the GW shape: nm=256 orbitals per slab, every element of the diagonal
    blocks, 2048 energies, one kz, on --gw-slabs slabs (the per-slab work of
    the 50-slab device).  Arrays are device resident (torch allocations
    handed to the library as device pointers)."""
    import torch
    from quantum_transport_amd import rgf

    nm, nx, nen = 256, args.gw_slabs, 2048
    rgf.init(0)
    dev = torch.device("cuda", 0)
    total = nm * nm * nx * nen
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)

    def arr():
        return torch.randn(total, dtype=torch.complex128, device=dev, generator=g)

    G = [arr() for _ in range(3)]
    Wt = [arr() for _ in range(3)]
    out = [torch.empty(total, dtype=torch.complex128, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    L = rgf.lib()
    ptr = lambda t: t.data_ptr()
    dEn = 8.0 / (nen - 1)

    def pol():
        rgf.check(L.qt_gw_polarization(nen, 1, 1, nx, nm, nm, dEn, 2.0, *(ptr(t) for t in G),
                                       *(ptr(t) for t in out), rgf.QT_F_DEVICE_OUTPUTS))

    def sig():
        rgf.check(L.qt_gw_selfenergy(nen, 1, 1, nx, nm, nm, dEn, *(ptr(t) for t in G),
                                     *(ptr(t) for t in Wt), *(ptr(t) for t in out),
                                     rgf.QT_F_DEVICE_OUTPUTS))

    times = {}
    for name, fn in (("polarization", pol), ("selfenergy", sig)):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        times[name] = (time.perf_counter() - t0) / args.steps
    series = nm * nm * nx
    # algorithmic HBM bytes: P reads G^R, G^<, G^> and writes P^R, P^<, P^>
    # once per element and slab (the element pair ij/ji shares its reads),
    # Sigma reads 6 series and writes 3 (16 B per complex sample)
    bytes_p = series * nen * 16 * 6
    bytes_s = series * nen * 16 * 9
    bw_p = bytes_p / times["polarization"] / 1e9
    bw_s = bytes_s / times["selfenergy"] / 1e9
    return {
        "metric": "GW P + Sigma assembly, element-slabs/s @nm=256, 2048 E",
        "value": round(series / (times["polarization"] + times["selfenergy"]), 1),
        "unit": "element-slabs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * (times["polarization"] + times["selfenergy"]), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (torch normal samples)",
        "config": {"workload": f"GW P and Sigma assembly, nm=256 (all elements), {nx} slabs, "
                               f"2048 energies, 1 kz", "nm": nm, "slabs": nx, "energies": nen},
        "roofline": {"bound": "hbm", "achieved": round(bw_p, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(bw_p / 8000.0, 4), "traffic": None,
                     "kernel": "gw_pol_pair_kernel (P assembly)"},
        "stages": {"polarization_ms": round(1e3 * times["polarization"], 2),
                   "selfenergy_ms": round(1e3 * times["selfenergy"], 2),
                   "selfenergy_GBps": round(bw_s, 1)},
        "cpu_baseline": None,
    }


def run_corrconv(args) -> dict:
    """Batched corr1d2 and conv1d2 (fft_mod.f90:58-91, 129-168) on device-resident
    series: 65536 series pairs of 2048 energies (FFT length 4096).  Algorithmic
    HBM bytes per series: two inputs and one output of n complex128 samples."""
    import torch
    from quantum_transport_amd import rgf

    n, ns = 2048, 65536
    rgf.init(0)
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    X = torch.randn(ns * n, dtype=torch.complex128, device=dev, generator=g)
    Y = torch.randn(ns * n, dtype=torch.complex128, device=dev, generator=g)
    Z = torch.empty(ns * n, dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    L = rgf.lib()
    res = {}
    for mode, fn in (("corr", L.qt_corr1d2_batched), ("conv", L.qt_conv1d2_batched)):
        def call():
            rgf.check(fn(n, ns, X.data_ptr(), Y.data_ptr(), Z.data_ptr(), rgf.QT_F_DEVICE_OUTPUTS))
        for _ in range(max(args.warmup, 1)):
            call()
        L.qt_synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            call()
        L.qt_synchronize()
        res[mode] = (time.perf_counter() - t0) / args.steps
    bytes_call = 3.0 * ns * n * 16
    bw = {m: bytes_call / t / 1e9 for m, t in res.items()}
    return {"metric": "batched corr1d2 series/s @n=2048 (FFT length 4096)",
            "value": round(ns / res["corr"], 1), "unit": "series/s", "n_gpus": 1,
            "steps": args.steps, "warmup": max(args.warmup, 1),
            "ms_per_step": round(1e3 * res["corr"], 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (torch normal samples)",
            "config": {"workload": f"corr1d2 / conv1d2, {ns} series pairs of {n} samples", "n": n,
                       "series": ns},
            "roofline": {"bound": "hbm", "achieved": round(bw["corr"], 1), "peak": 8000.0, "unit": "GB/s",
                         "frac": round(bw["corr"] / 8000.0, 4), "traffic": None,
                         "kernel": "corrconv_kernel (corr1d2)"},
            "stages": {"corr_ms": round(1e3 * res["corr"], 3), "conv_ms": round(1e3 * res["conv"], 3),
                       "conv_GBps": round(bw["conv"], 1)},
            "cpu_baseline": None}


def run_ephoton(args) -> dict:
    """e-photon self-energies in the monochromatic assumption
    (calc_sigma_ephoton_monochromatic_ext, gw_mod.f90:1075-1122) on
    device-resident arrays: nm = 128 orbitals, 32 slabs, 512 energies, the
    photon shift nop = 8 energies.  Algorithmic flops per (slab, energy): two
    complex products for each of Sigma^< and Sigma^> (4 x 8 nm^3); the rate is
    over the whole call (shifted sums, four GEMM launches, Sigma^R)."""
    import torch
    from quantum_transport_amd import rgf

    nm, length, nen, nop = 128, 32, 512, 8
    rgf.init(0)
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    nn = nm * nm
    M = torch.randn(nn * length, dtype=torch.complex128, device=dev, generator=g)
    G = [torch.randn(nn * length * (nen + 2 * nop), dtype=torch.complex128, device=dev, generator=g)
         for _ in range(2)]
    S = [torch.empty(nn * length * nen, dtype=torch.complex128, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    L = rgf.lib()

    def call():
        rgf.check(L.qt_gw_sigma_ephoton(nm, length, nen, nop, 1.0, 1.0, 0.1, M.data_ptr(),
                                        G[0].data_ptr(), G[1].data_ptr(), *(x.data_ptr() for x in S),
                                        rgf.QT_F_DEVICE_OUTPUTS))
    for _ in range(max(args.warmup, 1)):
        call()
    L.qt_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()
    L.qt_synchronize()
    t = (time.perf_counter() - t0) / args.steps
    flops = 32.0 * nm ** 3 * length * nen
    tf = flops / t / 1e12
    return {"metric": "e-photon self-energy (slab, energy) blocks/s @nm=128",
            "value": round(length * nen / t, 1), "unit": "blocks/s", "n_gpus": 1,
            "steps": args.steps, "warmup": max(args.warmup, 1),
            "ms_per_step": round(1e3 * t, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (torch normal samples)",
            "config": {"workload": f"Sigma^R/</> e-photon, nm={nm}, {length} slabs x {nen} energies, nop={nop}",
                       "nm": nm, "slabs": length, "energies": nen, "nop": nop},
            "roofline": {"bound": "mfma", "achieved": round(tf, 3), "peak": 78.6, "unit": "TFLOP/s",
                         "frac": round(tf / 78.6, 4), "traffic": None,
                         "kernel": "whole call: eph_sum + 4 x zgemm_mfma_kernel + eph_retarded"},
            "cpu_baseline": None if args.no_cpu else ephoton_cpu_baseline(args, nm, nop)}


def ephoton_cpu_baseline(args, nm: int, nop: int) -> dict:
    """The oracle on host cores: every thread computes one slab of 256 energies
    20 times (one BLAS thread each; ~10 s on 16 cores), all threads on the same
    seeded blocks."""
    import concurrent.futures as cf

    import oracle
    flavour = "mkl" if oracle.available("mkl") else "plain"
    if flavour == "mkl":
        for name in ("MKL_Set_Num_Threads", "mkl_set_num_threads_"):
            fn = getattr(oracle.lib("mkl"), name, None)
            if fn is not None:
                fn(C.c_int(1)) if name[0] == "M" else fn(C.byref(C.c_int(1)))
                break
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    nen, reps = 256, 20
    rng = np.random.default_rng(args.seed)
    cr = lambda *sh: np.asfortranarray(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    M, Gl, Gg = cr(nm, nm, 1), cr(nm, nm, 1, nen + 2 * nop), cr(nm, nm, 1, nen + 2 * nop)
    with cf.ThreadPoolExecutor(threads) as ex:
        # untimed warm-up: the BLAS library's first call in each thread
        w = 2 * nop + 2
        list(ex.map(lambda _: oracle.gw_sigma_ephoton(M, Gl[..., :w], Gg[..., :w], nop, 1.0, 1.0, 0.1,
                                                      flavour=flavour), range(threads)))
        t0 = time.perf_counter()
        list(ex.map(lambda _: [oracle.gw_sigma_ephoton(M, Gl, Gg, nop, 1.0, 1.0, 0.1, flavour=flavour)
                               for _ in range(reps)], range(threads)))
        tc = time.perf_counter() - t0
    return {"value": round(threads * reps * nen / tc, 1), "unit": "blocks/s", "cores": threads,
            "kind": "port",
            "sample": f"{threads} threads x {reps} calls of (1 slab x {nen} energies) at nm={nm} "
                      f"({flavour} ZGEMM, one BLAS thread each), {tc:.1f}s"}


def run_vendor_zgemm(args) -> dict:
    """The vendor library's batched complex128 GEMM (torch.bmm, which calls
    rocBLAS) at the sweep's product shape, for comparison with the
    ``roofline.achieved`` of the default RGF run (our batched GEMM kernel).
    Not a path of this library: a reference point only."""
    import torch
    dev = torch.device("cuda:0")
    B, n = args.batch, args.nb
    g = torch.Generator(device=dev).manual_seed(args.seed)
    A = torch.randn(B, n, n, dtype=torch.complex128, device=dev, generator=g)
    X = torch.randn(B, n, n, dtype=torch.complex128, device=dev, generator=g)
    C = torch.empty_like(A)
    for _ in range(max(args.warmup, 1)):
        torch.bmm(A, X, out=C)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(args.steps, 1) * 4
    e0.record()
    for _ in range(reps):
        torch.bmm(A, X, out=C)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    tf = 8.0 * n ** 3 * B / (ms * 1e-3) / 1e12
    return {"metric": f"vendor batched ZGEMM TFLOP/s @{B}x{n}^3 (torch.bmm -> rocBLAS)",
            "value": round(tf, 3), "unit": "TFLOP/s", "n_gpus": 1, "steps": reps,
            "warmup": max(args.warmup, 1), "ms_per_step": round(ms, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (torch normal samples)",
            "config": {"workload": f"complex128 batched GEMM, {B} members of {n}^3", "batch": B, "n": n},
            "roofline": {"bound": "mfma", "achieved": round(tf, 3), "peak": 78.6, "unit": "TFLOP/s",
                         "frac": round(tf / 78.6, 4), "traffic": None, "kernel": "rocBLAS zgemm (torch.bmm)"},
            "cpu_baseline": None}


def time_batches(rgf, H, Eall, Kall, B: int, nd: int, steps: int, warmup: int,
                 coupling_fix: bool = False, precision: int = 0):
    """Run ``warmup`` untimed and ``steps`` timed device batches of B points
    (outputs stay on the device); returns the wall time of the timed batches
    and the summed kernel timings."""
    bufs = {"tr": rgf.DeviceBuffer(8 * B), "tre": rgf.DeviceBuffer(8 * B)}
    for key in ("gr_diag", "gl_diag", "gg_diag"):
        bufs[key] = rgf.DeviceBuffer(16 * B * nd)
    cursor = [0]
    N = len(Eall)
    perm = (np.arange(N) * STRIDE) % N  # STRIDE is odd and N a power of two: a permutation
    sampled = []

    def step():
        idx = perm[(cursor[0] + np.arange(B)) % N]
        cursor[0] += B
        sampled.append(idx)
        rgf.solve_batch_device(H, Eall[idx], Kall[idx], bufs, mul=0.0, mur=-0.1,
                               coupling_fix=coupling_fix, precision=precision)

    for _ in range(warmup):
        step()
    rgf.lib().qt_synchronize()
    rgf.set_kernel_timing(True)
    acc = {"gemm_ms": 0.0, "gemm_flops": 0.0, "inv_ms": 0.0, "inv_flops": 0.0, "gemm_launches": 0,
           "lead_iterations": 0, "lead_ms": 0.0, "forward_ms": 0.0, "backward_ms": 0.0}
    rgf.lib().qt_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        t = rgf.last_timing()
        for k in acc:
            acc[k] += t[k]
    rgf.lib().qt_synchronize()
    elapsed = time.perf_counter() - t0
    rgf.set_kernel_timing(False)
    for b in bufs.values():
        b.free()
    timed = np.concatenate(sampled[warmup:]) if steps else np.zeros(0, int)
    acc["sample"] = {"points": int(len(timed)), "E_min": float(Eall[timed].min()) if len(timed) else None,
                     "E_max": float(Eall[timed].max()) if len(timed) else None,
                     "E_mean": float(Eall[timed].mean()) if len(timed) else None,
                     "kz_counts": np.bincount(Kall[timed], minlength=int(Kall.max()) + 1).tolist()}
    return elapsed, acc


def spectra_errors(mix: dict, ref: dict, nm) -> dict:
    """max |mixed - fp64| / max |fp64| of the per-slab LDOS, ndos, pdos
    (-2 Im Tr G^R, Im Tr G^<, -Im Tr G^>, negf_mod.f90:143-149) and of the
    contact currents tr, tre."""
    off = np.cumsum([0] + list(nm))
    out = {}
    for name, key, coef in (("ldos", "gr_diag", -2.0), ("ndos", "gl_diag", 1.0), ("pdos", "gg_diag", -1.0)):
        a = np.stack([mix[key][:, off[i]:off[i + 1]].sum(axis=1).imag for i in range(len(nm))], 1) * coef
        b = np.stack([ref[key][:, off[i]:off[i + 1]].sum(axis=1).imag for i in range(len(nm))], 1) * coef
        out[name] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    for key in ("tr", "tre"):
        out[key] = float(np.abs(mix[key] - ref[key]).max() / max(np.abs(ref[key]).max(), 1e-300))
    return out


def run_mixed_leg(args, rgf, H, Eall, Kall, nb: int, nx: int, fp64_value: float, digits: int = 6) -> dict:
    """The same workload with QT_PREC_MIXED (digits = 6: sweep products on the
    int8 matrix cores, operands rounded to 40 bits), QT_PREC_MIXED47
    (digits = 7, 47 bits) or QT_PREC_FP64_3M (digits = 0: FP64 products with
    three real matrix products per complex one); leads and inversions FP64:
    timed like the FP64 leg, and its error measured on 32 points spread over
    the grid."""
    B = args.batch
    prec = {6: rgf.QT_PREC_MIXED, 7: rgf.QT_PREC_MIXED47, 0: rgf.QT_PREC_FP64_3M}[digits]
    elapsed, acc = time_batches(rgf, H, Eall, Kall, B, nb * nx, args.steps, args.warmup, precision=prec)
    value = B * args.steps / elapsed
    idx = np.linspace(0, len(Eall) - 1, 32).astype(int)
    ref = rgf.solve_batch(H, Eall[idx], Kall[idx], mul=0.0, mur=-0.1)
    mix = rgf.solve_batch(H, Eall[idx], Kall[idx], mul=0.0, mur=-0.1, precision=prec)
    errs = spectra_errors(mix, ref, [nb] * nx)
    return {"value": round(value, 4), "unit": "energy points/s",
            "speedup_vs_fp64": round(value / fp64_value, 3),
            "ms_per_step": round(1e3 * elapsed / args.steps, 3),
            "gemm_ms_per_launch": round(acc["gemm_ms"] / max(1, acc["gemm_launches"]), 4),
            "max_rel_err": max(errs.values()), "errors": errs, "bar": 1e-8,
            "meets_bar": bool(max(errs.values()) <= 1e-8),
            "error_sample": "32 points evenly spread over the 4096-point (E, kz) list, against the FP64 run",
            "products": (f"int8 MFMA, {digits} base-128 digits per value ({7 * digits - 2}-bit operands), "
                         "Karatsuba complex product") if digits else
                        "FP64 MFMA, three real products per complex one ((Ar+Ai)(Br+Bi) - ArBr - AiBi)"}


def run_leads(args, hm, rgf) -> dict:
    """Contact surface Green's functions alone (the lead stage of every point;
    the shape a learned warm start would be measured against): lead block
    nb = 256 from the same generator, 4096 energies in [-3, 3] eV x 16 kz in
    [0, pi), eta = 1e-6, plain Sancho-Rubio (rgf_mod.f90:482-568, stop test
    sum |H01| < 1e-100, at most 50 steps) through qt_sancho.  A step is one
    batch of 4 x --batch points spread over the (E, kz) list.  qt_sancho takes
    host blocks, so the copies of H00/H10 in and G00 out are inside the timed
    region; path.kernel_share_of_wall is the GEMM + inversion kernel time over
    the wall time."""
    nb, B = 256, args.batch * 4
    kz = np.arange(16) * np.pi / 16
    w = hm.synthetic_wannier(nb, 1003, tz_density=0.02)
    H00s, H10s = zip(*[w.mat_def(0.0, 0.0, q, 1) for q in kz])
    E = hm.energy_grid(-3.0, 3.0, 4096)
    npts = 4096 * 16
    rgf.init(0)
    cursor = [0]

    def batch():
        idx = (np.arange(B) * 257 + cursor[0] * 7919) % npts
        cursor[0] += 1
        ie, ik = idx // 16, idx % 16
        return E[ie], [H00s[k] for k in ik], [H10s[k] for k in ik]

    for _ in range(args.warmup):
        En, h0, h1 = batch()
        rgf.sancho(h0, h1, En, eta=1e-6, want_gbb=False)
    rgf.set_kernel_timing(True)
    acc = {"gemm_ms": 0.0, "gemm_flops": 0.0, "inv_ms": 0.0, "inv_flops": 0.0}
    iters, conv = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        En, h0, h1 = batch()
        _, _, it, cv = rgf.sancho(h0, h1, En, eta=1e-6, want_gbb=False)
        t = rgf.last_timing()
        for k in acc:
            acc[k] += t[k]
        iters.append(it)
        conv.append(cv)
    elapsed = time.perf_counter() - t0
    rgf.set_kernel_timing(False)
    it = np.concatenate(iters)
    cv = np.concatenate(conv)
    points = B * args.steps
    n3 = float(nb) ** 3
    flops = float(np.sum(it * 56 + 16)) * n3  # K (INV + 6 GEMM) + INV + 1 GEMM (G00 and the final product)
    gemm_tf = acc["gemm_flops"] / (acc["gemm_ms"] * 1e-3) / 1e12 if acc["gemm_ms"] > 0 else 0.0
    res = {
        "metric": "lead surface Green's functions per second (plain Sancho-Rubio, nb=256, eta=1e-6)",
        "value": round(points / elapsed, 2), "unit": "lead points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 generator)",
        "config": {"workload": f"lead decimation, nb={nb}, 4096 E in [-3,3] eV x 16 kz, eta=1e-6, "
                               f"{B} points per step", "nb": nb, "points_per_step": B},
        "roofline": {"bound": "mfma", "achieved": round(gemm_tf, 3), "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": round(gemm_tf / PEAK_FP64, 4), "traffic": None,
                     "kernel": "zgemm_mfma_*_kernel (batched complex128 GEMM, autotuned tile; decimation products)"},
        "path": {"iterations_avg": round(float(it.mean()), 2), "iterations_max": int(it.max()),
                 "not_converged": int((~cv).sum()),
                 "algorithmic_tflops": round(flops / elapsed / 1e12, 3),
                 "inversion_tflops": round(acc["inv_flops"] / (acc["inv_ms"] * 1e-3) / 1e12, 3)
                 if acc["inv_ms"] > 0 else 0.0,
                 "kernel_share_of_wall": round((acc["gemm_ms"] + acc["inv_ms"]) / (1e3 * elapsed), 3)},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        import concurrent.futures as cf

        import oracle
        flavour = "mkl" if oracle.available("mkl") else "plain"
        if flavour == "mkl":  # one MKL thread per call; the points run in parallel threads
            for name in ("MKL_Set_Num_Threads", "mkl_set_num_threads_"):
                fn = getattr(oracle.lib("mkl"), name, None)
                if fn is not None:
                    fn(C.c_int(1)) if name[0] == "M" else fn(C.byref(C.c_int(1)))
                    break
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sample = [(float(E[i * 509 % 4096]), (i * 5) % 16) for i in range(16 * threads)]
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda ek: oracle.sancho(None, H00s[ek[1]], H10s[ek[1]], ek[0], eta=1e-6,
                                                 want_gbb=False, flavour=flavour), sample))
        tc = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(len(sample) / tc, 4), "unit": "lead points/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{len(sample)} points on {threads} threads ({flavour} LAPACK/BLAS, "
                                         f"one BLAS thread each), {tc:.1f}s"}
    return res


def run_warmstart(args, hm, rgf) -> dict:
    """Learned lead warm start at the lead shape of the leads workload (nb =
    256, 4096 energies in [-3, 3] eV x 16 kz, eta = 1e-6).  The run solves
    its first 10% of points with plain Sancho-Rubio (every 10th energy, all
    kz: a stride through the grid, so the network interpolates), trains the
    network on the device on them, and then solves a strided sample of
    --warm-points of the remaining points three ways: plain Sancho-Rubio,
    Newton from the network's g0, Newton from the bare layer's (z - H00)^-1
    (both Newton modes with the Sancho fallback).  value = lead points/s of
    the warm-started mode (network + Newton + fallback, device time)."""
    nb = 256
    kz = np.arange(16) * np.pi / 16
    w = hm.synthetic_wannier(nb, 1003, tz_density=0.02)
    H00s, H10s = zip(*[w.mat_def(0.0, 0.0, q, 1) for q in kz])
    E = hm.energy_grid(-3.0, 3.0, 4096)
    rgf.init(0)
    lead = rgf.Lead(list(H00s), list(H10s), kz, 1e-6, -3.0, 3.0, batch=128, epochs=args.warm_epochs)
    train = np.zeros(len(E), bool)
    train[::10] = True
    Et, kt = np.repeat(E[train], 16), np.tile(np.arange(16), int(train.sum()))
    t0 = time.perf_counter()
    loss = lead.train(Et, kt)
    t_train = time.perf_counter() - t0
    rest = np.where(~train)[0]
    pts = (np.arange(args.warm_points) * STRIDE) % (len(rest) * 16)
    Ee, ke = E[rest[pts // 16]], (pts % 16).astype(np.int32)
    modes = {}
    for mode in ("sancho", "newton", "bare"):
        r = lead.surface(mode, Ee, ke, compare=(mode != "sancho"))
        rep = r["report"]
        conv = r["fallback"] == 0
        modes[mode] = {
            "points_per_s": round(rep["points"] / (rep["total_ms"] * 1e-3), 2),
            "ms_per_point": round(rep["total_ms"] / rep["points"], 4),
            "network_ms": round(rep["network_ms"], 2), "newton_ms": round(rep["newton_ms"], 2),
            "fallback_ms": round(rep["fallback_ms"], 2),
            "fallback_rate": round(rep["fallbacks"] / rep["points"], 4) if mode != "sancho" else None,
            "newton_steps_per_point": round(rep["newton_steps"] / rep["points"], 3),
            "smith_doublings_per_point": round(rep["smith_doublings"] / rep["points"], 3),
            "sancho_iterations_per_point": round(rep["sancho_iterations"] / rep["points"], 2),
            "warm_start_err_median": round(float(np.median(r["warm_err"])), 4) if mode != "sancho" else None,
            "g00_err_vs_sancho_max": float(r["sancho_err"].max()) if mode != "sancho" else 0.0,
            "converged_points": int(conv.sum()) if mode != "sancho" else None,
        }
    lead.close()
    ws = modes["newton"]
    return {"metric": "lead surface Green's functions/s @nb=256, eta=1e-6: learned warm start vs Sancho-Rubio",
            "value": ws["points_per_s"], "unit": "lead points/s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": round(ws["ms_per_point"] * args.warm_points, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64 (network bf16, f32 accumulation)",
            "data": "synthetic (seeded SplitMix64 generator)",
            "config": {"workload": "learned lead warm start, nb=256, 4096 E in [-3,3] eV x 16 kz, eta=1e-6",
                       "nb": nb, "energies": 4096, "kz": 16, "train_points": int(len(Et)),
                       "eval_points": int(args.warm_points)},
            "training": {"seconds": round(t_train, 2), "epochs": int(args.warm_epochs),
                         "loss_first": round(float(loss[0]), 4), "loss_last": round(float(loss[-1]), 5)},
            "ablation": modes,
            "speed_vs_sancho": round(ws["points_per_s"] / modes["sancho"]["points_per_s"], 3),
            "cpu_baseline": None}


def run_uneven(args, hm, rgf) -> dict:
    """Uneven slabs: 80 slabs with sizes drawn uniformly from {128, 192, ..., 768}
    (direct block construction), 2048 energies in [-2, 2] eV x 8 kz on one
    k-point set (the 8 kz replicate the same blocks), plain FP64."""
    rng = np.random.default_rng(args.seed)
    sizes = [int(v) for v in rng.choice(np.arange(128, 769, 64), size=80)]
    dev = hm.uneven_device(sizes, args.seed)
    grid = hm.energy_grid(-2.0, 2.0, 2048)
    Eall = np.repeat(grid, 8)
    Kall = np.zeros(len(Eall), np.int32)
    rgf.init(0)
    H = rgf.Hamiltonian.from_device(dev)
    nd = int(sum(sizes))
    B = args.batch
    # uneven slabs need the coupling of the slab sizes (the reference indexes
    # H_{i,i+1} one slab off in its second pass, which only works for equal sizes)
    elapsed, acc = time_batches(rgf, H, Eall, Kall, B, nd, args.steps, args.warmup, coupling_fix=True)
    H.close()
    points = B * args.steps
    gemm_tf = acc["gemm_flops"] / (acc["gemm_ms"] * 1e-3) / 1e12 if acc["gemm_ms"] > 0 else 0.0
    inv_tf = acc["inv_flops"] / (acc["inv_ms"] * 1e-3) / 1e12 if acc["inv_ms"] > 0 else 0.0
    K = acc["lead_iterations"] / max(1, 2 * points)
    # 112 n^3 per slab (neighbours of unequal size) with n^3 standing for the
    # n_i n_{i-1} terms; leads at the end slab sizes
    sweep = sum(float(SLAB_FLOPS_UNEVEN) * float(n) ** 3 for n in sizes)
    lead = (K * 56 + 24) * (float(sizes[0]) ** 3 + float(sizes[-1]) ** 3)
    return {
        "metric": "RGF energy-points/s, uneven slabs {128..768} x 80, FP64",
        "value": round(points / elapsed, 4), "unit": "energy points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 blocks, numpy-drawn slab sizes)",
        "config": {"workload": f"ballistic RGF, 80 uneven slabs (sum n = {nd}), 8 kz x 2048 E in "
                               f"[-2,2] eV, G^R/G^</G^> diagonals + contact currents, plain FP64",
                   "slabs": 80, "sum_n": nd, "points_per_step": B, "precision": "fp64"},
        "roofline": {"bound": "mfma", "achieved": round(gemm_tf, 3), "peak": PEAK_FP64,
                     "unit": "TFLOP/s", "frac": round(gemm_tf / PEAK_FP64, 4), "traffic": None,
                     "kernel": "zgemm_mfma_*_kernel (batched complex128 GEMM, autotuned tile)",
                     "launches": acc["gemm_launches"]},
        "sample": acc["sample"],
        "path": {"lead_iterations_avg": round(K, 2),
                 "algorithmic_tflops_approx": round(points * (sweep + lead) / elapsed / 1e12, 3),
                 "inversion_tflops": round(inv_tf, 3),
                 "gemm_share_of_wall": round(acc["gemm_ms"] / (1e3 * elapsed), 3),
                 "inversion_share_of_wall": round(acc["inv_ms"] / (1e3 * elapsed), 3)},
        "cpu_baseline": None,
    }


def run_small(args, hm, rgf) -> dict:
    """The reference's CPU-case shape: nb=64 x 32 identical slabs, 256 energies
    in [-3, 3] eV, one kz, lead eta 1e-5; each step is the whole job (all 256
    points in one device batch), G^R/G^</G^> diagonals + contact currents."""
    nb, nx, nen = 64, 32, 256
    _, dev = hm.synthetic_device(nb, nx, seed=args.seed)
    Eall = hm.energy_grid(-3.0, 3.0, nen)
    Kall = np.zeros(nen, np.int32)
    rgf.init(0)
    H = rgf.Hamiltonian.from_device(dev)
    # kernel breakdown (HIP events around every product and inversion), then
    # the throughput from a run without the events
    _, acc = time_batches(rgf, H, Eall, Kall, nen, nb * nx, args.steps, args.warmup)
    bufs = {"tr": rgf.DeviceBuffer(8 * nen), "tre": rgf.DeviceBuffer(8 * nen)}
    for key in ("gr_diag", "gl_diag", "gg_diag"):
        bufs[key] = rgf.DeviceBuffer(16 * nen * nb * nx)
    step = lambda: rgf.solve_batch_device(H, Eall, Kall, bufs, mul=0.0, mur=-0.1)
    for _ in range(args.warmup):
        step()
    rgf.lib().qt_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    rgf.lib().qt_synchronize()
    elapsed = time.perf_counter() - t0
    for b in bufs.values():
        b.free()
    H.close()
    points = nen * args.steps
    gemm_tf = acc["gemm_flops"] / (acc["gemm_ms"] * 1e-3) / 1e12 if acc["gemm_ms"] > 0 else 0.0
    inv_tf = acc["inv_flops"] / (acc["inv_ms"] * 1e-3) / 1e12 if acc["inv_ms"] > 0 else 0.0
    K = acc["lead_iterations"] / max(1, 2 * points)
    n3 = float(nb) ** 3
    path_flops = nx * SLAB_FLOPS_EQUAL * n3 + 2 * (K * 56 + 24) * n3
    return {
        "metric": "RGF energy-points/s, nb=64 x 32 slabs, 256 E, FP64",
        "value": round(points / elapsed, 2), "unit": "energy points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 generator)",
        "config": {"workload": "ballistic RGF, nb=64 x 32 slabs, 256 E in [-3,3] eV, one kz, "
                               "G^R/G^</G^> diagonals + contact currents, plain FP64",
                   "nb": nb, "slabs": nx, "energies": nen, "points_per_step": nen, "precision": "fp64"},
        "roofline": {"bound": "mfma", "achieved": round(gemm_tf, 3), "peak": PEAK_FP64,
                     "unit": "TFLOP/s", "frac": round(gemm_tf / PEAK_FP64, 4), "traffic": None,
                     "kernel": "zgemm_mfma_*_kernel (batched complex128 GEMM, autotuned tile)",
                     "launches": acc["gemm_launches"]},
        "path": {"lead_iterations_avg": round(K, 2),
                 "algorithmic_tflops": round(points * path_flops / elapsed / 1e12, 3),
                 "inversion_tflops": round(inv_tf, 3),
                 "gemm_share_of_wall": round(acc["gemm_ms"] / (1e3 * elapsed), 3),
                 "inversion_share_of_wall": round(acc["inv_ms"] / (1e3 * elapsed), 3),
                 "lead_ms": round(acc["lead_ms"], 1), "forward_ms": round(acc["forward_ms"], 1),
                 "backward_ms": round(acc["backward_ms"], 1)},
        "cpu_baseline": None,
    }


def run_scba(args) -> dict:
    """One GW-SCBA iteration on the device (qt_gw_scba_step: G of every energy
    by the green_RGF_RS recursion with the current Sigma, P, W, Sigma, mixing;
    every array device resident) at the GW shape's block size and energy
    axis: nm = 256 orbitals per slab, 2048 energies in [-4, 4] eV, one kz, on
    --scba-slabs slabs (default 6: the element-major arrays of the 50-slab
    device would need 15 x 107 GB), bare Coulomb blocks with eps = 4 and
    r0 = 0.3 A, P and Sigma on the diagonal blocks (ndiag = nm), mixing 0.5.
    A step is one iteration; the first (warm-up) starts from Sigma = 0."""
    from quantum_transport_amd import hamiltonian as hm
    from quantum_transport_amd import rgf

    nm, nx, nen = 256, args.scba_slabs, 2048
    w = hm.synthetic_wannier(nm, args.seed)
    dev = hm.device_from_wannier(w, nx, 1, [(0.0, 0.0)])
    Vd, V1 = w.coulomb_blocks(0.0, 0.0, 0.0, 4.0, 0.3)
    Vii = np.repeat(np.asarray(Vd)[:, :, None], nx, axis=2)
    V1i = np.repeat(np.asarray(V1)[:, :, None], nx, axis=2)
    E = hm.energy_grid(-4.0, 4.0, nen)
    rgf.init(0)
    H = rgf.Hamiltonian.from_device(dev)
    st = rgf.GWState(nm, nx, nen, Vii, V1i)
    mu, T = [0.0, -0.1], [300.0, 300.0]
    for _ in range(args.warmup):
        rgf.gw_scba_step(H, st, E, mu, T, nm, 2.0, 0.5)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tr, tre, status = rgf.gw_scba_step(H, st, E, mu, T, nm, 2.0, 0.5)
    el = (time.perf_counter() - t0) / max(1, args.steps)
    st.free()
    H.close()
    return {"metric": "GW-SCBA iterations: energy x slab points per second (G, P, W, Sigma, mixing)",
            "value": round(nen * nx / el, 1), "unit": "energy-slab points/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * el, 1),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded SplitMix64 blocks, bare Coulomb on the orbital centres)",
            "config": {"workload": f"GW-SCBA iteration, nm=256, {nx} slabs, 2048 E in [-4,4] eV, 1 kz, "
                                   "ndiag=nm, eps=4, r0=0.3 A, mixing 0.5",
                       "nm": nm, "slabs": nx, "energies": nen},
            "iteration_s_per_slab": round(el / nx, 3),
            "status": int(status), "current_tr_sum": float(np.sum(tr)), "cpu_baseline": None}


def run_copy(args) -> dict:
    """Device-to-device copies of 256 MiB (torch's copy kernel): the plain
    copy of known size against which FETCH_SIZE / WRITE_SIZE are scaled, and
    the copy bandwidth the HBM actually sustains."""
    import torch

    n = 256 << 20
    dev = torch.device("cuda", 0)
    a = torch.ones(n, dtype=torch.uint8, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    reps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"metric": "device-to-device copy bandwidth (read + write)", "value": round(2 * n / dt / 1e9, 1),
            "unit": "GB/s", "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": round(1e3 * dt, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
            "data": "synthetic (ones)", "config": {"workload": "256 MiB device-to-device copy", "bytes": n},
            "cpu_baseline": None}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64, help="energy points per step")
    ap.add_argument("--nb", type=int, default=512)
    ap.add_argument("--slabs", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--seed", type=int, default=1002)
    ap.add_argument("--workload", choices=("rgf", "uneven", "leads", "warmstart", "gw", "corrconv", "ephoton",
                                               "vendor-zgemm", "copy", "scba", "small"),
                    default="rgf",
                    help="rgf: the headline RGF sweep; uneven: 80 slabs of 128..768 orbitals; "
                         "leads: Sancho-Rubio surface Green's functions at nb=256, eta=1e-6; "
                         "warmstart: the learned lead warm start against Sancho-Rubio at that shape; "
                         "gw: the GW energy-axis stage; corrconv: batched corr1d2/conv1d2; "
                         "vendor-zgemm: rocBLAS batched ZGEMM at the sweep's shape, for comparison; "
                         "copy: 256 MiB device-to-device copies, the HBM counter calibration; "
                         "scba: one GW-SCBA iteration on the device at nm=256, 2048 E; "
                         "small: the reference's CPU-case shape, nb=64 x 32 slabs x 256 E")
    ap.add_argument("--gw-slabs", type=int, default=2)
    ap.add_argument("--scba-slabs", type=int, default=6, help="scba: slabs of the device")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the legs of the other precision modes (mixed, mixed47, fp64_3m)")
    ap.add_argument("--warm-points", type=int, default=1024, help="warmstart: evaluated points")
    ap.add_argument("--warm-epochs", type=int, default=30, help="warmstart: training epochs")
    ap.add_argument("--cpu-validate", action="store_true",
                    help="also time the CPU baseline's points at the full slab count")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the RGF path is measured per GPU "
              "(energy points shard across GPUs with no communication but a final sum); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    if args.workload == "gw":
        print(json.dumps(run_gw(args)), flush=True)
        return 0
    if args.workload == "corrconv":
        print(json.dumps(run_corrconv(args)), flush=True)
        return 0
    if args.workload == "ephoton":
        print(json.dumps(run_ephoton(args)), flush=True)
        return 0
    if args.workload == "vendor-zgemm":
        print(json.dumps(run_vendor_zgemm(args)), flush=True)
        return 0
    if args.workload == "copy":
        print(json.dumps(run_copy(args)), flush=True)
        return 0
    if args.workload == "scba":
        print(json.dumps(run_scba(args)), flush=True)
        return 0

    from quantum_transport_amd import hamiltonian as hm
    from quantum_transport_amd import rgf

    if args.workload == "uneven":
        print(json.dumps(run_uneven(args, hm, rgf)), flush=True)
        return 0
    if args.workload == "leads":
        print(json.dumps(run_leads(args, hm, rgf)), flush=True)
        return 0
    if args.workload == "warmstart":
        print(json.dumps(run_warmstart(args, hm, rgf)), flush=True)
        return 0
    if args.workload == "small":
        print(json.dumps(run_small(args, hm, rgf)), flush=True)
        return 0

    nb, nx, B = args.nb, args.slabs, args.batch
    kz = [0.0, np.pi / 4, np.pi / 2, 3 * np.pi / 4]
    w = hm.synthetic_wannier(nb, args.seed, tz_density=0.02)
    dev = hm.device_from_wannier(w, nx, 1, [(0.0, q) for q in kz])
    grid = hm.energy_grid(-2.0, 2.0, 1024)
    Eall = np.repeat(grid, 4)
    Kall = np.tile(np.arange(4, dtype=np.int32), 1024)

    rgf.init(0)
    H = rgf.Hamiltonian.from_device(dev)
    del dev
    elapsed, acc = time_batches(rgf, H, Eall, Kall, B, nb * nx, args.steps, args.warmup)
    points = B * args.steps
    value = points / elapsed

    gemm_tf = acc["gemm_flops"] / (acc["gemm_ms"] * 1e-3) / 1e12 if acc["gemm_ms"] > 0 else 0.0
    inv_tf = acc["inv_flops"] / (acc["inv_ms"] * 1e-3) / 1e12 if acc["inv_ms"] > 0 else 0.0
    K = acc["lead_iterations"] / max(1, 2 * points)
    n3 = float(nb) ** 3
    lead_flops = (K * 56 + 24) * n3  # per lead: K (INV + 6 GEMM) + INV + 2 GEMM
    path_flops = nx * SLAB_FLOPS_EQUAL * n3 + 2 * lead_flops
    res = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "energy points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 generator)",
        "config": {"workload": f"ballistic RGF, nb={nb} x {nx} slabs, 4 kz x 1024 E in [-2,2] eV, "
                               f"G^R/G^</G^> diagonals + contact currents, plain FP64",
                   "nb": nb, "slabs": nx, "kz": 4, "energies": 1024, "points_per_step": B,
                   "precision": "fp64"},
        "roofline": {"bound": "mfma", "achieved": round(gemm_tf, 3), "peak": PEAK_FP64,
                     "unit": "TFLOP/s", "frac": round(gemm_tf / PEAK_FP64, 4),
                     "traffic": (GEMM_BYTES_PER_MEMBER_512 * B) if nb == 512 else None,
                     "kernel": "zgemm_mfma_*_kernel (batched complex128 GEMM, autotuned tile)",
                     "launches": acc["gemm_launches"]},
        "sample": acc["sample"],
        "path": {"lead_iterations_avg": round(K, 2),
                 "algorithmic_tflops": round(points * path_flops / elapsed / 1e12, 3),
                 "inversion_tflops": round(inv_tf, 3),
                 "gemm_share_of_wall": round(acc["gemm_ms"] / (1e3 * elapsed), 3),
                 "inversion_share_of_wall": round(acc["inv_ms"] / (1e3 * elapsed), 3),
                 "lead_ms": round(acc["lead_ms"], 1), "forward_ms": round(acc["forward_ms"], 1),
                 "backward_ms": round(acc["backward_ms"], 1)},
        "cpu_baseline": None,
    }
    if not args.no_mixed:
        res["mixed"] = run_mixed_leg(args, rgf, H, Eall, Kall, nb, nx, value)
        res["mixed47"] = run_mixed_leg(args, rgf, H, Eall, Kall, nb, nx, value, digits=7)
        res["fp64_3m"] = run_mixed_leg(args, rgf, H, Eall, Kall, nb, nx, value, digits=0)
    H.close()
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        res["cpu_baseline"] = cpu_baseline(nb, nx, threads, args.seed, validate=args.cpu_validate)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
