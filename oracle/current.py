"""CPU restatement of the cell-bond currents and collision integrals of the
GW loop.

TEST INFRASTRUCTURE ONLY: the checker of the library's qt_gw_block_current
and qt_gw_block_collision; imported by tests/, never by the product.

The reference calls calc_block_current (gw_mod.f90:384, :1001) and
calc_block_collision (gw_mod.f90:795, :926) with the argument lists kept
here, but neither routine is in its sources, so these are the definitions
(from NEGF theory; pinned in tests/test_oracle_current.py by current
conservation through every bond of a ballistic device, by a dense
whole-device G^< and by the equilibrium identity of the collision integral):

* ``block_current``: for cells of nb orbitals, ns per slab (nm = nb ns), the
  current matrix of bond c = (slab l, cell j) -> next cell,
    j < ns:  Re H_{j,j+1} G^<_{j+1,j} - Re H_{j+1,j} G^<_{j,j+1}
             (nb x nb products of the cell blocks of Hii(l) and G^<(l));
    j = ns:  the sum of the ns diagonal nb x nb blocks of green_RGF_RS's
             slab-bond matrix cur(l) (gw_mod.f90:1250-1273), zero for l = nx;
  its trace is the particle current through the bond.  Energy integrals
  with the reference's terminal-current weight (gw_mod.f90:424-428)
  w = dE e0 / (2 pi hbar) e0 spindeg: tot_cur = sum_e jdens(e) w,
  tot_ecur = sum_e E_e jdens(e) w.
* ``block_collision``: Ispec(E) = Sigma^<(E) G^>(E) - Sigma^>(E) G^<(E) per
  slab block, Itot = sum_e Ispec(e) dE spindeg / (2 pi).
"""
from __future__ import annotations

import numpy as np

E0, HBAR, TPI = 1.6022e-19, 1.0546e-34, 3.14159265359 * 2.0  # static.f90:10,14 and fft_mod.f90:19 there


def current_weight(En, spindeg: float) -> float:
    dE = En[1] - En[0] if len(En) > 1 else 0.0
    return dE * E0 / TPI / HBAR * E0 * spindeg


def block_current(Hii, G_lesser, cur, En, spindeg: float, nb: int, ns: int):
    """Hii (nm, nm, nx); G_lesser, cur (nm, nm, nx, nen) -> jdens (nb, nb,
    nx ns, nen), tot_cur, tot_ecur (nb, nb, nx ns)."""
    Hii = np.asarray(Hii)
    GL = np.asarray(G_lesser)
    cur = np.asarray(cur)
    nm, _, nx, nen = GL.shape
    assert nm == nb * ns
    jd = np.zeros((nb, nb, nx * ns, nen))
    blk = lambda M, i, j: M[i * nb:(i + 1) * nb, j * nb:(j + 1) * nb]
    for e in range(nen):
        for l in range(nx):
            H, G = Hii[:, :, l], GL[:, :, l, e]
            for j in range(ns):
                c = l * ns + j
                if j < ns - 1:
                    jd[:, :, c, e] = (blk(H, j, j + 1) @ blk(G, j + 1, j)).real \
                        - (blk(H, j + 1, j) @ blk(G, j, j + 1)).real
                elif l < nx - 1:
                    jd[:, :, c, e] = sum(blk(cur[:, :, l, e], k, k) for k in range(ns))
    w = current_weight(En, spindeg)
    tot = np.zeros((nb, nb, nx * ns))
    tote = np.zeros((nb, nb, nx * ns))
    for e in range(nen):  # energies in order, as the device sums them
        tot += jd[:, :, :, e] * w
        tote += jd[:, :, :, e] * (w * En[e])
    return jd, tot, tote


def block_collision(Sl, Sg, Gl, Gg, En, spindeg: float):
    """(nm, nm, nx, nen) blocks -> Itot (nm, nm, nx), Ispec (nm, nm, nx, nen)."""
    Sl, Sg, Gl, Gg = (np.asarray(x) for x in (Sl, Sg, Gl, Gg))
    nm, _, nx, nen = Sl.shape
    Ispec = np.zeros((nm, nm, nx, nen), dtype=complex)
    for e in range(nen):
        for l in range(nx):
            Ispec[:, :, l, e] = Sl[:, :, l, e] @ Gg[:, :, l, e] - Sg[:, :, l, e] @ Gl[:, :, l, e]
    f = (En[1] - En[0] if len(En) > 1 else 0.0) * spindeg / TPI
    Itot = np.zeros((nm, nm, nx), dtype=complex)
    for e in range(nen):
        Itot += Ispec[:, :, :, e] * f
    return Itot, Ispec
