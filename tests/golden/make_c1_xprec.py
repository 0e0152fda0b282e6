"""Writes c1_xprec.npz: the spectra of the nb=64 x 32-slab device (seed
1001, 256 energies in [-3, 3] eV, mul = 0, mur = -0.1 eV, 300 K, forward
variant) from the extended-precision sweep in tests/_xprec.py.

Stored per energy: LDOS, ndos, pdos per slab (the reductions of
negf_mod.f90:141-149: -2 Im Tr G^R, Im Tr G^<, -Im Tr G^> over the slab's
orbitals) and the contact currents tr, tre.  Run from the repository root:
    python tests/golden/make_c1_xprec.py  [processes]
"""
import os
import sys
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from quantum_transport_amd import hamiltonian as hm  # noqa: E402
from tests._xprec import xrgf_forward  # noqa: E402

NB, NX, NE = 64, 32, 256


def spectra(r):
    off = np.arange(NX + 1) * NB
    red = lambda d, c: np.stack([c * d[:, off[s]:off[s + 1]].sum(1).imag for s in range(NX)], 1)
    return dict(ldos=red(r["gr_diag"], -2.0), ndos=red(r["gl_diag"], 1.0), pdos=red(r["gg_diag"], -1.0),
                tr=r["tr"], tre=r["tre"])


def chunk(ix):
    _, dev = hm.synthetic_device(NB, NX, seed=1001)
    E = hm.energy_grid(-3.0, 3.0, NE)[ix]
    return ix, spectra(xrgf_forward(dev.Hii[0], dev.H1i[0], E, 0.0, -0.1))


def main():
    procs = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    parts = np.array_split(np.arange(NE), 4 * procs)
    out = {k: np.zeros((NE, NX)) for k in ("ldos", "ndos", "pdos")}
    out.update(tr=np.zeros(NE), tre=np.zeros(NE))
    with Pool(procs) as pool:
        for ix, r in pool.imap_unordered(chunk, parts):
            for k in out:
                out[k][ix] = r[k]
    out["E"] = hm.energy_grid(-3.0, 3.0, NE)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "c1_xprec.npz"), **out)


if __name__ == "__main__":
    main()
