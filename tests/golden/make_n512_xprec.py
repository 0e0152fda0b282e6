"""Writes n512_xprec.npz: the spectra of the production-block-size device
(nb = 512, 4 slabs, seed 1002, kz in {0, pi/2} with the 2%-dense T_z,
16 energies in [-1.9, 1.9] eV alternating between the two kz,
mul = 0, mur = -0.1 eV, 300 K, forward variant) from the extended-precision
sweep in tests/_xprec.py; stored per energy: LDOS, ndos, pdos per slab
(negf_mod.f90:141-149) and tr, tre.  Slow (about 25 minutes on 8 cores):
    python tests/golden/make_n512_xprec.py  [processes]
"""
import os
import sys
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from quantum_transport_amd import hamiltonian as hm  # noqa: E402
from tests._xprec import xrgf_forward  # noqa: E402

NB, NX = 512, 4
KZ = [0.0, np.pi / 2]
E = np.linspace(-1.9, 1.9, 16)
K = np.arange(16) % 2


def spectra(r):
    off = np.arange(NX + 1) * NB
    red = lambda d, c: np.stack([c * d[:, off[s]:off[s + 1]].sum(1).imag for s in range(NX)], 1)
    return dict(ldos=red(r["gr_diag"], -2.0), ndos=red(r["gl_diag"], 1.0), pdos=red(r["gg_diag"], -1.0),
                tr=r["tr"], tre=r["tre"])


def point(i):
    _, dev = hm.synthetic_device(NB, NX, seed=1002, kz=KZ, tz_density=0.02)
    k = int(K[i])
    return i, spectra(xrgf_forward(dev.Hii[k], dev.H1i[k], E[i:i + 1], 0.0, -0.1))


def main():
    procs = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    out = {k: np.zeros((len(E), NX)) for k in ("ldos", "ndos", "pdos")}
    out.update(tr=np.zeros(len(E)), tre=np.zeros(len(E)))
    with Pool(procs) as pool:
        for i, r in pool.imap_unordered(point, range(len(E))):
            for key in out:
                out[key][i] = r[key][0]
            print("point", i, flush=True)
    out["E"] = E
    out["k"] = K
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "n512_xprec.npz"), **out)


if __name__ == "__main__":
    main()
