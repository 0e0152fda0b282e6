"""GPU W solve (W^R = V (I - P^R V)^-1, W^<,> = W^R P^<,> W^R^H; the routine
called at gw_mod.f90:597) against the oracle, which is pinned by dense
inverses in test_oracle.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf
    rgf.init(0)
    return rgf


@pytest.mark.parametrize("nm,nx,nw", [(1, 1, 2), (3, 2, 3), (4, 6, 5), (16, 5, 4), (64, 4, 3), (5, 7, 9)])
def test_wsolve_matches_oracle(qt, oracle_mod, nm, nx, nw):
    print("wsolve", nm, nx, nw, flush=True)
    rng = np.random.default_rng(nm * 100 + nx)
    cr = lambda *sh, s=1.0: s * (rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    Vii = cr(nm, nm, nx)
    Vii = (Vii + Vii.conj().transpose(1, 0, 2)) / 2
    V1i = cr(nm, nm, nx, s=0.3)
    PR, PL, PG = (cr(nm, nm, nx, nw, s=0.2 / np.sqrt(nm)) for _ in range(3))
    WR, WL, WG = qt.gw_wsolve(Vii, V1i, PR, PL, PG)
    for w in range(nw):
        ref = oracle_mod.gw_wsolve(Vii, V1i, PR[..., w], PL[..., w], PG[..., w])
        for got, r, name in zip((WR[..., w], WL[..., w], WG[..., w]), ref, "R<>"):
            scale = np.abs(r).max()
            assert np.abs(got - r).max() <= 1e-10 * scale, (name, w)
