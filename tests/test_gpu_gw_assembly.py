"""GPU parity of the GW polarisation / self-energy assembly and mixing
(gw_mod.f90:444-749) against the oracle, whose restatement is pinned by
numpy convolutions and the momentum-pairing rule in test_oracle.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # nm, nx, nky, nkz, nen, ndiag
    (3, 2, 1, 1, 16, 1),
    (4, 2, 1, 4, 33, 2),
    (2, 2, 2, 2, 24, 1),
    (2, 1, 1, 1, 1000, 1),
    (2, 1, 1, 1, 2048, 0),
    (6, 3, 1, 1, 200, 5),
]


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf
    rgf.init(0)
    return rgf


def _arrays(seed, nm, nx, nk, nen, count):
    rng = np.random.default_rng(seed)
    sh = (nm * nm, nx, nk, nen)
    return [rng.standard_normal(sh) + 1j * rng.standard_normal(sh) for _ in range(count)]


def _close(got, ref, what):
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max()
    assert err <= 1e-12 * scale, (what, err, scale)


@pytest.mark.parametrize("case", CASES)
def test_polarization_matches_oracle(qt, oracle_mod, case):
    nm, nx, nky, nkz, nen, ndiag = case
    print("gw polarization", case, flush=True)
    Gr, Gl, Gg = _arrays(sum(case), nm, nx, nky * nkz, nen, 3)
    dEn, spin = 0.0125, 2.0
    got = qt.gw_polarization(Gr, Gl, Gg, nky, nkz, ndiag, dEn, spin)
    ref = oracle_mod.gw_polarization(Gr, Gl, Gg, nky, nkz, ndiag, dEn, spin)
    for g, r, name in zip(got, ref, ("P^R", "P^<", "P^>")):
        _close(g, r, name)
        # out-of-band entries are exactly zero
        for col in range(nm):
            for row in range(nm):
                if abs(row - col) > ndiag:
                    assert not g[row + col * nm].any()


@pytest.mark.parametrize("case", CASES)
def test_selfenergy_matches_oracle(qt, oracle_mod, case):
    nm, nx, nky, nkz, nen, ndiag = case
    print("gw self-energy", case, flush=True)
    arrs = _arrays(7 + sum(case), nm, nx, nky * nkz, nen, 6)
    dEn = 0.01
    got = qt.gw_selfenergy(*arrs, nky, nkz, ndiag, dEn)
    ref = oracle_mod.gw_selfenergy(*arrs, nky, nkz, ndiag, dEn)
    for g, r, name in zip(got, ref, ("Sigma^R", "Sigma^<", "Sigma^>")):
        _close(g, r, name)
        # (S + S^T)/2 makes the element pair equal
        g4 = g.reshape(nm, nm, nx, nky * nkz, nen)
        assert np.array_equal(g4, g4.transpose(1, 0, 2, 3, 4))


@pytest.mark.parametrize("budget", ["1", "230000"])
def test_slab_chunks_match_whole_call(qt, oracle_mod, monkeypatch, budget):
    """Host-path calls with the device working set capped (QT_WORKSET_BYTES):
    1 byte processes one slab per chunk, 230 kB three slabs for P and two for
    Sigma (the last chunk shorter) at nm = 3, 5 slabs, 2 k points, 40
    energies.  Same bits as the uncapped
    call, and the oracle's values."""
    nm, nx, nky, nkz, nen, ndiag = 3, 5, 1, 2, 40, 1
    G = _arrays(91, nm, nx, nky * nkz, nen, 3)
    W = _arrays(92, nm, nx, nky * nkz, nen, 3)
    P0 = qt.gw_polarization(*G, nky, nkz, ndiag, 0.02, 2.0)
    S0 = qt.gw_selfenergy(*G, *W, nky, nkz, ndiag, 0.02)
    monkeypatch.setenv("QT_WORKSET_BYTES", budget)
    print("gw chunked, budget", budget, flush=True)
    P1 = qt.gw_polarization(*G, nky, nkz, ndiag, 0.02, 2.0)
    S1 = qt.gw_selfenergy(*G, *W, nky, nkz, ndiag, 0.02)
    for a, b in zip(P0 + S0, P1 + S1):
        assert np.array_equal(a, b)
    for g, r in zip(P1, oracle_mod.gw_polarization(*G, nky, nkz, ndiag, 0.02, 2.0)):
        _close(g, r, "P")
    for g, r in zip(S1, oracle_mod.gw_selfenergy(*G, *W, nky, nkz, ndiag, 0.02)):
        _close(g, r, "Sigma")


def test_mix_matches_oracle(qt, oracle_mod):
    rng = np.random.default_rng(2)
    S = rng.standard_normal(5000) + 1j * rng.standard_normal(5000)
    N = rng.standard_normal(5000) + 1j * rng.standard_normal(5000)
    got = qt.gw_mix(S, N, 0.3)
    np.testing.assert_allclose(got, oracle_mod.gw_mix(S, N, 0.3), rtol=0, atol=1e-14)
