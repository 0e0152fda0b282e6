"""The learned lead warm start (qt_lead_*): its network against the CPU
restatement (oracle/mlp.py), training on the device, Newton refinement with
the Sancho-Rubio fallback, and the ablation of the network against the same
method without it.  Every surface Green's function returned must equal plain
Sancho-Rubio's (rgf_mod.f90:482-568) to the Newton tolerance."""
import numpy as np
import pytest

from oracle import mlp
from quantum_transport_amd import hamiltonian as hm

pytestmark = pytest.mark.gpu

KZ = [0.0, np.pi / 2, np.pi, 1.5 * np.pi]


@pytest.fixture(scope="module")
def qt():
    from quantum_transport_amd import rgf

    rgf.init(0)
    return rgf


def _lead_blocks(nb=64):
    """(H00, H10) per kz of the first contact (its decimation input)"""
    _, dev = hm.synthetic_device(nb, 3, seed=2004, kz=KZ, tz_density=0.05)
    return [(dev.Hii[k][0], dev.H1i[k][0].conj().T) for k in range(len(KZ))]


def _lead(qt, nb, eta=1e-6, **kw):
    blocks = _lead_blocks(nb)
    return qt.Lead([b[0] for b in blocks], [b[1] for b in blocks], KZ, eta, -3.0, 3.0, **kw)


def test_network_matches_cpu_oracle(qt):
    """Forward pass, loss and the gradient of every parameter from the
    device kernels (BF16 MFMA products, GELU, low-rank loss on the FP64
    GEMM) against the CPU restatement with the same BF16 roundings."""
    nb, B = 64, 32
    lead = _lead(qt, nb)
    rng = np.random.default_rng(0)
    dims = [32, 256, 256, 256, lead.width]
    q = []
    for i in range(4):
        q.append(rng.standard_normal(dims[i] * dims[i + 1]) / np.sqrt(dims[i]))
        q.append(0.1 * rng.standard_normal(dims[i + 1]))
    q = np.concatenate(q).astype(np.float32)
    lead.set_params(q)
    assert np.array_equal(lead.params(), q)
    E = rng.uniform(-3, 3, B)
    k = rng.integers(0, len(KZ), B)
    G = 0.3 * (rng.standard_normal((B, nb, nb)) + 1j * rng.standard_normal((B, nb, nb)))
    Y, loss, grad = lead.eval(E, k, G)
    net = mlp.MLP(dims, q, bf16_products=True)
    Yc = net.forward(mlp.features(E, np.array(KZ)[k], -3.0, 3.0))
    lc, dYc = mlp.loss_and_grad(Yc, G, nb, 32)
    gc = net.backward(dYc)
    ey = np.linalg.norm(Y - Yc) / np.linalg.norm(Yc)
    print(f"output rows: relative difference {ey:.1e}; loss {loss:.8f} vs {lc:.8f}")
    assert ey < 1e-3
    assert abs(loss - lc) <= 1e-5 * lc
    o = 0
    for i in range(4):
        for part, cnt in (("W", dims[i] * dims[i + 1]), ("b", dims[i + 1])):
            e = np.linalg.norm(grad[o:o + cnt] - gc[o:o + cnt]) / np.linalg.norm(gc[o:o + cnt])
            print(f"layer {i} {part}: gradient relative difference {e:.1e}")
            assert e < 5e-3, (i, part, e)
            o += cnt
    lead.close()


def test_newton_from_near_solutions_matches_sancho(qt):
    """Newton's machinery from starts 1e-3 away from the decimation result,
    at energies outside the lead's bands (spectral radius of the Stein
    operator well below one): every point converges within the Newton
    budget to the retarded root, equal to plain Sancho-Rubio."""
    lead = _lead(qt, 8, newton_max=4)
    # (E, kz) pairs where every Bloch factor of the lead decays (spectral
    # radius of g H01 between 0.41 and 0.89; inside a band it is 1 - O(eta))
    pts = [(-7.0, 2), (-6.0, 2), (-5.0, 0), (-5.0, 2), (3.8, 2), (5.0, 0), (5.0, 1), (6.0, 0), (6.0, 1), (6.0, 3)]
    E = np.array([p[0] for p in pts])
    k = np.array([p[1] for p in pts])
    ref = lead.surface("sancho", E, k, blocks=True)["G00"]
    rng = np.random.default_rng(5)
    start = [g * (1 + 1e-3 * (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape))) for g in ref]
    r = lead.surface("given", E, k, compare=True, start=start)
    print("newton steps", r["newton_steps"], "doublings", r["smith_doublings"], "residual", r["residual"])
    assert (r["fallback"] == 0).all()
    assert (r["newton_steps"] <= 4).all() and (r["newton_steps"] >= 1).all()
    assert (r["residual"] <= 1e-10).all()
    assert r["sancho_err"].max() <= 1e-9
    for g, s in zip(r["G00"], ref):
        assert -np.trace(g).imag > 0  # retarded branch
        assert np.abs(g - s).max() <= 1e-9 * np.abs(s).max()
    lead.close()


def test_warm_start_ablation(qt):
    """nb = 64 lead, 4 kz, 256 energies in [-3, 3] eV at eta = 1e-6: the
    network is trained on the device on plain Sancho-Rubio results of every
    10th energy (the points a run solves first); the other energies compare
    plain Sancho-Rubio, Newton from the network's g0 and Newton from the
    bare layer's (z - H00)^-1, both with the fallback.  Every returned g00
    equals plain Sancho-Rubio's; the network's start is closer to the
    solution than the bare one.  Whether that pays (fallback rate, Newton
    steps, wall time) is what the test reports."""
    lead = _lead(qt, 64, batch=32, epochs=200)
    E = hm.energy_grid(-3.0, 3.0, 256)
    train = np.zeros(len(E), bool)
    train[::10] = True
    nk = len(KZ)
    loss = lead.train(np.repeat(E[train], nk), np.tile(np.arange(nk), train.sum()))
    print(f"training loss (mean relative squared error): first epoch {loss[0]:.3f}, last {loss[-1]:.4f}")
    assert loss[-1] < 0.2 * loss[0]
    ev = np.where(~train)[0][::2]
    Ee, ke = np.repeat(E[ev], nk), np.tile(np.arange(nk), len(ev))
    rows = {}
    for mode in ("sancho", "newton", "bare"):
        r = lead.surface(mode, Ee, ke, compare=True)
        conv = r["fallback"] == 0
        if mode != "sancho":
            assert (r["sancho_err"][conv] <= 1e-8).all()
            assert (r["sancho_err"][~conv] == 0).all()  # fallback results are Sancho's own
        rows[mode] = r
        rep = r["report"]
        print(f"{mode:7s}: lead work {rep['total_ms'] / rep['points']:.3f} ms/point, fallback rate "
              f"{rep['fallbacks'] / rep['points']:.3f}, Newton steps {rep['newton_steps'] / rep['points']:.2f}, "
              f"doublings {rep['smith_doublings'] / rep['points']:.2f}, warm-start error median "
              f"{np.median(r['warm_err']) if mode != 'sancho' else 0:.2f}")
    assert np.median(rows["newton"]["warm_err"]) < 0.7 * np.median(rows["bare"]["warm_err"])
    lead.close()


def test_training_leaves_out_unconverged_decimations(qt):
    """With the decimation capped below what in-band points need at eta =
    1e-6 (out-of-band points converge in 7-11 steps), training on a mix of
    converged and unconverged points must equal,
    bit for bit, training on the converged ones alone (the minibatch order
    depends only on the seed and the number of points kept)."""
    nmax = 24
    E = np.linspace(-90.0, 60.0, 96)  # the bands span about -56..24 eV: 56 points lie outside
    k = np.arange(96) % len(KZ)
    lead = _lead(qt, 64, batch=32, epochs=3, sancho_nmax=nmax)
    blocks = _lead_blocks(64)
    H00, H10 = [b[0] for b in blocks], [b[1] for b in blocks]
    _, _, it, cv = qt.sancho([H00[i] for i in k], [H10[i] for i in k], E, eta=1e-6, nmax=nmax, want_gbb=False)
    print(f"{cv.sum()} of {len(E)} points converge within {nmax} steps")
    assert 32 <= cv.sum() < len(E)
    p0 = lead.params()
    loss_mix = lead.train(E, k)
    p_mix = lead.params()
    lead2 = _lead(qt, 64, batch=32, epochs=3, sancho_nmax=nmax)
    lead2.set_params(p0)
    loss_conv = lead2.train(E[cv], k[cv])
    print("loss, mixed set:", loss_mix, "converged only:", loss_conv)
    assert np.isfinite(loss_mix).all()
    assert np.array_equal(loss_mix, loss_conv)
    assert np.array_equal(p_mix, lead2.params())
    lead.close()
    lead2.close()
