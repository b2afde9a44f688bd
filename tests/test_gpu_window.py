"""GPU checks of the moment diagnostics and the differentiable window.

Moments (ipm_moments) against the oracle's WeightingParticleField, 1e-12.

Window forward (ipm_window_forward, record mode) against the plain method
restated on the CPU (oracle.window_plain: field E0 + c E1 formed on the host,
MoveESIm, MaterialAborptionParticle, loss from the deposited JzOne): per-step
losses 1e-12 relative, surviving particles 1e-12 as sets.

Adjoint (ipm_window_backward, hand-written HIP) against
* central finite differences of the GPU replay along random directions,
* central finite differences of the CPU plain method (no frozen events;
  the test checks that no absorption event flips between the two sides),
* torch.autograd.gradcheck of the window behind WindowLoss, chained with
  the matching circuit,
* itself under other checkpoint intervals and a permuted particle order.
"""
import numpy as np
import pytest
import torch

from helpers import AMU, ME, QE, assert_close, random_particles, small_geometry, sort_set
from ipm2d import HotPath
from ipm2d import circuit as ck
from ipm2d.geometry import set_e_from_phi
from ipm2d.window import Window, window_loss

pytestmark = pytest.mark.gpu

DX = 1e-3
SPECIES = [(-QE, ME), (QE, 69.0059096 * AMU), (-QE, 18.9984030 * AMU)]


def fields(geom, amp0=25.0):
    nz, nr = geom.nz, geom.nr
    L = (nz - 1) * geom.dx
    z = np.arange(nz)[:, None] * geom.dx
    r = np.arange(nr)[None, :] * geom.dx
    R = (nr - 1) * geom.dx
    phi0 = amp0 * np.sin(np.pi * z / L) ** 2 * np.cos(np.pi * r / (2 * R)) ** 2
    phi1 = (1.0 - z / L) + 0.0 * r
    return set_e_from_phi(phi0, geom.dx), set_e_from_phi(phi1, geom.dx)


def near_electrode(geom, n, rng, vscale, zmax=9.0):
    P = random_particles(geom.nz, geom.nr, n, rng, vscale=vscale)
    P[0] = rng.uniform(1.02, zmax, n)
    return P


def setup(rng, dt, loss_species, n=(1500, 1500, 1000), vscales=(0.05, 0.002, 0.002), layers=4, shuffle=False):
    geom = small_geometry(dielectric=False)
    parts = [near_electrode(geom, n[s], rng, vscales[s]) for s in range(3)]
    if shuffle:
        parts = [P[:, rng.permutation(P.shape[1])] for P in parts]
    hp = HotPath(geom.nz, geom.nr, geom.dx, dt, 3, 1)
    for s, (q, m) in enumerate(SPECIES):
        hp.set_species(s, q, m, 10000)
    hp.set_geometry(geom)
    for s in range(3):
        hp.upload(s, parts[s])
    mask = np.zeros((geom.nz, geom.nr))
    mask[1:1 + layers, :] = 1.0
    mu = mask * geom.inv_volume
    vf = geom.dx / dt
    weights = np.array([(-1.0 / SPECIES[s][0]) if s in loss_species else 0.0 for s in range(3)])
    alpha = np.array([dt * weights[s] * SPECIES[s][0] * vf for s in range(3)])
    return geom, hp, parts, mask, mu, weights, alpha


def coefs(T, amp=150.0, period=40.0, phase=0.3):
    n = np.arange(T)
    return amp * np.cos(2 * np.pi * n / period + phase) - 20.0


# ------------------------------------------------------------------ moments
@pytest.mark.parametrize("implicit_j", [False, True])
def test_moments_match_oracle(oracle, rng, implicit_j):
    geom = small_geometry()
    nz, nr = geom.nz, geom.nr
    dt = 5e-11
    (ez, er), _ = fields(geom, amp0=400.0)
    hp = HotPath(nz, nr, geom.dx, dt, 2, 1)
    for s in range(2):
        hp.set_species(s, *SPECIES[s], 50000)
    hp.set_geometry(geom)
    hp.set_fields(ez, er)
    parts = [random_particles(nz, nr, 20000, rng, vscale=0.1), random_particles(nz, nr, 5000, rng, vscale=0.01)]
    for s in range(2):
        hp.upload(s, parts[s])
    for s in range(2):
        got = hp.moments(s, implicit_j=implicit_j)
        ref = oracle.moments(nz, nr, geom.dx, dt, *SPECIES[s], ez, er, parts[s], geom.inv_volume,
                             implicit_j=implicit_j)
        for k in ("rho", "energy", "mp", "jz", "jr"):
            a, b = got[k], ref[k]
            assert_close(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max(), what=f"{k} species {s}")


# ------------------------------------------------------------ window forward
@pytest.mark.parametrize("dt,loss_species", [(2e-9, (1, 2)), (5e-11, (0,))])
def test_window_forward_matches_plain_oracle(oracle, rng, dt, loss_species):
    geom, hp, parts, mask, mu, weights, alpha = setup(rng, dt, loss_species)
    e0, e1 = fields(geom)
    T = 30
    c = coefs(T)
    w = Window(hp, T, 7, e0, e1, mu, alpha)
    lw = w.record(c)
    species = [(q, m, parts[s]) for s, (q, m) in enumerate(SPECIES)]
    lo, _, Pfinal = oracle.window_plain(geom, species, e0, e1, c, mask, weights, dt)
    assert np.abs(lo).max() > 0
    assert_close(lw, lo, rtol=1e-12, atol=1e-12 * np.abs(lo).max(), what="per-step loss")
    st, dead = w.checkpoint(w.info()["segments"])
    off = 0
    for s in range(3):
        n = parts[s].shape[1]
        alive = dead[off:off + n] == T
        got = np.vstack([st[:, off:off + n][:, alive], parts[s][7:8, alive]])
        assert got.shape[1] == Pfinal[s].shape[1], f"survivors of species {s}"
        assert_close(sort_set(got), sort_set(Pfinal[s]), rtol=1e-12, atol=1e-300, what=f"final state species {s}")
        off += n
    # replay with the recorded events gives the same losses
    assert_close(w.forward(c), lw, rtol=1e-13, atol=1e-13 * np.abs(lw).max(), what="replay")
    w.close()


# -------------------------------------------------------------------- adjoint
def _directional_fd(f, c, v, h):
    return (f(c + h * v) - f(c - h * v)) / (2 * h)


@pytest.mark.parametrize("dt,loss_species", [(2e-9, (1, 2)), (5e-11, (0,)), (5e-11, (0, 1, 2))])
def test_adjoint_matches_finite_differences(rng, dt, loss_species):
    geom, hp, parts, mask, mu, weights, alpha = setup(rng, dt, loss_species)
    e0, e1 = fields(geom)
    T = 30
    c = coefs(T)
    w = Window(hp, T, 7, e0, e1, mu, alpha)
    w.record(c)
    g = rng.uniform(0.5, 1.5, T)
    dc = w.backward(c, g)
    assert np.abs(dc).max() > 0
    for trial in range(3):
        v = rng.standard_normal(T)
        fd = _directional_fd(lambda x: float(g @ w.forward(x)), c, v, 1e-6 * np.abs(c).max())
        ad = float(dc @ v)
        assert ad == pytest.approx(fd, rel=1e-6, abs=1e-9 * np.abs(dc).max() * np.abs(v).sum())
    w.close()


def test_adjoint_matches_plain_method_fd(oracle, rng):
    """Reference gradient from the plain method: central differences of the
    CPU restatement of the reference's step (events not frozen)."""
    dt = 5e-11
    geom, hp, parts, mask, mu, weights, alpha = setup(rng, dt, (0, 1, 2), n=(600, 600, 400))
    e0, e1 = fields(geom)
    T = 20
    c = coefs(T)
    w = Window(hp, T, 5, e0, e1, mu, alpha)
    w.record(c)
    dc = w.backward(c, np.ones(T))
    species = [(q, m, parts[s]) for s, (q, m) in enumerate(SPECIES)]
    v = rng.standard_normal(T)
    h = 1e-6 * np.abs(c).max()
    lp, _, Pp = oracle.window_plain(geom, species, e0, e1, c + h * v, mask, weights, dt)
    lm, _, Pm = oracle.window_plain(geom, species, e0, e1, c - h * v, mask, weights, dt)
    assert [P.shape[1] for P in Pp] == [P.shape[1] for P in Pm], "an absorption event flipped"
    fd = (lp.sum() - lm.sum()) / (2 * h)
    assert float(dc @ v) == pytest.approx(fd, rel=1e-6)
    w.close()


def test_adjoint_invariant_to_interval_and_order(rng):
    dt = 5e-11
    T = 25
    c = coefs(T)
    g = np.linspace(0.5, 1.5, T)
    grads = []
    for K, shuffle in ((1, False), (4, False), (25, False), (64, False), (6, True)):
        r = np.random.default_rng(7)
        geom, hp, parts, mask, mu, weights, alpha = setup(r, dt, (0, 1, 2), shuffle=shuffle)
        e0, e1 = fields(geom)
        w = Window(hp, T, K, e0, e1, mu, alpha)
        w.record(c)
        grads.append(w.backward(c, g))
        w.close()
    for k, gk in enumerate(grads[1:], 1):
        assert_close(gk, grads[0], rtol=1e-11, atol=1e-11 * np.abs(grads[0]).max(), what=f"variant {k}")


def test_gradcheck_through_circuit(rng):
    dt = 5e-11
    geom, hp, parts, mask, mu, weights, alpha = setup(rng, dt, (0, 1, 2), n=(400, 400, 300))
    e0, e1 = fields(geom)
    T = 16
    p = ck.IMN(eqc=1.39e-12)
    scale = torch.tensor([1.0, 1.0, 1e-12, 1e-12], dtype=torch.float64)
    th0 = p.theta()
    w = Window(hp, T, 4, e0, e1, mu, alpha)
    w.record(ck.voltages(p, th0, dt, T).detach().numpy())

    def f(u):
        return window_loss(ck.voltages(p, u * scale, dt, T), w)

    # losses are O(1e12): gradcheck sees them divided by that scale
    lscale = float(np.abs(w.forward(ck.voltages(p, th0, dt, T).detach().numpy())).max())

    def fn(u):
        return f(u) / lscale

    u = (th0 / scale).requires_grad_(True)
    # the adjoint sums with atomics: repeated backwards agree to rounding only
    assert torch.autograd.gradcheck(fn, (u,), eps=1e-4, atol=1e-8, rtol=1e-5, nondet_tol=1e-12)
    cc = torch.tensor(coefs(T), dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: window_loss(x, w) / lscale, (cc,), eps=1e-3, atol=1e-9, rtol=1e-5,
                                    nondet_tol=1e-12)
    w.close()
