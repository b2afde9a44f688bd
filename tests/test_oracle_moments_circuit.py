"""CPU checks of two restatements the differentiable window builds on.

ora_moments (WeightingParticleField, R/src/diag/diagParticleField.F90:196-318):
* sum over nodes of RhoOne * V_node = sum of WQ (all-vacuum grid)
* sum of MPOne = particle count
* uniform velocity: JzOne = RhoOne * Charge * Vz * VFactor node by node
* uniform speed: TOne = 0.5 m (|v| VFactor)^2 / JtoeV wherever RhoOne > 0
* zero field, Az = Ar = 0: the implicit JzOne equals the explicit one

The matching circuit (R/src/pic/circuit/IMN.F90:60-103, trapezoidal steps):
* agrees with a tight-tolerance integration of the same ODE (scipy)
* the periodic start satisfies the step recurrence exactly
* it approaches the continuous phasor steady state to O((omega dt)^2)
* torch gradcheck of the electrode voltages in (amplitude, phase, Cm1, Cm2)
"""
import math

import numpy as np
import pytest
import torch

from ipm2d import circuit as ck
from ipm2d import geometry as geo

from helpers import random_particles, QE, ME, AMU


def _vacuum(nz=17, nr=9, dx=1e-3):
    return geo.build_geometry(nz, nr, dx, np.zeros((nz - 1, nr - 1), dtype=np.int32), nmetal=0, ndiel=0)


def test_moments_conservation_and_counts(oracle, rng):
    g = _vacuum()
    P = random_particles(g.nz, g.nr, 3000, rng)
    P[7] = rng.uniform(0.5, 2.0, P.shape[1]) * 1e6
    ez = np.zeros((g.nz + 1) * g.nr)
    er = np.zeros(g.nz * (g.nr + 1))
    m = oracle.moments(g.nz, g.nr, g.dx, 5e-11, -QE, ME, ez, er, P, g.inv_volume, implicit_j=False)
    assert np.sum(m["rho"] * g.volume) == pytest.approx(P[7].sum(), rel=1e-12)
    assert m["mp"].sum() == pytest.approx(P.shape[1], rel=1e-12)


def test_moments_uniform_velocity_and_speed(oracle, rng):
    g = _vacuum()
    dt = 5e-11
    vf = g.dx / dt
    P = random_particles(g.nz, g.nr, 2000, rng)
    P[2], P[3], P[4] = 0.03, -0.02, 0.01
    P[7] = rng.uniform(0.5, 2.0, P.shape[1])
    ez = np.zeros((g.nz + 1) * g.nr)
    er = np.zeros(g.nz * (g.nr + 1))
    q, m_ = 1.0 * QE, 69.0 * AMU
    m = oracle.moments(g.nz, g.nr, g.dx, dt, q, m_, ez, er, P, g.inv_volume, implicit_j=False)
    on = m["rho"] > 0
    np.testing.assert_allclose(m["jz"][on], m["rho"][on] * q * 0.03 * vf, rtol=1e-12)
    np.testing.assert_allclose(m["jr"][on], m["rho"][on] * q * -0.02 * vf, rtol=1e-12)
    e_ev = 0.5 * m_ * (math.sqrt(0.03 ** 2 + 0.02 ** 2 + 0.01 ** 2) * vf) ** 2 / 1.6022e-19
    np.testing.assert_allclose(m["energy"][on], e_ev, rtol=1e-12)


def test_moments_implicit_current_in_zero_field(oracle, rng):
    g = _vacuum()
    P = random_particles(g.nz, g.nr, 1000, rng)
    P[5] = 0.0
    P[6] = 0.0
    ez = np.zeros((g.nz + 1) * g.nr)
    er = np.zeros(g.nz * (g.nr + 1))
    a = oracle.moments(g.nz, g.nr, g.dx, 5e-11, QE, 19 * AMU, ez, er, P, g.inv_volume, implicit_j=False)
    b = oracle.moments(g.nz, g.nr, g.dx, 5e-11, QE, 19 * AMU, ez, er, P, g.inv_volume, implicit_j=True)
    np.testing.assert_array_equal(a["jz"], b["jz"])
    np.testing.assert_array_equal(a["rho"], b["rho"])


# ------------------------------------------------------------- circuit
DT = 5e-11


def _imn():
    return ck.IMN(eqc=1.39e-12)


def test_circuit_matches_tight_integration():
    from scipy.integrate import solve_ivp
    p = _imn()
    th = p.theta()
    n = 600
    V = ck.voltages(p, th, DT, n).numpy()
    A, bV, _, _ = ck.system(p, th[2], th[3])
    A, bV = A.numpy(), bV.numpy()
    st = ck.Stepper(p, th, DT)
    y0 = st.periodic_state().numpy()
    w = 2 * math.pi * p.frequency

    def f(t, y):
        return A @ y + bV * p.amplitude * math.sin(w * t + p.phase_deg * math.pi / 180)

    t = np.arange(n) * DT
    sol = solve_ivp(f, (0.0, t[-1]), y0, t_eval=t, method="Radau", rtol=1e-11, atol=1e-22)
    v_ref = sol.y[5] / p.eqc
    # the 13.56 MHz response is resolved to (omega dt)^2 ~ 2e-5
    np.testing.assert_allclose(V, v_ref, rtol=0, atol=2e-3 * np.abs(v_ref).max())


def test_periodic_start_solves_the_recurrence():
    p = _imn()
    th = p.theta()
    a = ck.Stepper(p, th, DT, t0=0.0)
    b = ck.Stepper(p, th, DT, t0=DT)
    y1 = a.advance(a.periodic_state(), 0, 0.0)
    np.testing.assert_allclose(y1.numpy(), b.periodic_state().numpy(), rtol=1e-9,
                               atol=1e-12 * float(y1.abs().max()))


def test_periodic_start_near_continuous_phasor():
    p = _imn()
    th = p.theta()
    A, bV, _, _ = ck.system(p, th[2], th[3])
    w = 2 * math.pi * p.frequency
    yc = np.linalg.solve(1j * w * np.eye(6) - A.numpy().astype(complex), bV.numpy() * p.amplitude)
    y_cont = yc.imag                                    # phase 0 at t0 = 0
    y_disc = ck.Stepper(p, th, DT).periodic_state().numpy()
    assert abs(y_disc[5] - y_cont[5]) <= 1e-3 * np.abs(yc[5])


def test_circuit_gradcheck():
    p = ck.IMN(eqc=1.39e-12)
    scale = torch.tensor([1.0, 1.0, 1e-12, 1e-12], dtype=torch.float64)

    def f(u):
        return ck.voltages(p, u * scale, DT, 40)

    u = (p.theta() / scale).requires_grad_(True)
    # eps in scaled units (1e-4 V, 1e-4 degree, 1e-4 pF): below that the
    # central difference of the capacitance columns loses digits to rounding
    assert torch.autograd.gradcheck(f, (u,), eps=1e-4, atol=1e-6, rtol=1e-5)
