"""Matching network of the powered electrode (the reference's IMN circuit,
R/src/pic/circuit/IMN.F90:60-103) as a fixed-step integrator that torch
can differentiate.

State (IMN.F90:79-93, indices 1..6 there, 0..5 here):
    y0' = (Vs - (y0 - y1)/Cm1) / Rs
    y1' = y2
    y2' = ((y0 - y1)/Cm1 - y1/Cm2 - y2*Rm - (y5 - q0)/eqc) / Lm
    y3' = y2 - y4                                  (stray branch; else 0)
    y4' = ((y5 - q0)/eqc - (y1 - y3)/C_stray - y4*R_stray) / L_stray
    y5' = y2 - y4 + Iconv
with the source Vs(t) = A sin(2 pi f t + phase * pi/180) (IMN.F90:75-78)
and the electrode voltage V = (y5 - q0)/eqc (R/src/pic/circuit/Circuits.F90:170-172).
Iconv is the charge the electrode collected in the step divided by dt.

The reference advances the ODE with CVODE; here every particle step is one
trapezoidal (Crank-Nicolson) step, A-stable for the stiff stray branch and
differentiable through plain linear algebra.  The window can start from the
discrete periodic steady state of the source alone (no plasma current),
found in closed form from the complex amplitude of the recurrence.
"""
import math
from dataclasses import dataclass

import torch


@dataclass
class IMN:
    Rs: float = 50.0
    Cm1: float = 500e-12
    Cm2: float = 200e-12
    Lm: float = 4.3e-6          # R/config.json circuits_list[0]
    Rm: float = 0.5
    stray: bool = True
    C_stray: float = 24e-12
    L_stray: float = 10e-9
    R_stray: float = 0.5
    frequency: float = 13.56e6
    amplitude: float = 200.0
    phase_deg: float = 0.0
    eqc: float = 1.0e-12        # electrode capacitance (calculateCapacitance)
    q0: float = 0.0             # induced charge of the frozen space charge

    def theta(self):
        """The differentiable parameters (amplitude, phase in degrees, Cm1, Cm2)."""
        return torch.tensor([self.amplitude, self.phase_deg, self.Cm1, self.Cm2], dtype=torch.float64)


def system(p: IMN, cm1, cm2):
    """y' = A y + bV Vs + bI Iconv + bq q0 (torch, float64; cm1/cm2 may carry grad)."""
    one = torch.ones((), dtype=torch.float64)
    z = torch.zeros((), dtype=torch.float64)
    e, Lm, Rs = p.eqc, p.Lm, p.Rs
    rows = [[z] * 6 for _ in range(6)]
    rows[0][0] = -1.0 / (cm1 * Rs)
    rows[0][1] = 1.0 / (cm1 * Rs)
    rows[1][2] = one
    rows[2][0] = 1.0 / (cm1 * Lm)
    rows[2][1] = (-1.0 / cm1 - 1.0 / cm2) / Lm
    rows[2][2] = -p.Rm / Lm * one
    rows[2][5] = -1.0 / (e * Lm) * one
    if p.stray:
        Ls, Cs = p.L_stray, p.C_stray
        rows[3][2] = one
        rows[3][4] = -one
        rows[4][5] = 1.0 / (e * Ls) * one
        rows[4][1] = -1.0 / (Cs * Ls) * one
        rows[4][3] = 1.0 / (Cs * Ls) * one
        rows[4][4] = -p.R_stray / Ls * one
    rows[5][2] = one
    rows[5][4] = -one
    A = torch.stack([torch.stack(r) for r in rows])
    bV = torch.zeros(6, dtype=torch.float64)
    bV[0] = 1.0 / Rs
    bI = torch.zeros(6, dtype=torch.float64)
    bI[5] = 1.0
    bq = torch.zeros(6, dtype=torch.float64)
    bq[2] = 1.0 / (e * Lm)
    if p.stray:
        bq[4] = -1.0 / (e * p.L_stray)
    return A, bV, bI, bq


class Stepper:
    """Trapezoidal steps of the circuit at step dt for parameters theta =
    (amplitude, phase_deg, Cm1, Cm2).  Differentiable when theta requires grad."""

    def __init__(self, p: IMN, theta, dt, t0=0.0):
        self.p, self.dt, self.t0 = p, dt, t0
        amp, ph, c1, c2 = theta[0], theta[1], theta[2], theta[3]
        self.amp, self.ph = amp, ph
        A, bV, bI, bq = system(p, c1, c2)
        I6 = torch.eye(6, dtype=torch.float64)
        Mm = I6 - 0.5 * dt * A
        Mp = I6 + 0.5 * dt * A
        self.N = torch.linalg.solve(Mm, Mp)
        self.G = torch.linalg.solve(Mm, (0.5 * dt) * bV)
        self.GI = torch.linalg.solve(Mm, dt * bI)
        self.Gq = torch.linalg.solve(Mm, (dt * p.q0) * bq)
        self.omega = 2.0 * math.pi * p.frequency

    def source(self, n):
        t = self.t0 + n * self.dt
        return self.amp * torch.sin(self.omega * t + self.ph * (math.pi / 180.0))

    def periodic_state(self):
        """y0 of the discrete periodic response to the source alone: the
        recurrence y[n+1] = N y[n] + G (Vs[n] + Vs[n+1]) with Vs[n] =
        Im(a z^n), z = exp(i omega dt), is solved by y[n] = Im(Y z^n),
        Y = (z I - N)^-1 G a (1 + z).  Needs q0 = 0."""
        if self.p.q0 != 0.0:
            raise ValueError("the periodic start assumes q0 = 0")
        zeta = complex(math.cos(self.omega * self.dt), math.sin(self.omega * self.dt))
        ang = self.omega * self.t0 + self.ph * (math.pi / 180.0)
        a = torch.polar(self.amp * torch.ones((), dtype=torch.float64), ang * torch.ones((), dtype=torch.float64))
        Nc = self.N.to(torch.complex128)
        lhs = zeta * torch.eye(6, dtype=torch.complex128) - Nc
        rhs = self.G.to(torch.complex128) * (a * (1.0 + zeta))
        return torch.linalg.solve(lhs, rhs).imag

    def voltage(self, y):
        return (y[5] - self.p.q0) / self.p.eqc

    def advance(self, y, n, iconv):
        """y at step n -> y at step n + 1, with the plasma current iconv over the step."""
        return self.N @ y + self.G * (self.source(n) + self.source(n + 1)) + self.GI * iconv + self.Gq


def voltages(p: IMN, theta, dt, nsteps, iconv=None, t0=0.0, periodic=True):
    """Electrode voltage at steps 0..nsteps-1 (the voltage the push of step n
    sees) with the plasma current iconv[n] collected during step n."""
    st = Stepper(p, theta, dt, t0)
    y = st.periodic_state() if periodic else torch.zeros(6, dtype=torch.float64)
    out = []
    for n in range(nsteps):
        out.append(st.voltage(y))
        ic = 0.0 if iconv is None else float(iconv[n])
        y = st.advance(y, n, ic)
    return torch.stack(out)
