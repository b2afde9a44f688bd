"""Gradient of the ion flux to the powered electrode with respect to the
drive and matching network: RF amplitude, phase, Cm1 and Cm2.

Chain, for a window of T steps that starts from resident particles:

  theta --(matching circuit, circuit.py)--> V[n] --(window, window.py)--> l[n]

The field of step n is E0 + V[n] E1 (E0: the fixed plasma part of the
potential; E1: per volt on the powered electrode).  The plasma current into
the electrode, Iconv[n] (charge collected in step n over dt, as the
reference feeds its circuit, R/src/pic/circuit/Circuits.F90:166-170), and
every absorption event are taken from a base run by the plain method and
then held fixed: collisions and absorption are frozen events of the window,
and the space-charge field is frozen in E0.  The electrode capacitance is
the vacuum-plate value of E1's geometry with q0 = 0, and nothing is
differentiated through a Poisson solve: the gradients returned here are
those of this reduced model, not of the self-consistent discharge.

``plain_coupled_run`` is that base run: per step the field from the present
circuit voltage, PB%MoveES and MaterialAborptionParticle for every species
(ipm_move_absorb), the electrode tally, the loss from the deposited JzOne
(ipm_moments), and one circuit step, exactly in the order of OneStep
(R/src/onestep.F90:141-172).
"""
import numpy as np
import torch

from . import circuit as ck
from .window import Window, ion_flux_loss, split_fields, window_loss


def plain_coupled_run(hp, cfg, particles, e0, e1, mask, weights, imn, theta, T, t0=0.0):
    """Returns (loss[n], voltage[n], iconv[n]).  Uploads `particles` first."""
    for s in range(cfg.ns):
        hp.upload(s, particles[s])
    hp.take_tallies()
    st = ck.Stepper(imn, torch.as_tensor(theta, dtype=torch.float64), cfg.dt, t0)
    y = st.periodic_state()
    ez0, er0 = (np.asarray(x, dtype=np.float64).ravel() for x in e0)
    ez1, er1 = (np.asarray(x, dtype=np.float64).ravel() for x in e1)
    loss, volts, iconv = [], [], []
    for n in range(T):
        v = float(st.voltage(y))
        hp.set_fields(ez0 + v * ez1, er0 + v * er1)
        hp.move_absorb(-1)
        mc, _ = hp.take_tallies()
        i_n = float(mc[0]) / cfg.dt if len(mc) else 0.0
        ln = 0.0
        for s in range(cfg.ns):
            if weights[s] != 0.0:
                ln += weights[s] * cfg.dt * float(np.sum(mask * hp.moments(s, implicit_j=False)["jz"]))
        y = st.advance(y, n, i_n)
        loss.append(ln)
        volts.append(v)
        iconv.append(i_n)
    return np.array(loss), np.array(volts), np.array(iconv)


def electrode_capacitance(cfg):
    """Vacuum capacitance of the plate pair that E1 describes: eps0 pi R^2 / L."""
    return 8.8542e-12 * np.pi * cfg.radius ** 2 / cfg.length


class FluxGradient:
    """Sets up the base run and the window; ``loss(theta)`` is a torch scalar
    whose backward runs the circuit in torch and the window adjoint on the GPU."""

    def __init__(self, cfg, hp, geom, particles, imn, T, interval=20, layers=4):
        self.cfg, self.hp, self.imn, self.T = cfg, hp, imn, T
        self.e0, self.e1 = split_fields(cfg)
        mu, alpha, self.mask = ion_flux_loss(cfg, geom, layers)
        from . import workload as W
        self.weights = np.array([0.0 if sp.name == "e" else -1.0 / W.species_params(cfg, s)[0]
                                 for s, sp in enumerate(cfg.species)])
        self.theta0 = imn.theta()
        self.base_loss, self.base_volts, self.iconv = plain_coupled_run(
            hp, cfg, particles, self.e0, self.e1, self.mask, self.weights, imn, self.theta0, T)
        for s in range(cfg.ns):
            hp.upload(s, particles[s])
        hp.sort(-1)
        self.window = Window(hp, T, interval, self.e0, self.e1, mu, alpha)
        self.record_loss = self.window.record(self.base_volts)

    def voltages(self, theta):
        return ck.voltages(self.imn, theta, self.cfg.dt, self.T, iconv=self.iconv)

    def loss(self, theta):
        return window_loss(self.voltages(theta), self.window).sum()

    def gradient(self, theta=None):
        th = (self.theta0 if theta is None else theta).clone().requires_grad_(True)
        L = self.loss(th)
        L.backward()
        return float(L.detach()), th.grad.clone()

    def central_fd(self, theta=None, rel=1e-4):
        """Central differences with h_k = rel * |theta_k| per parameter."""
        th = (self.theta0 if theta is None else theta).clone()
        out = torch.zeros(4, dtype=torch.float64)
        with torch.no_grad():
            for k in range(4):
                h = rel * abs(float(th[k]))
                e = torch.zeros(4, dtype=torch.float64)
                e[k] = h
                out[k] = (self.loss(th + e) - self.loss(th - e)) / (2 * h)
        return out

    def close(self):
        self.window.close()
