"""Differentiable collisionless window: a torch.autograd.Function over the
ipm_window_* entry points (forward replay and hand-written HIP adjoint).

A window advances a snapshot of the resident particles through T steps of
the push (MoveESIm, R/src/pic/particle/ParticleBundle.f90:671-748) and the
absorption rule (R/src/pic/particleboundary/ParticleMaterial.F90:22-41) in
the field E0 + c[n] E1 and returns one loss per step,

    l[n] = sum_i alpha[s_i] WQ_i Vz_i M(Z_i, R_i),

M being the P2CES-weighted interpolation (R/src/pic/field/FieldOne.f90:151-166)
of a node field mu.  ``ion_flux_loss`` builds mu and alpha so that
sum_n l[n] is the time-integrated ion flux density toward the powered
electrode (z = 0), measured with the deposit on the nodes next to it:
l[n] = dt * sum_ions (-JzOne / Charge) summed over the masked nodes.

``WindowLoss.apply(c, window)`` maps the coefficients c (T,) to the losses
(T,); its backward runs the adjoint kernels.  Absorption events are frozen
at those of the recording forward (``Window.record``).
"""
import ctypes as C

import numpy as np
import torch

from . import _native as N
from .geometry import fortran

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


class Window:
    def __init__(self, hp, nsteps, interval, e0, e1, mu, alpha):
        """hp: HotPath with particles resident (sort them first); e0, e1:
        (Ez, Er) pairs in the Fortran layouts; mu: (nz, nr) node field;
        alpha: per-species loss factors."""
        self.hp, self.T, self.K = hp, int(nsteps), int(interval)
        ez0, er0 = (np.ascontiguousarray(x, dtype=np.float64).ravel() for x in e0)
        ez1, er1 = (np.ascontiguousarray(x, dtype=np.float64).ravel() for x in e1)
        assert ez0.size == ez1.size == (hp.nz + 1) * hp.nr and er0.size == er1.size == hp.nz * (hp.nr + 1)
        mu = np.ascontiguousarray(fortran(np.asarray(mu, dtype=np.float64)))
        assert mu.size == hp.nz * hp.nr
        al = np.zeros(max(hp.ns, 1))
        al[:len(alpha)] = alpha
        N.call("ipm_window_create", hp._h, self.T, self.K, _d(ez0), _d(er0), _d(ez1), _d(er1), _d(mu), _d(al))
        self.recorded = False

    def record(self, coef):
        """Forward with absorption by the reference's rule; stores each
        particle's deletion step for later replays.  Returns the losses."""
        out = self._forward(coef, 1)
        self.recorded = True
        return out

    def forward(self, coef):
        """Replay with the recorded deletion steps."""
        return self._forward(coef, 0)

    def _forward(self, coef, record):
        c = np.ascontiguousarray(coef, dtype=np.float64)
        assert c.size == self.T
        loss = np.zeros(self.T)
        N.call("ipm_window_forward", self.hp._h, _d(c), record, _d(loss))
        return loss

    def backward(self, coef, gloss):
        c = np.ascontiguousarray(coef, dtype=np.float64)
        g = np.ascontiguousarray(gloss, dtype=np.float64)
        assert c.size == self.T and g.size == self.T
        dc = np.zeros(self.T)
        N.call("ipm_window_backward", self.hp._h, _d(c), _d(g), _d(dc))
        return dc

    def info(self):
        n = C.c_int64()
        ns = C.c_int()
        f = C.c_float()
        b = C.c_float()
        N.call("ipm_window_info", self.hp._h, C.byref(n), C.byref(ns), C.byref(f), C.byref(b))
        return {"particles": n.value, "segments": ns.value, "forward_ms": f.value, "adjoint_ms": b.value}

    def checkpoint(self, k):
        """State at checkpoint k (7, N) as Z R Vz Vr Vt Az Ar, and deletion steps."""
        n = self.info()["particles"]
        st = np.zeros(7 * max(n, 1))
        dead = np.zeros(max(n, 1), dtype=np.int32)
        N.call("ipm_window_download", self.hp._h, int(k), _d(st), dead.ctypes.data_as(_ip))
        return st[:7 * n].reshape(7, n), dead[:n]

    def close(self):
        if self.hp is not None and self.hp._h is not None:
            N.call("ipm_window_destroy", self.hp._h)
        self.hp = None


class WindowLoss(torch.autograd.Function):
    """c (T,) float64 -> per-step losses (T,); gradient by the HIP adjoint."""

    @staticmethod
    def forward(ctx, coef, window):
        c = coef.detach().cpu().numpy().astype(np.float64)
        ctx.window = window
        ctx.save_for_backward(coef)
        return torch.from_numpy(window.forward(c)).to(coef.dtype)

    @staticmethod
    def backward(ctx, gl):
        (coef,) = ctx.saved_tensors
        c = coef.detach().cpu().numpy().astype(np.float64)
        g = gl.detach().cpu().numpy().astype(np.float64)
        dc = ctx.window.backward(c, g)
        return torch.from_numpy(dc).to(coef.dtype), None


def window_loss(coef, window):
    return WindowLoss.apply(coef, window)


def ion_flux_loss(cfg, geom, layers=4):
    """(mu, alpha) of the ion flux toward the powered electrode at z = 0:
    mask = 1 on the `layers` node planes next to the electrode's metal cells
    (node indices 2 .. 1 + layers, 1-based), mu = mask * point_inv_volume;
    alpha[s] = dt * w_s * Charge_s * VFactor with w_s = -1/Charge_s for ions
    and 0 for electrons, i.e. alpha = -dt * VFactor for every ion species.
    Then sum_n l[n] = dt * sum_n sum_ions sum_masked (-JzOne / Charge)."""
    nz, nr = cfg.nz, cfg.nr
    mask = np.zeros((nz, nr))
    mask[1:1 + layers, :] = 1.0
    mu = mask * np.asarray(geom.inv_volume, dtype=np.float64)
    vf = cfg.dx / cfg.dt
    alpha = np.array([0.0 if sp.name == "e" else -cfg.dt * vf for sp in cfg.species])
    return mu, alpha, mask


def split_fields(cfg):
    """E0 (the fixed plasma part) and E1 (per volt on the powered electrode)
    of the synthetic potential, so that the field at voltage V is E0 + V E1."""
    import dataclasses
    from . import workload as W
    from .geometry import set_e_from_phi
    phi0 = W.potential(dataclasses.replace(cfg, drive=[(f, 0.0) for f, _ in cfg.drive]), 0.0)
    z = np.arange(cfg.nz)[:, None] * cfg.dx
    phi1 = (1.0 - z / cfg.length) + 0.0 * np.arange(cfg.nr)[None, :]
    return set_e_from_phi(phi0, cfg.dx), set_e_from_phi(phi1, cfg.dx)
