"""Particle weight rezoning control, the host side of
R/src/base/Rezoning.F90 (ModuleParticleRezoning).

``Rezoning.control(timer)`` is particle_split_merge_control
(Rezoning.F90:17-50): every step the electrons (species 0) are split and
merged; every ``period`` steps each bundle's base weight is first adjusted by
weighting_update_base_particle_number (Rezoning.F90:95-119) and the ions are
split and merged.  The per-particle work runs on the device
(``HotPath.split_merge``); this class keeps what the reference keeps in each
ParticleBundle (Weight, alpha, grid_r = Nr - 1) and the module's limits.

With ``fused_electrons=True`` the electrons' every-step split/merge runs
inside the fused step (``HotPath.set_step_rezone``: after the deposit and
before the collisions, OneStep's order) instead of as a separate pass;
``control(timer)`` then only updates the electrons' Weight every ``period``
steps (taking effect from the next step, as in the reference) and splits and
merges the ions as before, after the step's collisions rather than before
them.
"""
import dataclasses

import numpy as np


@dataclasses.dataclass
class RezoningLimits:
    """Module variables of ModuleParticleRezoning (Rezoning.F90:8-13), read
    from plasma.* in the reference's input (R/src/base/init.F90:196-209)."""
    is_open_split_merge: bool = True
    paritlc_upper_limit: float = 1.01
    paritlc_lower_limit: float = 0.99
    min_mp_to_split_k: float = 0.1
    merge_limit: float = 0.34


class Rezoning:
    def __init__(self, hp, nz, nr, weights, ppg_ref, alpha=0.01, period=1, limits=None, fused_electrons=False):
        self.hp = hp
        self.nz, self.nr = int(nz), int(nr)
        self.weight = [float(w) for w in weights]          # PB(k)%Weight
        self.ppg_ref = [int(p) for p in ppg_ref]           # ppg_dif(k)
        self.alpha = float(alpha)                          # PB%alpha (plasma.alpha)
        self.grid_r = float(self.nr - 1)                   # PB%grid_r, ParticleBundle.f90:347
        self.period = int(period)                          # CF%Period
        self.limits = limits or RezoningLimits()
        if len(self.weight) != hp.ns or len(self.ppg_ref) != hp.ns:
            raise ValueError("one weight and one ppg_ref per species")
        if self.period < 1:
            raise ValueError("period must be >= 1")
        self.fused_electrons = bool(fused_electrons) and self.limits.is_open_split_merge
        if self.fused_electrons:
            self._fold_electrons()

    def _fold_electrons(self):
        self.hp.set_step_rezone(0, self.weight[0], self.alpha, self.grid_r, self.limits.merge_limit)

    def reference_weight(self, s, r):
        """CalRW, R/src/pic/particle/ParticleBundle.f90:393-402."""
        return self.weight[s] * ((1.0 - self.alpha) * np.asarray(r) / self.grid_r + self.alpha)

    def split_merge(self, s):
        """split_particle_weight then merge_particle_random on bundle s."""
        return self.hp.split_merge(s, self.weight[s], self.alpha, self.grid_r, self.limits.merge_limit)

    def update_base_weight(self, s, npar=None):
        """weighting_update_base_particle_number, Rezoning.F90:95-119 (the
        MPI_ALLREDUCE over ranks is the one GPU's own count here)."""
        lim = self.limits
        npar = self.hp.count(s) if npar is None else int(npar)
        cells = float((self.nz - 1) * (self.nr - 1) * self.ppg_ref[s])
        npar_max = int(cells * lim.paritlc_upper_limit)
        npar_min = int(cells * lim.paritlc_lower_limit)
        limit = float(int(cells * lim.min_mp_to_split_k))
        if npar > npar_max:
            self.weight[s] = self.weight[s] * lim.paritlc_upper_limit
        elif npar < npar_min and npar > limit:
            self.weight[s] = self.weight[s] * lim.paritlc_lower_limit
        return self.weight[s]

    def control(self, timer):
        """particle_split_merge_control, Rezoning.F90:17-50.  Returns the
        device counts (copies, kept, deleted) per species touched."""
        out = {}
        if not self.limits.is_open_split_merge:
            return out
        if not self.fused_electrons:
            out[0] = self.split_merge(0)
        if timer % self.period == 0:
            self.update_base_weight(0)
            if self.fused_electrons:
                self._fold_electrons()
            for k in range(1, self.hp.ns):
                self.update_base_weight(k)
                out[k] = self.split_merge(k)
        return out
