"""Recombination control, the host side of R/src/mc/Recombination.F90.

``Recombination`` holds what initRecombination2D stores (reaction list and
rates, Recombination.F90:43-140, read from gas.recb.* in the reference's
input, R/src/base/init.F90:819-878) and ``control(timer)`` is the call in
OneStep (R/src/onestep.F90:180-182): every recombination_time_step_number
steps, recombination over that many time steps.  The per-cell work runs on
the device (``HotPath.recombine``).
"""
import numpy as np


class Recombination:
    def __init__(self, hp, reactions, rates, time_step_number, dt, enabled=True):
        rl = np.asarray(reactions, dtype=np.int64).reshape(-1, 3)
        if np.any(rl[:, :2] < 0) or np.any(rl[:, :2] >= hp.ns):
            raise ValueError("The specified particle index is invalid in recombination reactions.")
        if np.any((rl[:, 2] != 0) & (rl[:, 2] != 1)):
            raise ValueError("recombination type must be 0 (constant) or 1 (temperature)")
        if len(rates) != rl.shape[0]:
            raise ValueError("one rate per reaction")
        if int(time_step_number) < 1:
            raise ValueError("recb_time_step_num must be >= 1")
        self.hp = hp
        self.reactions = [tuple(int(x) for x in r) for r in rl]
        self.rates = [float(r) for r in rates]
        self.time_step_number = int(time_step_number)
        self.dt = float(dt)
        self.enabled = bool(enabled)

    def control(self, timer):
        """Deleted macro-particles per species, or None when not due."""
        if not self.enabled or timer % self.time_step_number != 0:
            return None
        return self.hp.recombine(self.reactions, self.rates, self.time_step_number * self.dt)
