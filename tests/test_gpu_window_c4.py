"""The differentiable window at its full size: 129x257 nodes, ~4M particles
(e, CF3+, F-), 200 steps, checkpoint every 20 steps, the ion flux to the
powered electrode as the loss, gradients with respect to RF amplitude, phase,
Cm1 and Cm2 of the matching circuit (Rs = 50 Ohm, Cm1 = 500 pF, Cm2 = 200 pF).

* the window's recording forward reproduces the plain method (product entry
  points ipm_move_absorb + ipm_moments, circuit stepped on the host with the
  live electrode current) to 1e-12 relative;
* the adjoint gradient matches central finite differences with
  h = 1e-4 * |theta_k| to 1e-5 relative for every parameter.
"""
import numpy as np
import pytest

import ipm2d
from ipm2d import circuit as ck
from ipm2d import workload as W
from ipm2d.matching import FluxGradient, electrode_capacitance

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4():
    cfg = W.make_config("c4")
    geom = W.make_geometry(cfg)
    rng = np.random.default_rng(cfg.seed)
    parts = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    hp, _, _, _ = ipm2d.from_config(cfg, geom=geom, particles=parts, capacity_factor=1.2)
    imn = ck.IMN(Cm1=500e-12, Cm2=200e-12, Rs=50.0, amplitude=200.0, phase_deg=30.0,
                 eqc=electrode_capacitance(cfg))
    fg = FluxGradient(cfg, hp, geom, parts, imn, T=cfg.steps, interval=20)
    yield fg
    fg.close()
    hp.close()


def test_window_reproduces_plain_method(c4):
    assert c4.window.info()["segments"] == 10
    a, b = c4.base_loss, c4.record_loss
    assert np.abs(a).max() > 0
    np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_adjoint_matches_central_differences(c4):
    L, g = c4.gradient()
    fd = c4.central_fd(rel=1e-4)
    rel = (g - fd).abs() / fd.abs()
    assert float(rel.max()) < 1e-5, f"grad {g.numpy()} fd {fd.numpy()} rel {rel.numpy()}"
