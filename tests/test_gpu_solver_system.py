"""The field solver's linear system assembled on the device
(ipm_solver_system: SetFieldSolverCoe / SetFieldSolverCoeWithoutChi,
SetFieldSolverSource, SetSolverBoundary; R/src/pic/field/FieldSolver.f90:136-301)
against the oracle, bit for bit, on the Rho and Chi of real steps: c1 and
the full c2 grid, a permittivity that differs in every cell, random surface
charge and boundary values."""
import numpy as np
import pytest

import ipm2d
from ipm2d import workload as W

pytestmark = pytest.mark.gpu


def _run(cfg_name, oracle, steps, ppc=None):
    cfg = W.make_config(cfg_name)
    if ppc is not None:
        import dataclasses
        cfg = dataclasses.replace(cfg, species=[dataclasses.replace(x, ppc=ppc) for x in cfg.species])
    geom = W.make_geometry(cfg)
    rng = np.random.default_rng(31)
    parts = [W.load_species(cfg, geom, s, np.random.default_rng(5 + s)) for s in range(cfg.ns)]
    hp, _, _, _ = ipm2d.from_config(cfg, geom=geom, particles=parts, capacity_factor=1.5)
    ez, er = W.fields(cfg, 0.0)
    hp.set_fields(ez, er)
    hp.set_rng(7, 0)
    eps = rng.uniform(1.0, 12.0, (cfg.nz - 1, cfg.nr - 1))
    pt = np.array(geom.point_type, dtype=np.int32)
    hp.set_solver_geometry(eps, pt)
    for _ in range(steps):
        hp.one_step()
    rho, chi = hp.field_source_sum()
    assert np.abs(chi).max() > 0 and np.abs(rho).max() > 0
    sigma = rng.normal(size=(cfg.nz, cfg.nr)) * 1e-8
    bound = rng.normal(size=(cfg.nz, cfg.nr)) * 100
    mask = ((pt > 0) | (rng.uniform(size=pt.shape) < 0.05)).astype(np.int32)
    print(f"{cfg_name}: {cfg.nz}x{cfg.nr}, metal nodes {int((pt > 0).sum())}", flush=True)

    for with_chi in (True, False):
        coe, src = hp.solver_system(with_chi=with_chi, sigma=sigma, bound=bound, bound_mask=mask)
        ref = oracle.solver_coe(chi, eps, pt, with_chi=with_chi)
        for k in "ABCDE":
            assert np.array_equal(coe[k], ref[k]), (cfg_name, with_chi, k)
        assert np.array_equal(src, oracle.solver_source(cfg.dx, rho, geom.inv_volume, sigma=sigma, bound=bound,
                                                        bound_mask=mask))
    coe, src = hp.solver_system()                     # no surface charge, no boundary merge
    assert np.array_equal(src, oracle.solver_source(cfg.dx, rho, geom.inv_volume))
    # Chi of a plasma raises the coefficients above the vacuum form off the metal
    wn = oracle.solver_coe(chi, eps, pt, with_chi=False)
    inner = (pt[1:-1, :-1] == 0)
    assert (coe["C"][1:-1, :-1][inner] >= wn["C"][1:-1, :-1][inner]).all()
    hp.close()


def test_solver_system_c1(oracle):
    _run("c1", oracle, steps=5, ppc=20.0)


def test_solver_system_c2_grid(oracle):
    # the full 257x513 grid of c2 with a lighter load (the system depends on
    # the grid, not on the particle count)
    _run("c2", oracle, steps=2, ppc=2.0)
