"""Statistical parity of the full step (push, absorption, deposit, MCC) over
many steps: the GPU against the oracle restating the reference's routines,
on the c1 geometry (65x129 nodes, e / CF3+ / F-, 12 macro-particles per cell
per species, CF4 50 mTorr, synthetic cross sections), 100 steps, seed 0.

Both sides draw from the same counter-based streams, so they start out event
for event identical; any later divergence (a last-ulp difference of a
transcendental flipping one comparison) makes them two statistically
equivalent runs.  The bounds are therefore those of independent runs:

* per-process event counts (species, reaction): |a - b| <= 5 sqrt(a + b + 1)
* time-averaged macro-particle profile along z (RhoOne's MPOne summed over
  r) per species: |a - b| <= 5 sqrt(2 mean + 1) per node
* electron and ion mean energy (EEDF first moment) at the end: within
  5 sigma / sqrt(N) of each other, and the second moment likewise
* the fraction of electrons selected per step (null events included)
  equals the table's CollisionRatio within 5 binomial sigma (closed form)
"""
import dataclasses

import numpy as np
import pytest

import ipm2d
from ipm2d import workload as W

pytestmark = pytest.mark.gpu

STEPS = 100


def _energies(P, mass, vfactor):
    v2 = (P[2] ** 2 + P[3] ** 2 + P[4] ** 2) * vfactor ** 2
    return 0.5 * mass * v2 / 1.6022e-19


def test_mcc_statistics_match_oracle(oracle):
    cfg = W.make_config("c1")
    cfg.species = [dataclasses.replace(sp, ppc=12.0) for sp in cfg.species]
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    rng = np.random.default_rng(cfg.seed)
    parts = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    hp, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.5)
    hp.set_event_counting(True)
    hp.take_mcc_counts()
    ez, er = W.fields(cfg, 0.0)
    sp = [W.species_params(cfg, s) for s in range(cfg.ns)]
    setup = oracle.MccSetup(cfg.dx, cfg.dt, [q for q, _ in sp], [m for _, m in sp], tables)
    st = oracle.StepState(geom, setup, parts, ez, er, extra=0.5, fast=True)
    counts_o = np.zeros(cfg.ns * 128, dtype=np.int64)

    prof_g = np.zeros((cfg.ns, cfg.nz))
    prof_o = np.zeros((cfg.ns, cfg.nz))
    electrons_seen = 0
    for n in range(STEPS):
        electrons_seen += hp.count(0)
        hp.one_step()
        st.step(cfg.seed, n, counts_o)
        for s in range(cfg.ns):
            q, m = sp[s]
            prof_g[s] += hp.moments(s, implicit_j=False)["mp"].sum(axis=1)
            prof_o[s] += oracle.moments(cfg.nz, cfg.nr, cfg.dx, cfg.dt, q, m, ez, er, st.species(s),
                                        geom.inv_volume, implicit_j=False)["mp"].sum(axis=1)
    counts_g = hp.take_mcc_counts().reshape(-1)

    # per-process counts
    assert counts_o.sum() > 1000
    for k in np.nonzero(counts_g + counts_o)[0]:
        a, b = int(counts_g[k]), int(counts_o[k])
        assert abs(a - b) <= 5 * np.sqrt(a + b + 1), f"process {divmod(k, 128)}: {a} vs {b}"

    # selection fraction of electrons vs CollisionRatio (null events included)
    ratio = tables[(0, 0)].collision_ratio
    sel = counts_g[:128].sum()
    sigma = np.sqrt(electrons_seen * ratio * (1 - ratio))
    assert abs(sel - electrons_seen * ratio) <= 5 * sigma, (sel, electrons_seen * ratio, sigma)

    # time-averaged profiles along z
    prof_g /= STEPS
    prof_o /= STEPS
    for s in range(cfg.ns):
        bound = 5 * np.sqrt(2 * np.maximum(prof_g[s], prof_o[s]) + 1)
        bad = np.abs(prof_g[s] - prof_o[s]) > bound
        assert not bad.any(), f"species {s} profile off at nodes {np.nonzero(bad)[0][:5]}"

    # EEDF moments at the end
    vf = cfg.dx / cfg.dt
    for s in range(cfg.ns):
        eg = _energies(hp.download(s), sp[s][1], vf)
        eo = _energies(st.species(s), sp[s][1], vf)
        for k, (a, b) in enumerate(((eg, eo), (eg ** 2, eo ** 2))):
            sig = np.sqrt(a.var() / a.size + b.var() / b.size)
            assert abs(a.mean() - b.mean()) <= 5 * sig + 1e-12 * abs(b.mean()), \
                f"species {s} moment {k + 1}: {a.mean()} vs {b.mean()}"
    hp.close()
