"""GPU parity of the collision path with two background gases, and of twin
products, against the oracle event for event.

The reference loops over gases outside the species loop (R/src/mc/MCCPublic.F90:52-115):
each gas draws its own selection against its own CollisionRatio and sees
the velocity the earlier gases left.  With more than one gas the device
repeats the selection per gas in the collision pass, writes back null
events' VelRes rounding itself, and runs the gases after a real reaction
in its second part; this file is what exercises those branches.
"""
import numpy as np
import pytest

from helpers import AMU, QE, assert_close, random_particles, sort_set
from ipm2d import HotPath
from ipm2d import workload as W
from ipm2d import xsec

pytestmark = pytest.mark.gpu


def _second_gas_tables(cfg):
    """A lighter, hotter second gas at a third of the density: its own
    electron set (different thresholds and products), a charge-exchange set
    for CF3+, a reactive set for F-, and no table (model 0) for the rest."""
    gas = dict(mass=39.948 * AMU, temperature=450.0, density=xsec.gas_density(cfg.pressure_mtorr, 450.0, 1.0 / 3.0),
               natom=1.0, beta_max=2.0)
    names = [sp.name for sp in cfg.species]
    idx = {nm: k for k, nm in enumerate(names)}
    tabs = {}
    e = xsec.synthetic_electron_reactions(ion_th=12.0, ion_peak=8e-20, ion_product=idx["CF2+"],
                                          att_product=idx["CF3-"], excitation=(9.0, 3e-20))
    tabs[(0, 1)] = xsec.build_bundle(xsec.SigmaFile(1, e), W.species_params(cfg, 0)[1], 1.0, gas, cfg.dt)
    k = idx["CF3+"]
    ions = W._ion_model2(3e-19, 2e-19, 102)
    for r in ions:
        r.resultant = k
    tabs[(k, 1)] = xsec.build_bundle(xsec.SigmaFile(2, ions), W.species_params(cfg, k)[1], 4.0, gas, cfg.dt)
    k = idx["F-"]
    tabs[(k, 1)] = xsec.build_bundle(xsec.SigmaFile(3, [xsec.Reaction(0, 101, k, 0.0), xsec.Reaction(0, 113, 0, 2.5),
                                                        xsec.Reaction(0, 111, k, 4.0)]),
                                     W.species_params(cfg, k)[1], 1.0, gas, cfg.dt)
    return tabs


def _setup(pressure=1500.0):
    cfg = W.make_config("c3", nz=33, nr=17, length=32e-3, pressure_mtorr=pressure)
    tables = dict(W.build_tables(cfg))
    tables.update(_second_gas_tables(cfg))
    species = [W.species_params(cfg, s) for s in range(cfg.ns)]
    return cfg, tables, species


def _hp(cfg, geom, tables, species, ng, cap=400000):
    hp = HotPath(cfg.nz, cfg.nr, cfg.dx, cfg.dt, cfg.ns, ng)
    for s, (q, m) in enumerate(species):
        hp.set_species(s, q, m, cap)
    hp.set_geometry(geom)
    for s in range(cfg.ns):
        for g in range(ng):
            hp.set_mcc(s, g, tables.get((s, g)))
    return hp


def _energetic(cfg, s, n, rng, ev, inside=False):
    P = random_particles(cfg.nz, cfg.nr, n, rng)
    if inside:
        P[0] = rng.uniform(4.0, cfg.nz - 5.0, n)
        P[1] = rng.uniform(0.7, cfg.nr - 4.0, n)
    q, m = W.species_params(cfg, s)
    v = np.sqrt(2.0 * ev * QE / m) / (cfg.dx / cfg.dt)
    P[2:5] = rng.standard_normal((3, n)) * v / np.sqrt(3.0)
    return P


def _compare(hp, ref, P, ns, what):
    for s in range(ns):
        got = hp.download(s)
        n0 = P[s].shape[1]
        assert got.shape == ref[s].shape, f"{what}: species {s} {got.shape} vs {ref[s].shape}"
        g0, r0 = got[:, :n0], ref[s][:, :n0]
        err = np.linalg.norm(g0[2:5] - r0[2:5], axis=0)
        assert np.all(err <= 1e-10 * np.linalg.norm(r0[2:5], axis=0) + 1e-300), f"{what}: species {s} velocities"
        assert_close(g0[[0, 1, 7]], r0[[0, 1, 7]], rtol=1e-12, atol=1e-300, what=f"{what}: species {s} Z R WQ")
        assert_close(g0[5:7], r0[5:7], rtol=1e-12, atol=1e-300, what=f"{what}: species {s} Az/Ar")
        pg, pr = sort_set(got[:, n0:]), sort_set(ref[s][:, n0:])
        assert pg.shape == pr.shape
        if pg.shape[1]:
            assert_close(pg[[0, 1, 5, 6, 7]], pr[[0, 1, 5, 6, 7]], rtol=1e-12, atol=1e-300,
                         what=f"{what}: species {s} products")
            e = np.linalg.norm(pg[2:5] - pr[2:5], axis=0)
            assert np.all(e <= 1e-10 * np.linalg.norm(pr[2:5], axis=0) + 1e-300)


def test_two_gas_mcc_matches_oracle(oracle, rng):
    cfg, tables, species = _setup()
    geom = W.make_geometry(cfg)
    hp = _hp(cfg, geom, tables, species, 2)
    energies = [30.0, 2.0, 5.0, 5.0, 8.0, 10.0]
    P = [_energetic(cfg, s, 60000 if s == 0 else 30000, rng, energies[s]) for s in range(cfg.ns)]
    for s in range(cfg.ns):
        hp.upload(s, P[s])
    seed, step = 0x5EED0002, 11
    hp.set_rng(seed, step)
    hp.set_event_counting(True)
    hp.mcc()
    counts = hp.take_mcc_counts()
    setup = oracle.MccSetup(cfg.dx, cfg.dt, [q for q, _ in species], [m for _, m in species], tables, ng=2)
    ref, cref = oracle.mcc(setup, seed, step, P, counts=True)
    np.testing.assert_array_equal(counts, cref)
    # both gases collide, with real reactions on each, and particles that
    # collide with both in one pass exist
    assert counts[0, 0, 1:].sum() > 100 and counts[0, 1, 1:].sum() > 100
    assert counts[1, 1, 1:].sum() > 10 and counts[4, 1, 1:].sum() > 0
    assert counts[2, 1].sum() == 0                     # model 0 on gas 2
    _compare(hp, ref, P, cfg.ns, "two-gas MCC")
    hp.close()


def test_two_gas_fused_step_matches_oracle(oracle, rng):
    """The fused step (push, absorb, deposit, per-gas selection, collisions)
    with two gases against the oracle's push -> deposit -> MCC."""
    cfg, tables, species = _setup(pressure=800.0)
    geom = W.make_geometry(cfg)
    nz, nr = cfg.nz, cfg.nr
    P = [_energetic(cfg, s, 15000, rng, 20.0 if s == 0 else 1.0, inside=True) for s in range(cfg.ns)]
    rngf = np.random.default_rng(3)
    ez = rngf.standard_normal((nz + 1) * nr) * 50.0
    er = rngf.standard_normal(nz * (nr + 1)) * 50.0
    hp = _hp(cfg, geom, tables, species, 2)
    for s in range(cfg.ns):
        hp.upload(s, P[s])
    hp.set_fields(ez, er)
    hp.set_rng(21, 4)
    hp.set_event_counting(True)
    hp.one_step()
    rho, _ = hp.field_source_sum()
    counts = hp.take_mcc_counts()
    pushed = []
    rho_ref = np.zeros((nz, nr))
    for s in range(cfg.ns):
        q, m = species[s]
        X = oracle.push(nz, nr, cfg.dx, cfg.dt, q, m, ez, er, P[s])
        X2, _, _ = oracle.absorb(geom, q, X)
        assert X2.shape == X.shape, "test setup lost particles"
        r_s, _ = oracle.p2ces(nz, nr, cfg.dt, q, m, X, geom.inv_volume)
        rho_ref = rho_ref + r_s
        pushed.append(X)
    assert_close(rho, rho_ref, rtol=1e-12, atol=1e-12 * np.abs(rho_ref).max(), what="Rho")
    setup = oracle.MccSetup(cfg.dx, cfg.dt, [q for q, _ in species], [m for _, m in species], tables, ng=2)
    ref, cref = oracle.mcc(setup, 21, 4, pushed, counts=True)
    np.testing.assert_array_equal(counts, cref)
    assert counts[0, 1, 1:].sum() > 10
    _compare(hp, ref, pushed, cfg.ns, "two-gas step")
    hp.close()


def test_two_gas_graph_equals_eager(rng):
    cfg, tables, species = _setup(pressure=800.0)
    geom = W.make_geometry(cfg)
    P = [_energetic(cfg, s, 12000, rng, 20.0 if s == 0 else 1.0) for s in range(cfg.ns)]
    ez, er = W.fields(cfg, 1e-8)
    a = _hp(cfg, geom, tables, species, 2)
    b = _hp(cfg, geom, tables, species, 2)
    for hp in (a, b):
        for s in range(cfg.ns):
            hp.upload(s, P[s])
        hp.set_fields(ez * 30, er * 30)
        hp.set_rng(5, 0)
    for _ in range(4):
        a.one_step()
        b.one_step_graph()
    b.synchronize()
    for s in range(cfg.ns):
        ga, gb = a.download(s), b.download(s)
        assert ga.shape == gb.shape
        np.testing.assert_array_equal(sort_set(ga), sort_set(gb))
    a.close()
    b.close()


def _vt_bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def test_twin_products_match_oracle(oracle, rng):
    """Dissociation into two fragments of one species and 113 giving back the
    parent's species: the device tags the twin like the oracle (Vt's low bit)."""
    cfg, _, species = _setup()
    geom = W.make_geometry(cfg)
    idx = {sp.name: k for k, sp in enumerate(cfg.species)}
    gas = W.gas_dict(cfg)
    fm = idx["F-"]
    e = xsec.log_grid()
    el = [xsec.Reaction(0, 11, 0, 0.0, e, np.full(e.size, 5e-20)),
          xsec.Reaction(0, 15, fm * 10 + fm, 5.0, e, np.where(e > 5.0, 2e-19, 0.0))]
    tables = {(0, 0): xsec.build_bundle(xsec.SigmaFile(1, el), species[0][1], 1.0, gas, cfg.dt),
              (fm, 0): xsec.build_bundle(xsec.SigmaFile(3, [xsec.Reaction(0, 101, fm, 0.0),
                                                            xsec.Reaction(0, 113, fm, 0.0)]),
                                         species[fm][1], 1.0, gas, cfg.dt)}
    hp = _hp(cfg, geom, tables, species, 1)
    P = [_energetic(cfg, s, 40000 if s in (0, fm) else 10, rng, 30.0 if s == 0 else 3.0) for s in range(cfg.ns)]
    for s in range(cfg.ns):
        hp.upload(s, P[s])
    hp.set_rng(77, 2)
    hp.set_event_counting(True)
    hp.mcc()
    counts = hp.take_mcc_counts()
    setup = oracle.MccSetup(cfg.dx, cfg.dt, [q for q, _ in species], [m for _, m in species], tables)
    ref, cref = oracle.mcc(setup, 77, 2, P, counts=True)
    np.testing.assert_array_equal(counts, cref)
    assert counts[0, 0, 2] > 100 and counts[fm, 0, 2] > 10
    _compare(hp, ref, P, cfg.ns, "twins")
    got = hp.download(fm)
    prods = got[:, P[fm].shape[1]:]
    # no two products share a bit pattern
    assert np.unique(prods.T, axis=0).shape[1] == 8 and np.unique(prods.T, axis=0).shape[0] == prods.shape[1]
    hp.close()


def test_hard_sphere_tables_match_oracle(oracle, rng):
    """Ions with the hard-sphere fallback (ProbilityHardShpere,
    R/src/mc/MCCInit.F90:172-270): the elastic/charge-exchange pair of a
    one-ion gas on gas 1 and the elastic-only table of a several-ion gas on
    gas 2, event for event against the oracle."""
    cfg, tables, species = _setup(pressure=3000.0)
    geom = W.make_geometry(cfg)
    g1 = dict(mass=39.948 * AMU, temperature=300.0, density=xsec.gas_density(cfg.pressure_mtorr, 300.0), radius=1.9e-10,
              ns=1, index_start=0, beta_max=0.0)
    g2 = dict(g1, mass=20.18 * AMU, radius=1.5e-10, ns=3, density=g1["density"] / 2)
    tabs = {(0, 0): tables[(0, 0)]}
    for s in range(1, cfg.ns):
        tabs[(s, 0)] = xsec.hard_sphere_bundle(s, species[s][1], 2e-10, g1, cfg.dt)
        tabs[(s, 1)] = xsec.hard_sphere_bundle(s, species[s][1], 2e-10, g2, cfg.dt)
    hp = _hp(cfg, geom, tabs, species, 2)
    P = [_energetic(cfg, s, 40000, rng, 30.0 if s == 0 else 10.0) for s in range(cfg.ns)]
    for s in range(cfg.ns):
        hp.upload(s, P[s])
    hp.set_rng(404, 9)
    hp.set_event_counting(True)
    hp.mcc()
    counts = hp.take_mcc_counts()
    setup = oracle.MccSetup(cfg.dx, cfg.dt, [q for q, _ in species], [m for _, m in species], tabs, ng=2)
    ref, cref = oracle.mcc(setup, 404, 9, P, counts=True)
    np.testing.assert_array_equal(counts, cref)
    assert counts[1, 0, 1] > 50 and counts[1, 0, 2] > 50 and counts[1, 1, 1] > 50
    _compare(hp, ref, P, cfg.ns, "hard sphere")
    hp.close()
