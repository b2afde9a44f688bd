"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: push and deposit 1e-12 relative in FP64 (deposit allows for atomic
summation order); absorption exact as a particle set plus 1e-12 on the
tallies; MCC with the same counter-based random stream reproduces the
oracle's events, so velocities agree to 1e-10 of |v| (device and host libm
transcendentals may differ in the last ulp) and products agree as sets.
"""
import dataclasses

import numpy as np
import pytest

from helpers import (AMU, ME, QE, assert_close, random_fields, random_particles, small_geometry, sort_set)
from ipm2d import HotPath, CapacityError, IPMError
from ipm2d import workload as W

pytestmark = pytest.mark.gpu

DX, DT = 1e-3, 5e-11
SPECIES = [(-QE, ME), (QE, 69.0059096 * AMU)]


def make_hp(geom, nspecies=2, cap=200000, ng=1, species=SPECIES):
    hp = HotPath(geom.nz, geom.nr, geom.dx, DT, nspecies, ng)
    for s in range(nspecies):
        q, m = species[s]
        hp.set_species(s, q, m, cap)
    hp.set_geometry(geom)
    return hp


def test_push_matches_oracle(oracle, rng):
    geom = small_geometry()
    nz, nr = geom.nz, geom.nr
    ez, er = random_fields(nz, nr, rng)
    hp = make_hp(geom)
    hp.set_fields(ez, er)
    parts = [random_particles(nz, nr, 20000, rng), random_particles(nz, nr, 7000, rng, vscale=0.01)]
    for s in range(2):
        hp.upload(s, parts[s])
    hp.move_es(-1)
    for s in range(2):
        got = hp.download(s)
        q, m = SPECIES[s]
        ref = oracle.push(nz, nr, DX, DT, q, m, ez, er, parts[s])
        scale = np.abs(ref).max(axis=1, keepdims=True)
        assert_close(got, ref, rtol=1e-12, atol=0.0, what=f"push species {s}")
        assert np.all(np.abs(got - ref) <= 1e-12 * np.maximum(np.abs(ref), 1e-300) + 1e-15 * scale)


def test_push_single_species_and_aos_layout(oracle, rng):
    geom = small_geometry()
    nz, nr = geom.nz, geom.nr
    ez, er = random_fields(nz, nr, rng)
    hp = make_hp(geom)
    hp.set_fields(ez, er)
    P0 = random_particles(nz, nr, 3000, rng)
    P1 = random_particles(nz, nr, 500, rng)
    hp.upload_aos(0, P0.T)
    hp.upload(1, P1)
    hp.move_es(0)                       # species 1 untouched
    assert_close(hp.download(1), P1, rtol=0, what="untouched species")
    ref = oracle.push(nz, nr, DX, DT, SPECIES[0][0], SPECIES[0][1], ez, er, P0)
    assert_close(hp.download_aos(0).T, ref, rtol=1e-12, what="push AoS")


def _absorb_particles(geom, rng, n):
    nz, nr = geom.nz, geom.nr
    P = np.zeros((8, n))
    P[0] = rng.uniform(-1.0, nz, n)          # includes Z <= 0 and Z >= nz-1
    P[1] = rng.uniform(-0.5, nr - 0.5, n)    # includes R <= 0 (reflect) and R >= nr-1
    P[0, :5] = [0.0, nz - 1.0, 5.5, 7.25, 3.0]
    P[1, :5] = [3.0, 4.0, nr - 1.0, 0.0, -0.0]
    P[2:5] = rng.standard_normal((3, n))
    P[5:7] = rng.standard_normal((2, n))
    P[7] = rng.uniform(1.0, 3.0, n)
    return P


def test_absorption_matches_oracle(oracle, rng):
    geom = small_geometry()
    hp = make_hp(geom)
    P = [_absorb_particles(geom, rng, 30000), _absorb_particles(geom, rng, 9000)]
    for s in range(2):
        hp.upload(s, P[s])
    hp.material_absorption(-1)
    mc, sg = hp.take_tallies()
    mc_ref = np.zeros(geom.nmetal)
    sg_ref = np.zeros((geom.nz, geom.nr))
    for s in range(2):
        got = hp.download(s)
        ref, m1, s1 = oracle.absorb(geom, SPECIES[s][0], P[s])
        mc_ref += m1
        sg_ref += s1
        assert got.shape == ref.shape, f"survivors {got.shape} vs {ref.shape}"
        np.testing.assert_array_equal(sort_set(got), sort_set(ref))
    assert (np.abs(sg_ref) > 0).sum() > 10, "dielectric edges were not exercised"
    assert_close(mc, mc_ref, rtol=1e-12, what="metal charge")
    # sigma and metal charge add both species' charges: opposite signs cancel,
    # so the atomic summation order shows at 1e-12 of the magnitude, not of the net
    assert_close(sg, sg_ref, rtol=1e-12, atol=1e-12 * np.abs(sg_ref).max(), what="sigma")
    # tallies are cleared after being taken
    mc2, sg2 = hp.take_tallies()
    assert not mc2.any() and not sg2.any()


def test_absorption_vacuum_runs_after_push(oracle, rng):
    """The particle pass skips the material lookup for a cell inside its
    row's vacuum run (the longest run of cells that are neither metal nor
    dielectric, read for the four rows around each wave's first particle).
    Cell-sorted particles, so whole waves take that path, over a geometry
    whose rows hold two or three vacuum runs (a dielectric block and a metal
    block inside the domain, the dielectric ring, both electrodes): the
    survivors after push + absorption and every tally equal the oracle's."""
    nz, nr = 65, 33
    cm = np.zeros((nz - 1, nr - 1), dtype=np.int32)
    cm[:, nr - 4:] = -1
    cm[20:26, 6:12] = -1          # dielectric block: rows 6-11 hold two vacuum runs
    cm[40:43, 9:20] = 2           # metal block: rows 9-11 hold three
    cm[0, :] = 1
    cm[nz - 2, :] = 2
    from ipm2d import geometry as geo
    geom = geo.build_geometry(nz, nr, 1e-3, cm, nmetal=2, ndiel=1, permittivity={-1: 2.1})
    hp = make_hp(geom)
    ez, er = random_fields(nz, nr, rng, scale=3e3)
    hp.set_fields(ez, er)
    P = []
    for s, n in ((0, 60000), (1, 20000)):
        X = random_particles(nz, nr, n, rng, vscale=0.3)
        X[0] = rng.uniform(0.2, nz - 1.2, n)
        X[1] = rng.uniform(0.0, nr - 1.5, n)
        key = (np.ceil(X[0]) - 1) + (np.ceil(X[1]) - 1) * (nz - 1)
        P.append(np.ascontiguousarray(X[:, np.argsort(key, kind="stable")]))
    for s in range(2):
        hp.upload(s, P[s])
    hp.move_absorb(-1)
    mc, sg = hp.take_tallies()
    mc_ref = np.zeros(geom.nmetal)
    sg_ref = np.zeros((nz, nr))
    for s in range(2):
        q, m = SPECIES[s]
        X = oracle.push(nz, nr, DX, DT, q, m, ez, er, P[s])
        ref, m1, s1 = oracle.absorb(geom, q, X)
        mc_ref += m1
        sg_ref += s1
        got = hp.download(s)
        assert got.shape == ref.shape, f"species {s}: survivors {got.shape} vs {ref.shape}"
        assert_close(sort_set(got), sort_set(ref), rtol=1e-12, atol=1e-300, what=f"species {s}")
    assert (np.abs(sg_ref) > 0).sum() > 20, "dielectric edges were not exercised"
    assert (mc_ref != 0).all(), "both metals take charge"
    assert_close(mc, mc_ref, rtol=1e-12, what="metal charge")
    assert_close(sg, sg_ref, rtol=1e-12, atol=1e-12 * np.abs(sg_ref).max(), what="sigma")


def test_deposit_matches_oracle(oracle, rng):
    geom = small_geometry()
    nz, nr = geom.nz, geom.nr
    hp = make_hp(geom)
    P = [random_particles(nz, nr, 40000, rng), random_particles(nz, nr, 10000, rng)]
    P[0][:, :20000] = P[0][:, np.argsort(P[0][0, :20000] + 1000 * np.ceil(P[0][1, :20000]))]   # a sorted run
    for s in range(2):
        hp.upload(s, P[s])
    for s in range(2):
        rho, chi = hp.p2ces(s)
        q, m = SPECIES[s]
        rref, cref = oracle.p2ces(nz, nr, DT, q, m, P[s], geom.inv_volume)
        assert_close(rho, rref, rtol=1e-12, atol=1e-300, what=f"RhoOne {s}")
        assert_close(chi, cref, rtol=1e-12, atol=1e-300, what=f"ChiOne {s}")
    rho_all, chi_all = hp.p2ces_all()
    rho, chi = hp.field_source_sum()
    assert_close(rho, rho_all[0] + rho_all[1], rtol=1e-15, atol=1e-300, what="FieldSource Rho")
    assert chi.shape == (nz + 2, nr + 2)
    assert_close(chi[1:-1, 1:-1], chi_all[0] + chi_all[1], rtol=1e-15, atol=1e-300, what="FieldSource Chi")
    assert not chi[0].any() and not chi[-1].any() and not chi[:, 0].any() and not chi[:, -1].any()


def _six_species_setup(pressure=2000.0):
    cfg = W.make_config("c3", nz=33, nr=17, length=32e-3, pressure_mtorr=pressure)
    tables = W.build_tables(cfg)
    species = [W.species_params(cfg, s) for s in range(cfg.ns)]
    return cfg, tables, species


def _energetic(cfg, s, n, rng, ev):
    P = random_particles(cfg.nz, cfg.nr, n, rng)
    q, m = W.species_params(cfg, s)
    v = np.sqrt(2.0 * ev * QE / m) / (cfg.dx / cfg.dt)
    P[2:5] = rng.standard_normal((3, n)) * v / np.sqrt(3.0)
    return P


def _non_cumulative(tables):
    """The same tables with the first reaction's column raised above the
    others: Probility(1, i) > Probility(j, i), so the columns are no longer
    nondecreasing in j (the null test from the last value alone would be
    wrong; the device must walk every reaction)."""
    out = {}
    for k, b in tables.items():
        if b is not None and b.model != 0 and b.nreaction >= 2:
            pr = np.array(b.prob, copy=True)
            pr[:, 0] = np.minimum(1.0, pr[:, -1] * 1.1)
            b = dataclasses.replace(b, prob=pr)
        out[k] = b
    return out


@pytest.mark.parametrize("cumulative", [True, False])
def test_mcc_matches_oracle(oracle, rng, cumulative):
    cfg, tables, species = _six_species_setup()
    if not cumulative:
        tables = _non_cumulative(tables)
    geom = W.make_geometry(cfg)
    hp = HotPath(cfg.nz, cfg.nr, cfg.dx, cfg.dt, cfg.ns, 1)
    for s, (q, m) in enumerate(species):
        hp.set_species(s, q, m, 400000)
    hp.set_geometry(geom)
    for s in range(cfg.ns):
        hp.set_mcc(s, 0, tables.get((s, 0)))
    energies = [30.0, 2.0, 5.0, 5.0, 8.0, 10.0]
    P = [_energetic(cfg, s, 60000 if s == 0 else 30000, rng, energies[s]) for s in range(cfg.ns)]
    for s in range(cfg.ns):
        hp.upload(s, P[s])
    seed, step = 0x1234ABCD5678, 17
    hp.set_rng(seed, step)
    hp.set_event_counting(True)
    hp.mcc()
    counts = hp.take_mcc_counts()
    setup = oracle.MccSetup(cfg.dx, cfg.dt, [q for q, _ in species], [m for _, m in species], tables)
    ref, cref = oracle.mcc(setup, seed, step, P, counts=True)
    np.testing.assert_array_equal(counts, cref)
    assert counts[0, 0, 1:].sum() > 100
    for s in range(cfg.ns):
        got = hp.download(s)
        n0 = P[s].shape[1]
        assert got.shape == ref[s].shape, f"species {s}: {got.shape} vs {ref[s].shape}"
        # in-place particles keep their slots
        g0, r0 = got[:, :n0], ref[s][:, :n0]
        vg, vr = g0[2:5], r0[2:5]
        err = np.linalg.norm(vg - vr, axis=0)
        assert np.all(err <= 1e-10 * np.linalg.norm(vr, axis=0) + 1e-300), f"species {s} velocities"
        assert_close(g0[[0, 1, 7]], r0[[0, 1, 7]], rtol=0, what=f"species {s} positions/weights")
        assert_close(g0[5:7], r0[5:7], rtol=1e-12, atol=1e-300, what=f"species {s} Az/Ar")
        # products (appended) agree as sets
        pg, pr = sort_set(got[:, n0:]), sort_set(ref[s][:, n0:])
        assert pg.shape == pr.shape
        if pg.shape[1]:
            assert_close(pg[[0, 1, 7]], pr[[0, 1, 7]], rtol=0, what=f"species {s} products pos")
            e = np.linalg.norm(pg[2:5] - pr[2:5], axis=0)
            assert np.all(e <= 1e-10 * np.linalg.norm(pr[2:5], axis=0) + 1e-300)


def _config_hp(cfg, geom, tables, particles, cap_factor=2.0, extra=4096):
    hp = HotPath(cfg.nz, cfg.nr, cfg.dx, cfg.dt, cfg.ns, 1)
    for s in range(cfg.ns):
        q, m = W.species_params(cfg, s)
        hp.set_species(s, q, m, int(particles[s].shape[1] * cap_factor) + extra)
    hp.set_geometry(geom)
    for s in range(cfg.ns):
        hp.set_mcc(s, 0, tables.get((s, 0)))
        hp.upload(s, particles[s])
    return hp


def test_fused_step_equals_separate_calls(rng):
    cfg, tables, species = _six_species_setup(pressure=500.0)
    geom = W.make_geometry(cfg, dielectric_ring=2)
    P = [_energetic(cfg, s, 20000, rng, 20.0 if s == 0 else 3.0) for s in range(cfg.ns)]
    ez, er = W.fields(cfg, 1e-8)
    ez, er = ez * 50, er * 50
    a = _config_hp(cfg, geom, tables, P)
    b = _config_hp(cfg, geom, tables, P)
    for hp in (a, b):
        hp.set_fields(ez, er)
        hp.set_rng(99, 5)
    a.one_step()
    rho_a, chi_a = a.field_source_sum()
    b.move_absorb(-1)
    b.p2ces_all()
    rho_b, chi_b = b.field_source_sum()
    b.mcc()
    # net Rho/Chi are differences of species terms (cancellation): absolute
    # tolerance scaled to the field's magnitude; per-species sums meet 1e-12
    assert_close(rho_a, rho_b, rtol=1e-12, atol=1e-12 * np.abs(rho_b).max(), what="Rho fused vs separate")
    assert_close(chi_a, chi_b, rtol=1e-12, atol=1e-12 * np.abs(chi_b).max(), what="Chi fused vs separate")
    for s in range(cfg.ns):
        ga, gb = a.download(s), b.download(s)
        assert ga.shape == gb.shape
        assert_close(sort_set(ga), sort_set(gb), rtol=1e-14, atol=1e-300, what=f"species {s}")
    ma, sa = a.take_tallies()
    mb, sb = b.take_tallies()
    assert_close(ma, mb, rtol=1e-12, atol=1e-12 * np.abs(ma).max(), what="metal charge")
    assert_close(sa, sb, rtol=1e-12, atol=1e-12 * np.abs(sa).max(), what="sigma")


def test_full_step_matches_oracle_without_losses(oracle, rng):
    """No particle leaves, so slots stay put and the oracle's push -> absorb
    -> deposit -> MCC sequence sees the same slots as the fused step."""
    cfg, tables, species = _six_species_setup(pressure=1000.0)
    geom = W.make_geometry(cfg)
    nz, nr = cfg.nz, cfg.nr
    P = []
    for s in range(cfg.ns):
        X = _energetic(cfg, s, 15000, rng, 20.0 if s == 0 else 1.0)
        X[0] = rng.uniform(4.0, nz - 5.0, X.shape[1])
        X[1] = rng.uniform(0.7, nr - 4.0, X.shape[1])
        P.append(X)
    ez, er = random_fields(nz, nr, rng, scale=50.0)
    hp = _config_hp(cfg, geom, tables, P)
    hp.set_fields(ez, er)
    hp.set_rng(7, 3)
    hp.one_step()
    rho, chi = hp.field_source_sum()
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
    setup = oracle.MccSetup(cfg.dx, cfg.dt, [q for q, _ in species], [m for _, m in species], tables)
    ref, _ = oracle.mcc(setup, 7, 3, pushed)
    for s in range(cfg.ns):
        got = hp.download(s)
        assert got.shape == ref[s].shape
        n0 = P[s].shape[1]
        assert_close(got[[0, 1, 5, 6, 7], :n0], ref[s][[0, 1, 5, 6, 7], :n0], rtol=1e-12, atol=1e-300,
                     what=f"species {s} state")
        err = np.linalg.norm(got[2:5, :n0] - ref[s][2:5, :n0], axis=0)
        assert np.all(err <= 1e-10 * np.linalg.norm(ref[s][2:5, :n0], axis=0) + 1e-300)


def test_sort_preserves_particles_and_deposit(oracle, rng):
    cfg = W.make_config("tiny", nz=65, nr=33, length=64e-3)
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    P = [random_particles(cfg.nz, cfg.nr, 30000, rng) for _ in range(cfg.ns)]
    hp = _config_hp(cfg, geom, tables, P)
    ez, er = W.fields(cfg, 0.0)
    hp.set_fields(ez, er)
    hp.move_absorb(-1)                 # creates dead slots
    before = [hp.download(s) for s in range(cfg.ns)]
    r0, _ = hp.p2ces_all()
    hp.sort(-1)
    for s in range(cfg.ns):
        after = hp.download(s)
        np.testing.assert_array_equal(sort_set(after), sort_set(before[s]))
        cell = (np.ceil(after[0]) - 1) + (np.maximum(np.ceil(after[1]), 1) - 1) * (cfg.nz - 1)
        assert np.all(np.diff(cell) >= 0), "not sorted by cell"
        assert hp.count(s) == before[s].shape[1]
    r1, _ = hp.p2ces_all()
    assert_close(r1, r0, rtol=1e-12, atol=1e-30 * np.abs(r0).max(), what="deposit after sort")


def test_graph_replay_equals_eager(rng):
    cfg, tables, species = _six_species_setup(pressure=300.0)
    geom = W.make_geometry(cfg)
    P = [_energetic(cfg, s, 8000, rng, 15.0 if s == 0 else 1.0) for s in range(cfg.ns)]
    ez, er = W.fields(cfg, 2e-8)
    a = _config_hp(cfg, geom, tables, P)
    b = _config_hp(cfg, geom, tables, P)
    for hp in (a, b):
        hp.set_fields(ez, er)
        hp.set_rng(5, 0)
    for k in range(3):
        a.one_step()
        b.one_step_graph()
        b.synchronize()
        ra, _ = a.field_source_sum()
        rb, _ = b.field_source_sum()
        # net Rho is a difference of species terms: tolerance scaled to its magnitude
        assert_close(ra, rb, rtol=1e-12, atol=1e-12 * np.abs(ra).max(), what=f"Rho replay {k}")
        for s in range(cfg.ns):
            assert a.count(s) == b.count(s)
    for s in range(cfg.ns):
        assert_close(sort_set(a.download(s)), sort_set(b.download(s)), rtol=1e-12, atol=1e-300, what=f"sp {s}")


def test_capacity_above_the_addressing_limit_rejected():
    geom = small_geometry()
    hp = HotPath(geom.nz, geom.nr, geom.dx, DT, 1, 1)
    with pytest.raises(IPMError):
        hp.set_species(0, -QE, ME, (1 << 29))
    hp.close()


def test_capacity_overflow_raises(rng):
    geom = small_geometry()
    hp = make_hp(geom, cap=1000)
    with pytest.raises(CapacityError):
        hp.upload(0, random_particles(geom.nz, geom.nr, 1001, rng))
    cfg, tables, species = _six_species_setup(pressure=20000.0)
    geom = W.make_geometry(cfg)
    P = [_energetic(cfg, s, 5000, rng, 40.0 if s == 0 else 10.0) for s in range(cfg.ns)]
    hp = _config_hp(cfg, geom, tables, P, cap_factor=1.0, extra=0)
    hp.set_rng(1, 0)
    with pytest.raises(CapacityError):
        hp.mcc()
    with pytest.raises(CapacityError):       # the report stays until the caller acts on it
        hp.count(0)
    # re-uploading smaller populations clears the report: the context recovers
    for s in range(cfg.ns):
        hp.upload(s, P[s][:, :1000])
    hp.set_rng(2, 0)
    hp.mcc()
    assert all(hp.count(s) >= 1000 for s in range(cfg.ns))


def test_mcc_independent_of_storage_order(rng):
    """Streams follow particle state: a permuted upload gives the same set."""
    cfg, tables, species = _six_species_setup(pressure=800.0)
    geom = W.make_geometry(cfg)
    P = [_energetic(cfg, s, 12000, rng, 25.0 if s == 0 else 3.0) for s in range(cfg.ns)]
    Q = [X[:, rng.permutation(X.shape[1])] for X in P]
    a = _config_hp(cfg, geom, tables, P)
    b = _config_hp(cfg, geom, tables, Q)
    for hp in (a, b):
        hp.set_rng(21, 4)
        hp.mcc()
    for s in range(cfg.ns):
        np.testing.assert_array_equal(sort_set(a.download(s)), sort_set(b.download(s)))


def test_set_potential_matches_oracle_set_e(oracle, rng):
    """ipm_set_potential (SetE on the device) feeds the push exactly as the
    oracle's SetE restatement does."""
    geom = small_geometry()
    nz, nr = geom.nz, geom.nr
    phi = rng.standard_normal((nz, nr)) * 30.0
    hp = make_hp(geom)
    hp.set_potential(phi)
    P = random_particles(nz, nr, 5000, rng)
    hp.upload(0, P)
    hp.move_es(0)
    ez, er = oracle.set_e(phi, DX)
    ref = oracle.push(nz, nr, DX, DT, SPECIES[0][0], SPECIES[0][1], ez, er, P)
    np.testing.assert_array_equal(hp.download(0), ref)
