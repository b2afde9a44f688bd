"""Pins of the oracle's absorption rules and MCC against hand-worked cases,
closed forms and statistics.

Absorption (R/src/pic/particleboundary/ParticleMaterial.F90:20-88): one
hand-built particle per rule with its exact outcome and tally.

MCC: Philox known answers; the Maxwellian sampler's mean energy 3/2 kT; the
Vahedi and isotropic cos(chi) CDFs; selection fraction = CollisionRatio;
channel shares = sigma ratios at fixed energy; excitation removes the
threshold; ionization shares E - threshold between the two electrons;
attachment keeps the electron and adds one negative ion; products sit at the
parent's position with its weight.  Statistical bounds are 6 standard errors.
"""
import math

import numpy as np
import pytest

from helpers import AMU, ME, QE
from ipm2d import geometry as geo
from ipm2d import xsec

# ------------------------------------------------------------ absorption


def _geom():
    nz, nr = 12, 8
    cm = np.zeros((nz - 1, nr - 1), dtype=np.int32)
    cm[0, :] = 1                     # metal 1: cells iz = 1
    cm[nz - 2, :] = 2                # metal 2: cells iz = nz-1
    cm[5, 3] = -1                    # a dielectric cell at (iz, ir) = (6, 4) (1-based)
    return geo.build_geometry(nz, nr, 1e-3, cm, nmetal=2, ndiel=1, permittivity={-1: 3.0})


def test_absorption_rule_table(oracle):
    g = _geom()
    nz, nr = g.nz, g.nr
    # hedge/vedge of the dielectric cell (6, 4): hedge(6,4) vacuum->diel, hedge(7,4) diel->vacuum,
    # vedge(6,4) vacuum->diel, vedge(6,5) diel->vacuum
    assert g.hedge_type[5, 3] == geo.EDGE_V2D and g.hedge_type[6, 3] == geo.EDGE_D2V
    assert g.vedge_type[5, 3] == geo.EDGE_V2D and g.vedge_type[5, 4] == geo.EDGE_D2V
    q = -2.0
    rows = [
        # Z,    R,    outcome
        (0.0, 2.0),          # Z <= 0: deleted, no tally
        (nz - 1.0, 2.0),     # Z >= Nz-1: deleted
        (4.0, nr - 1.0),     # R >= Nr-1: deleted
        (4.5, -0.25),        # R <= 0: reflected
        (0.5, 2.5),          # metal 1 cell: deleted, tally metal 1
        (nz - 1.5, 3.5),     # metal 2 cell: deleted, tally metal 2
        (5.25, 3.75),        # dielectric cell, hedge(NzL, NrL) rule: sigma on (6,4),(6,5)
        (4.5, 4.5),          # vacuum: kept
    ]
    P = np.zeros((8, len(rows)))
    for k, (z, r) in enumerate(rows):
        P[0, k], P[1, k] = z, r
    P[3] = -1.5          # Vr
    P[6] = -0.5          # Ar
    P[7] = np.arange(1, len(rows) + 1, dtype=float)
    out, mc, sg = oracle.absorb(g, q, P)
    kept = sorted(out[7].tolist())
    assert kept == [4.0, 8.0]
    refl = out[:, out[7] == 4.0][:, 0]
    assert refl[1] == 0.25 and refl[3] == 1.5 and refl[6] == 0.5
    assert mc[0] == 5.0 * q and mc[1] == 6.0 * q
    RL = 4.0 - 3.75
    assert sg[5, 3] == pytest.approx(RL * 7.0 * q) and sg[5, 4] == pytest.approx((1 - RL) * 7.0 * q)
    assert np.count_nonzero(sg) == 2


def test_absorption_delete_order_is_swap_with_last(oracle):
    """DelOne copies the last particle into the hole (ParticleBundle.f90:562)."""
    g = _geom()
    P = np.zeros((8, 5))
    P[0] = [2.5, -1.0, 3.5, -1.0, 4.5]
    P[1] = 2.5
    P[7] = [1, 2, 3, 4, 5]
    out, _, _ = oracle.absorb(g, 1.0, P)
    assert out[7].tolist() == [1.0, 5.0, 3.0]


# ------------------------------------------------------------------- RNG


def test_philox_known_answers(oracle):
    # Random123 kat_vectors for philox4x32-10
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_uniform_is_open_interval_and_flat(oracle):
    u = np.array([oracle.uniform(7, 3, 1, i, 0, 0, 0) for i in range(40000)])
    assert u.min() > 0.0 and u.max() < 1.0
    h, _ = np.histogram(u, bins=20, range=(0, 1))
    exp = len(u) / 20
    assert ((h - exp) ** 2 / exp).sum() < 60.0          # chi2(19) far tail


# ------------------------------------------------------------- samplers


def test_maxwellian_mean_energy(oracle):
    m, T = 69.0 * AMU, 300.0
    e = []
    for i in range(30000):
        vz, vr, vt, k = oracle.vel_maxwell(m, T, 5, i, 0, 0, 9)
        e.append(0.5 * m * (vz * vz + vr * vr + vt * vt))
    e = np.array(e) / (xsec.KB * T)
    se = e.std() / math.sqrt(len(e))
    assert abs(e.mean() - 1.5) < 6 * se


def test_maxwellian_acceptance_rate(oracle):
    """VelMaxInit's rejection loop (ParticleOne.f90:76-95) accepts a pair
    when R1^2 <= -e R2 ln R2.  Its acceptance probability is
    sqrt(e) * int_0^1 sqrt(-r ln r) dr = sqrt(e) Gamma(3/2) / (3/2)^(3/2)
    = 0.7953; each try takes two draws and the direction two more, so the
    mean draw count per sample is 2/p + 2.  (The accepted speeds do not
    depend on the e factor -- it rescales the acceptance uniformly -- so the
    mean-energy test cannot see it; the draw count can.)"""
    p = math.sqrt(math.e) * math.gamma(1.5) / 1.5 ** 1.5
    n = 30000
    tries = np.array([(oracle.vel_maxwell(69.0 * AMU, 300.0, 11, i, 0, 0, 9)[3] - 2) / 2 for i in range(n)])
    assert np.all(tries >= 1) and np.all(tries == np.round(tries))
    se = math.sqrt((1.0 - p) / p ** 2 / n)
    assert abs(tries.mean() - 1.0 / p) < 6 * se, (tries.mean(), 1.0 / p)


@pytest.mark.parametrize("energy", [0.5, 10.0, 200.0])
def test_vahedi_cos_chi_cdf(oracle, energy):
    c = np.array([oracle.lib().ora_cos_chi(1, energy, 3, i, 0) for i in range(20000)])
    assert c.min() >= -1.0 and c.max() <= 1.0
    for x in (-0.5, 0.0, 0.5, 0.9):
        cdf = 1.0 - math.log((energy + 2.0 - x * energy) / 2.0) / math.log1p(energy)
        p = (c <= x).mean()
        assert abs(p - cdf) < 6 * math.sqrt(cdf * (1 - cdf) / len(c)) + 1e-9


def test_isotropic_cos_chi_uniform(oracle):
    c = np.array([oracle.lib().ora_cos_chi(0, 1.0, 4, i, 0) for i in range(20000)])
    assert abs(c.mean()) < 6 * (1 / math.sqrt(3)) / math.sqrt(len(c))


# ------------------------------------------------------------------ MCC

DX, DT = 1e-3, 5e-11


def _electron_setup(oracle, reactions, density=1e22):
    gas = dict(mass=88.0043128 * AMU, temperature=300.0, density=density, natom=5.0, beta_max=3.0)
    sf = xsec.SigmaFile(1, reactions)
    b = xsec.build_bundle(sf, ME, 1.0, gas, DT)
    ion = xsec.MCCBundle(model=0)
    return b, oracle.MccSetup(DX, DT, [-QE, QE, -QE], [ME, 69 * AMU, 19 * AMU], {(0, 0): b, (1, 0): ion})


def _mono(n, ev, rng, z=5.5, r=3.5):
    P = np.zeros((8, n))
    P[0], P[1], P[7] = z, r, 2.0
    v = math.sqrt(2 * ev * QE / ME) / (DX / DT)
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    P[2:5] = d * v
    return P


def _energy_ev(P, m=ME):
    v = np.linalg.norm(P[2:5], axis=0) * (DX / DT)
    return 0.5 * m * v * v / QE


def test_selection_fraction_is_collision_ratio(oracle, rng):
    e = xsec.log_grid()
    b, st = _electron_setup(oracle, [xsec.Reaction(0, 11, 0, 0.0, e, np.full(e.size, 1e-19))])
    n = 100000
    P = _mono(n, 5.0, rng)
    out, cnt = oracle.mcc(st, 1, 0, [P, np.zeros((8, 0)), np.zeros((8, 0))], counts=True)
    sel = cnt[0, 0].sum()
    p = b.collision_ratio
    assert abs(sel - n * p) < 6 * math.sqrt(n * p * (1 - p))


def test_channel_shares_follow_cross_sections(oracle, rng):
    e = xsec.log_grid()
    r1 = xsec.Reaction(0, 11, 0, 0.0, e, np.full(e.size, 3e-20))
    r2 = xsec.Reaction(0, 12, 0, 1.0, e, np.where(e > 1.0, 1e-20, 0.0))
    b, st = _electron_setup(oracle, [r1, r2], density=5e22)
    P = _mono(200000, 20.0, rng)
    _, cnt = oracle.mcc(st, 2, 0, [P, np.zeros((8, 0)), np.zeros((8, 0))], counts=True)
    n1, n2 = cnt[0, 0, 1], cnt[0, 0, 2]
    frac = n2 / (n1 + n2)
    se = math.sqrt(0.25 * 0.75 / (n1 + n2))
    assert abs(frac - 0.25) < 6 * se + 0.01        # gas motion smears E slightly


def test_excitation_removes_threshold(oracle, rng):
    e = xsec.log_grid()
    th = 7.0
    r = xsec.Reaction(0, 12, 0, th, e, np.where(e > th, 5e-19, 0.0))
    b, st = _electron_setup(oracle, [r], density=1e23)
    P = _mono(40000, 15.0, rng)
    out, cnt = oracle.mcc(st, 3, 0, [P, np.zeros((8, 0)), np.zeros((8, 0))], counts=True)
    E1 = _energy_ev(out[0])
    hit = E1 < 12.0
    assert hit.sum() == cnt[0, 0, 1] > 1000
    assert abs(E1[hit].mean() - (15.0 - th)) < 0.05


def test_ionization_shares_energy_and_creates_pair(oracle, rng):
    e = xsec.log_grid()
    th = 16.0
    r = xsec.Reaction(0, 13, 1, th, e, np.where(e > th, 5e-19, 0.0))
    b, st = _electron_setup(oracle, [r], density=1e23)
    n = 30000
    P = _mono(n, 40.0, rng)
    out, cnt = oracle.mcc(st, 4, 0, [P, np.zeros((8, 0)), np.zeros((8, 0))], counts=True)
    nion = cnt[0, 0, 1]
    assert out[0].shape[1] == n + nion and out[1].shape[1] == nion
    E_prim = _energy_ev(out[0][:, :n])
    E_sec = _energy_ev(out[0][:, n:])
    hit = E_prim < 39.0
    assert hit.sum() == nion
    tot = E_prim[hit].sum() + E_sec.sum()
    assert abs(tot / nion - (40.0 - th)) < 0.05
    # the secondary sits at the primary's place with its weight; the ion too
    assert np.all(out[0][0, n:] == 5.5) and np.all(out[1][1] == 3.5) and np.all(out[1][7] == 2.0)
    # the ion is a gas Maxwellian sample (mean energy 3/2 kT of CF4 at 300 K)
    Ei = _energy_ev(out[1], 88.0043128 * AMU) * QE / (xsec.KB * 300.0)
    assert abs(Ei.mean() - 1.5) < 6 * Ei.std() / math.sqrt(nion)


def test_attachment_keeps_electron_and_adds_ion(oracle, rng):
    e = xsec.log_grid()
    r = xsec.Reaction(0, 14, 2, 0.0, e, np.full(e.size, 1e-19))
    b, st = _electron_setup(oracle, [r], density=1e23)
    n = 20000
    P = _mono(n, 5.0, rng)
    out, cnt = oracle.mcc(st, 5, 0, [P, np.zeros((8, 0)), np.zeros((8, 0))], counts=True)
    na = cnt[0, 0, 1]
    assert na > 100 and out[0].shape[1] == n and out[2].shape[1] == na and out[1].shape[1] == 0
    # electrons only pass through the unit conversion and back
    np.testing.assert_allclose(out[0][2:5], P[2:5], rtol=1e-15)


def test_stream_follows_state_not_slot(oracle, rng):
    """Permuting the particle array permutes the MCC result the same way."""
    e = xsec.log_grid()
    b, st = _electron_setup(oracle, [xsec.Reaction(0, 11, 0, 0.0, e, np.full(e.size, 1e-19))], density=1e23)
    P = _mono(5000, 5.0, rng)
    P[0] = rng.uniform(1, 9, P.shape[1])
    perm = rng.permutation(P.shape[1])
    out1, _ = oracle.mcc(st, 9, 2, [P, np.zeros((8, 0)), np.zeros((8, 0))])
    out2, _ = oracle.mcc(st, 9, 2, [P[:, perm], np.zeros((8, 0)), np.zeros((8, 0))])
    np.testing.assert_array_equal(out1[0][:, perm], out2[0])


def test_state_keys_distinguish_parent_and_secondary(oracle):
    a = oracle.state_key(1.5, 2.5, 0.1, 0.2, 0.3)
    b = oracle.state_key(1.5, 2.5, 0.1 * 0.7, 0.2 * 0.7, 0.3 * 0.7)
    assert a != b


def test_paired_draws_are_independent_uniforms(oracle):
    """Draws 2j and 2j+1 come from one Philox block (words 1,0 and 3,2): the
    pairs fill the unit square evenly, and consecutive blocks differ."""
    n = 40000
    a = np.array([oracle.uniform(11, 2, 3, i, 7, 9, 0) for i in range(n)])
    b = np.array([oracle.uniform(11, 2, 3, i, 7, 9, 1) for i in range(n)])
    c = np.array([oracle.uniform(11, 2, 3, i, 7, 9, 2) for i in range(n)])
    h, _, _ = np.histogram2d(a, b, bins=10, range=[[0, 1], [0, 1]])
    exp = n / 100
    assert ((h - exp) ** 2 / exp).sum() < 180.0             # chi2(99) far tail
    assert abs(np.corrcoef(a, b)[0, 1]) < 6 / math.sqrt(n)
    assert abs(np.corrcoef(b, c)[0, 1]) < 6 / math.sqrt(n)
    assert not np.any(a == b) and not np.any(b == c)
