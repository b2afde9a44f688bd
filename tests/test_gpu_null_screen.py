"""The null-event screen (ipm_set_null_screen) is exact.

A selected particle whose reaction-choice uniform lies at or above every
probability its table can interpolate over the relative energies the target
sampler allows is a null event; the particle pass drops it from the
collision queue (one gas) or the collision pass skips its target sampling
(several gases).  With the screen on and off the particles must agree bit for
bit and the event counts exactly, and both must agree with the oracle, which
has no screen.

Tables with sharp features make a wrong bin range show: probabilities that
alternate between neighbouring energy bins, a threshold reaction, energies
below the grid (the N < 1 row), above it (the E/emax extrapolation) and
below the target's speed bound (electrons slower than the gas's sampled
atoms), and a model-2 ion table whose reachable energies span many bins (the
sparse table's upper levels).
"""
import numpy as np
import pytest

import ipm2d
from helpers import AMU, ME, QE, random_particles, small_geometry
from ipm2d import HotPath
from ipm2d import workload as W
from ipm2d.xsec import MCCBundle

pytestmark = pytest.mark.gpu

DT = 5e-11
GAS_MASS = 88.0043 * AMU
ION_MASS = 69.0059096 * AMU


def _hashes(P):
    b = np.ascontiguousarray(P).view(np.uint64)
    h = np.full(b.shape[1], 0x9E3779B97F4A7C15, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for f in range(8):
            h ^= b[f] * np.uint64(0xBF58476D1CE4E5B9 + 2 * f)
            h = (h << np.uint64(27)) | (h >> np.uint64(37))
            h *= np.uint64(0x94D049BB133111EB)
    return np.sort(h)


def _electron_table():
    ns, emin, eint = 3001, -3.0, 0.0015
    i = np.arange(ns)
    c0 = 0.35 + 0.3 * ((i // 2) % 2)
    c1 = c0 + 0.2 * (i % 5 < 2)
    return MCCBundle(model=1, nreaction=2, nsigma=ns, emin=emin, einterval=eint,
                     emax=10.0 ** (emin + (ns - 1) * eint), collision_ratio=0.9,
                     rtype=np.array([11, 12], dtype=np.int32), resultant=np.array([0, 0], dtype=np.int32),
                     threshold=np.array([0.0, 2.0]), prob=np.stack([c0, c1], axis=1),
                     gas_mass=GAS_MASS, gas_temperature=300.0, beta_max=3.0)


def _ion_table():
    ns, emin, eint = 1001, -1.0, 0.004
    i = np.arange(ns)
    c0 = 0.3 + 0.4 * ((i // 3) % 2)
    c1 = c0 + 0.1 + 0.2 * (i % 4 == 0)
    return MCCBundle(model=2, nreaction=2, nsigma=ns, emin=emin, einterval=eint, emax=10.0 ** 3,
                     collision_ratio=0.9, rtype=np.array([101, 103], dtype=np.int32),
                     resultant=np.array([1, 1], dtype=np.int32), threshold=np.zeros(2),
                     prob=np.stack([c0, c1], axis=1), gas_mass=GAS_MASS, gas_temperature=300.0, beta_max=3.0)


def _particles(geom, n, mass, lo_ev, hi_ev, rng):
    P = random_particles(geom.nz, geom.nr, n, rng)
    ev = 10.0 ** rng.uniform(np.log10(lo_ev), np.log10(hi_ev), n)
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    P[2:5] = d * np.sqrt(2.0 * ev * QE / mass) / (geom.dx / DT)
    return P


def _hp(geom, tables, P, screen):
    hp = HotPath(geom.nz, geom.nr, geom.dx, DT, 2, 1)
    for s, (q, m) in enumerate([(-QE, ME), (QE, ION_MASS)]):
        hp.set_species(s, q, m, 4 * P[s].shape[1] + 4096)
    hp.set_geometry(geom)
    for s in range(2):
        hp.set_mcc(s, 0, tables[s])
        hp.upload(s, P[s])
    hp.set_null_screen(screen)
    hp.set_rng(0x5EED5EED1234, 3)
    hp.set_event_counting(True)
    return hp


def test_screen_is_exact_on_sharp_tables(oracle, rng):
    geom = small_geometry()
    tables = [_electron_table(), _ion_table()]
    P = [_particles(geom, 120000, ME, 1e-7, 100.0, rng), _particles(geom, 60000, ION_MASS, 1e-4, 2000.0, rng)]
    on, off = _hp(geom, tables, P, True), _hp(geom, tables, P, False)
    setup = oracle.MccSetup(geom.dx, DT, [-QE, QE], [ME, ION_MASS], {(0, 0): tables[0], (1, 0): tables[1]})
    ref, cref = oracle.mcc(setup, 0x5EED5EED1234, 3, P, counts=True)
    for k in range(3):
        on.mcc()
        off.mcc()
        c_on, c_off = on.take_mcc_counts(), off.take_mcc_counts()
        print(f"pass {k}: null events e {c_on[0, 0, 0]}, ion {c_on[1, 0, 0]}; reactions e {c_on[0, 0, 1:].sum()}, "
              f"ion {c_on[1, 0, 1:].sum()}", flush=True)
        np.testing.assert_array_equal(c_on, c_off)
        if k == 0:
            np.testing.assert_array_equal(c_on, cref)
            assert c_on[0, 0, 0] > 20000 and c_on[0, 0, 1] > 5000 and c_on[0, 0, 2] > 1000
            assert c_on[1, 0, 0] > 5000 and c_on[1, 0, 1:].sum() > 5000
        for s in range(2):
            a, b = on.download(s), off.download(s)
            assert a.shape == b.shape, (k, s)
            np.testing.assert_array_equal(_hashes(a), _hashes(b), err_msg=f"pass {k} species {s}")
            if k == 0:
                assert a.shape == ref[s].shape
    on.close()
    off.close()


def test_screen_is_exact_in_fused_steps():
    """c1 (e / CF3+ / F-, the electrons' table with its thresholds): five
    fused steps as a graph with the screen, eagerly without it."""
    cfg = W.make_config("c1")
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    rng = np.random.default_rng(cfg.seed)
    parts = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    hps = []
    for screen in (True, False):
        hp, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.5)
        hp.set_null_screen(screen)
        hp.set_event_counting(True)
        hp.take_mcc_counts()
        hps.append(hp)
    a, b = hps
    for _ in range(5):
        a.one_step_graph()
        b.one_step()
    a.synchronize()
    ca, cb = a.take_mcc_counts(), b.take_mcc_counts()
    np.testing.assert_array_equal(ca, cb)
    assert ca[0, 0, 0] > 1000 and ca[0, 0, 1:].sum() > 100
    for s in range(cfg.ns):
        assert a.count(s) == b.count(s)
        np.testing.assert_array_equal(_hashes(a.download(s)), _hashes(b.download(s)), err_msg=f"species {s}")
    a.close()
    b.close()
