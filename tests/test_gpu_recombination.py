"""GPU parity of volume recombination (ipm_recombine,
R/src/mc/Recombination.F90:164-273) against the oracle.  Given the same
storage order, every density, budget and draw is the oracle's, so the
surviving particles are equal as sets, bit for bit:

* c1 geometry (65x129 nodes, e / CF3+ / F-, 100 per cell each) after 5 full
  steps (products appended, dead slots in place), reactions F- + CF3+
  (constant rate) and e + CF3+ (temperature form), rates chosen so that each
  cell's budget is about a fifth of its particles: sets and deletion counts
  equal, then a further step runs on the regrouped bundles;
* the full-size c2 bundles (257x513 nodes, 50M particles): deletion counts
  and the multiset of per-particle 64-bit hashes equal the oracle's;
* invalid reactions are rejected;
* the scenario of the reference's own recombination test
  (R/test/test_recb.F90): survivors and deletion counts equal the oracle's.
"""
import numpy as np
import pytest

import ipm2d
from ipm2d import IPMError
from ipm2d import workload as W

pytestmark = pytest.mark.gpu

REACTIONS = [(2, 1, 0), (0, 1, 1)]


def _hashes(P):
    b = np.ascontiguousarray(P).view(np.uint64)
    h = np.full(b.shape[1], 0x9E3779B97F4A7C15, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for f in range(8):
            h ^= b[f] * np.uint64(0xBF58476D1CE4E5B9 + 2 * f)
            h = (h << np.uint64(27)) | (h >> np.uint64(37))
            h *= np.uint64(0x94D049BB133111EB)
    return np.sort(h)


def _rates(oracle, cfg, parts, masses, ts):
    """Rates that make each cell's budget about 0.2 of its particles."""
    d, t = zip(*[oracle.recomb_density(cfg.nz, cfg.nr, cfg.dx, cfg.dt, masses[s], parts[s], fast=True)
                 for s in range(cfg.ns)])
    med = lambda a: float(np.median(a[a > 0]))          # noqa: E731
    r0 = 0.2 / (med(d[1]) * ts)
    r1 = 0.2 / (med(d[1]) * ts * np.sqrt(med(t[0])) * med(t[1]))
    return [r0, r1]


def _setup(name, steps, capacity_factor=1.5):
    cfg = W.make_config(name)
    hp, geom, tables, parts = ipm2d.from_config(cfg, capacity_factor=capacity_factor)
    for _ in range(steps):
        hp.one_step()
    masses = [W.species_params(cfg, s)[1] for s in range(cfg.ns)]
    return cfg, hp, masses


def test_recombination_matches_oracle(oracle):
    cfg, hp, masses = _setup("c1", 5)
    before = [hp.download(s) for s in range(cfg.ns)]
    ts = 50 * cfg.dt
    rates = _rates(oracle, cfg, before, masses, ts)
    seed, step = 11, 123
    want, dl_o = oracle.recombine(cfg.nz, cfg.nr, cfg.dx, cfg.dt, before, masses, REACTIONS, rates, ts, seed, step,
                                  fast=True)
    hp.set_rng(seed, step)
    dl_g = hp.recombine(REACTIONS, rates, ts)
    assert tuple(dl_g) == tuple(dl_o)
    assert dl_g[0] > 1000 and dl_g[1] > 1000 and dl_g[2] > 1000
    for s in range(cfg.ns):
        got = hp.download(s)
        assert hp.count(s) == want[s].shape[1] == before[s].shape[1] - dl_g[s]
        assert got.shape == want[s].shape
        np.testing.assert_array_equal(_hashes(got), _hashes(want[s]))
    hp.one_step()
    for s in range(cfg.ns):
        assert hp.count(s) > 0
    hp.close()


def test_recombination_full_size(oracle):
    cfg, hp, masses = _setup("c2", 0, capacity_factor=1.2)
    before = [hp.download(s) for s in range(cfg.ns)]
    ts = 50 * cfg.dt
    rates = _rates(oracle, cfg, before, masses, ts)
    want, dl_o = oracle.recombine(cfg.nz, cfg.nr, cfg.dx, cfg.dt, before, masses, REACTIONS, rates, ts, cfg.seed, 9,
                                  fast=True)
    del before
    hp.set_rng(cfg.seed, 9)
    dl_g = hp.recombine(REACTIONS, rates, ts)
    assert tuple(dl_g) == tuple(dl_o)
    for s in range(cfg.ns):
        got = hp.download(s)
        assert got.shape == want[s].shape
        np.testing.assert_array_equal(_hashes(got), _hashes(want[s]))
        del got
    hp.close()


def test_invalid_reactions_rejected():
    cfg, hp, masses = _setup("c1", 0)
    with pytest.raises(IPMError):
        hp.recombine([(0, 3, 0)], [1e-13], 1e-9)
    with pytest.raises(IPMError):
        hp.recombine([(0, 1, 2)], [1e-13], 1e-9)
    assert tuple(hp.recombine([], [], 1e-9)) == (0, 0, 0)
    hp.close()


def test_reference_recb_scenario(oracle):
    """The scenario of the reference's own recombination test
    (R/test/test_recb.F90): 4 bundles (e, F-, CF3+, CF3-) of 500 000
    particles uniform over 10 x 20 cells of 0.1 m (Z in [0, 10), R in
    [0, 20), WQ = R, at rest), reactions e + CF3+ and F- + CF3+ at constant
    rates 10000 and 5000, one call over 1e-10 s.  The reference writes the
    density maps and does not check them; here the GPU's survivors and
    deletion counts equal the oracle's, bit for bit."""
    from ipm2d import HotPath, geometry
    from ipm2d import xsec
    gen = np.random.default_rng(21)
    nz, nr, dx, dt = 11, 21, 0.1, 1e-10
    q = xsec.ELECTRON_CHARGE
    amu = xsec.ATOMIC_MASS
    species = [(-q, xsec.ELECTRON_MASS), (-q, 18.9984030 * amu), (q, 69.0059096 * amu), (-q, 69.0059096 * amu)]
    hp = HotPath(nz, nr, dx, dt, 4, 1)
    parts = []
    for s, (qs, ms) in enumerate(species):
        hp.set_species(s, qs, ms, 600000)
        P = np.zeros((8, 500000))
        P[0] = 10.0 * gen.random(500000)
        P[1] = 20.0 * gen.random(500000)
        P[7] = P[1]
        parts.append(P)
    hp.set_geometry(geometry.build_geometry(nz, nr, dx, np.zeros((nz - 1, nr - 1), np.int32), nmetal=0, ndiel=0))
    for s in range(4):
        hp.upload(s, parts[s])
    reactions, rates = [(0, 2, 0), (1, 2, 0)], [10000.0, 5000.0]
    masses = [m for _, m in species]
    want, dl_o = oracle.recombine(nz, nr, dx, dt, parts, masses, reactions, rates, 1e-10, 5, 0, fast=True)
    hp.set_rng(5, 0)
    dl_g = hp.recombine(reactions, rates, 1e-10)
    assert tuple(dl_g) == tuple(dl_o)
    assert dl_g[0] > 0 and dl_g[1] > 0 and dl_g[2] > 0 and dl_g[3] == 0
    for s in range(4):
        got = hp.download(s)
        assert got.shape == want[s].shape
        np.testing.assert_array_equal(_hashes(got), _hashes(want[s]))
    hp.close()
