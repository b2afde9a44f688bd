"""GPU parity of weight rezoning (ipm_split_merge, R/src/base/Rezoning.F90:53-92)
against the oracle.  Every decision depends on the particle alone (its weight,
radius and state-keyed draw), so the device result equals the oracle's as a
set, bit for bit, whatever the storage order:

* one call on a mixture of merge candidates, split candidates and untouched
  particles, then a second call on the device's own output (dead slots and
  appended copies in place), each against the oracle;
* split only and merge only;
* the full-size electron bundle (257x513 nodes, 200 per cell, 26M): counts
  add up, the merge keep fraction is WQ/ref = 0.2 within 5 binomial sigma,
  and the multiset of per-particle 64-bit hashes equals the oracle's;
* copies beyond the capacity raise CapacityError.
"""
import numpy as np
import pytest

from helpers import QE, ME, random_particles, small_geometry, sort_set
from ipm2d import CapacityError, HotPath
from ipm2d import workload as W

pytestmark = pytest.mark.gpu

DT = 5e-11
ALPHA, LIMIT = 0.01, 0.34


def _mixture(geom, n, rng, weight):
    P = random_particles(geom.nz, geom.nr, n, rng)
    grid_r = geom.nr - 1.0
    ref = weight * ((1.0 - ALPHA) * P[1] / grid_r + ALPHA)
    u = rng.uniform(size=n)
    frac = np.where(u < 0.3, rng.uniform(0.005, 0.34, n),
                    np.where(u < 0.6, rng.uniform(2.0, 4.7, n), rng.uniform(0.34, 2.0, n)))
    P[7] = frac * ref
    return P, grid_r


def _hashes(P):
    """Sorted 64-bit hashes of the particles' 8 bit patterns (order-free)."""
    b = np.ascontiguousarray(P).view(np.uint64)
    h = np.full(b.shape[1], 0x9E3779B97F4A7C15, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for f in range(8):
            h ^= b[f] * np.uint64(0xBF58476D1CE4E5B9 + 2 * f)
            h = (h << np.uint64(27)) | (h >> np.uint64(37))
            h *= np.uint64(0x94D049BB133111EB)
    return np.sort(h)


def _hp(geom, cap):
    hp = HotPath(geom.nz, geom.nr, geom.dx, DT, 1, 1)
    hp.set_species(0, -QE, ME, cap)
    hp.set_geometry(geom)
    return hp


@pytest.mark.parametrize("split,merge", [(True, True), (True, False), (False, True)])
def test_split_merge_matches_oracle(oracle, rng, split, merge):
    geom = small_geometry()
    weight = 3.0e5
    P, grid_r = _mixture(geom, 30000, rng, weight)
    hp = _hp(geom, 120000)
    hp.upload(0, P)
    seed, step = 77, 5
    mode = (1 if split else 0) | (2 if merge else 0)
    want, st_o = oracle.split_merge(P, weight, ALPHA, grid_r, LIMIT, seed, step, 0, mode=mode)
    hp.set_rng(seed, step)
    st_g = hp.split_merge(0, weight, ALPHA, grid_r, LIMIT, split=split, merge=merge)
    assert st_g == tuple(int(x) for x in st_o)
    got = hp.download(0)
    assert hp.count(0) == want.shape[1]
    np.testing.assert_array_equal(sort_set(got), sort_set(want))
    if split:
        assert st_g[0] > 1000
    if merge:
        assert st_g[1] > 100 and st_g[2] > 100
    # again on the device's own output (dead slots and copies in the arrays)
    want2, st_o2 = oracle.split_merge(want, weight * 0.97, ALPHA, grid_r, LIMIT, seed, step + 1, 0, mode=mode)
    hp.set_rng(seed, step + 1)
    st_g2 = hp.split_merge(0, weight * 0.97, ALPHA, grid_r, LIMIT, split=split, merge=merge)
    assert st_g2 == tuple(int(x) for x in st_o2)
    np.testing.assert_array_equal(sort_set(hp.download(0)), sort_set(want2))
    hp.close()


def test_split_merge_full_size(oracle):
    cfg = W.make_config("c2")
    geom = W.make_geometry(cfg)
    rng = np.random.default_rng(cfg.seed)
    P = W.load_species(cfg, geom, 0, rng)
    n = P.shape[1]
    grid_r = cfg.nr - 1.0
    weight = float(np.median(P[7]))
    ref = weight * ((1.0 - ALPHA) * P[1] / grid_r + ALPHA)
    frac = rng.choice([0.2, 1.0, 2.5], size=n, p=[0.2, 0.6, 0.2])
    P[7] = frac * ref
    hp = HotPath(cfg.nz, cfg.nr, cfg.dx, cfg.dt, 1, 1)
    hp.set_species(0, -QE, ME, int(n * 1.3) + 4096)
    hp.set_geometry(geom)
    hp.upload(0, P)
    hp.set_rng(cfg.seed, 0)
    copies, kept, deleted = hp.split_merge(0, weight, ALPHA, grid_r, LIMIT)
    nsplit = int((frac == 2.5).sum())
    ncand = int((frac == 0.2).sum())
    assert copies == nsplit                    # int(2.5) - 1 = 1 copy each
    assert kept + deleted == ncand
    assert hp.count(0) == n + copies - deleted
    sigma = np.sqrt(ncand * 0.2 * 0.8)
    assert abs(kept - 0.2 * ncand) <= 5 * sigma
    want, st = oracle.split_merge(P, weight, ALPHA, grid_r, LIMIT, cfg.seed, 0, 0, fast=True)
    assert (copies, kept, deleted) == tuple(int(x) for x in st)
    got = hp.download(0)
    assert got.shape == want.shape
    np.testing.assert_array_equal(_hashes(got), _hashes(want))
    hp.close()


def test_split_beyond_capacity_raises(rng):
    geom = small_geometry()
    weight = 1.0e5
    P = random_particles(geom.nz, geom.nr, 1000, rng)
    grid_r = geom.nr - 1.0
    P[7] = 3.0 * weight * ((1.0 - ALPHA) * P[1] / grid_r + ALPHA)
    hp = _hp(geom, 1500)
    hp.upload(0, P)
    with pytest.raises(CapacityError):
        hp.split_merge(0, weight, ALPHA, grid_r, LIMIT)
    hp.close()
