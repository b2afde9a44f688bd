"""Cell sort of a drifted population (ipm_sort; the reference has no
counterpart: its bundles keep insertion order, and neither the deposit nor
the collisions depend on slot order).

c1 (65x129 nodes, 2.46 M particles, e / CF3+ / F-) after 30 fused steps, so
that tiles hold many runs of different cells and neighbouring tiles share
cells:

* each species keeps exactly its live particles (multiset of 64-bit hashes
  of the 8 fields) and its live count;
* slots are in nondecreasing cell order;
* the deposit after the sort equals the deposit before it to 1e-12 of its
  magnitude (atomic summation order);
* the next 5 fused steps from the sorted state give the same particle
  multisets as the same steps from an unsorted twin (collision draws follow
  the particle's state, not its slot).
"""
import numpy as np
import pytest

import ipm2d
from ipm2d import workload as W

pytestmark = pytest.mark.gpu


def _hashes(P):
    b = np.ascontiguousarray(P).view(np.uint64)
    h = np.full(b.shape[1], 0x9E3779B97F4A7C15, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for f in range(8):
            h ^= b[f] * np.uint64(0xBF58476D1CE4E5B9 + 2 * f)
            h = (h << np.uint64(27)) | (h >> np.uint64(37))
            h *= np.uint64(0x94D049BB133111EB)
    return np.sort(h)


@pytest.mark.parametrize("tile", [2048, 4096])
def test_sort_after_drift(tile):
    cfg = W.make_config("c1")
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    rng = np.random.default_rng(cfg.seed)
    parts = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    a, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.5)
    b, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.5)
    a.set_sort_tile(tile)
    for _ in range(30):
        a.one_step()
        b.one_step()
    before = [a.download(s) for s in range(cfg.ns)]
    for s in range(cfg.ns):
        np.testing.assert_array_equal(_hashes(before[s]), _hashes(b.download(s)), err_msg=f"twin {s}")
    r0, _ = a.p2ces_all()
    a.sort(-1)
    for s in range(cfg.ns):
        after = a.download(s)
        assert after.shape[1] == before[s].shape[1] == a.count(s)
        np.testing.assert_array_equal(_hashes(after), _hashes(before[s]), err_msg=f"species {s}")
        cell = (np.ceil(after[0]) - 1) + (np.maximum(np.ceil(after[1]), 1) - 1) * (cfg.nz - 1)
        assert np.all(np.diff(cell) >= 0), f"species {s} not in cell order"
    r1, _ = a.p2ces_all()
    np.testing.assert_allclose(r1, r0, rtol=1e-12, atol=1e-12 * np.abs(r0).max())
    for _ in range(5):
        a.one_step()
        b.one_step()
    for s in range(cfg.ns):
        np.testing.assert_array_equal(_hashes(a.download(s)), _hashes(b.download(s)),
                                      err_msg=f"species {s} after the sort")
    a.close()
    b.close()


@pytest.mark.parametrize("tile", [2048, 4096])
def test_sort_of_shuffled_population(tile):
    """Slots in random order: every tile spans nearly the whole grid, so the
    tiles are radix-sorted rather than counted in LDS bins.  The sorted set
    and its cell order must not depend on the tile size."""
    cfg = W.make_config("c1")
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    rng = np.random.default_rng(7)
    parts = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    parts = [P[:, rng.permutation(P.shape[1])] for P in parts]
    a, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.5)
    a.set_sort_tile(tile)
    a.sort(-1)
    for s in range(cfg.ns):
        P = a.download(s)
        assert P.shape[1] == parts[s].shape[1] == a.count(s)
        np.testing.assert_array_equal(_hashes(P), _hashes(parts[s]), err_msg=f"species {s}")
        cell = (np.ceil(P[0]) - 1) + (np.maximum(np.ceil(P[1]), 1) - 1) * (cfg.nz - 1)
        assert np.all(np.diff(cell) >= 0), f"species {s} not in cell order"
    a.close()


@pytest.mark.parametrize("tile", [2048, 4096])
def test_sort_of_degenerate_boxes(tile):
    """Tiles whose cells form degenerate boxes: one cell (a box of one bin),
    one full row of cells in z, and the grid's four corner cells (a box of
    the whole grid: radix-sorted on every bit of the cell index), with
    particles on the outermost cells' faces.  Each species keeps its
    particles and ends in cell order."""
    cfg = W.make_config("c1")
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    rng = np.random.default_rng(11)
    parts = [W.load_species(cfg, geom, s, rng)[:, :6000] for s in range(cfg.ns)]
    n = parts[0].shape[1]
    parts[0][0] = 10.0 + rng.uniform(0.01, 0.99, n)              # all in cell (11, 21)
    parts[0][1] = 20.0 + rng.uniform(0.01, 0.99, n)
    n = parts[1].shape[1]
    parts[1][0] = rng.uniform(0.0, cfg.nz - 1.0, n)               # one row, every z
    parts[1][1] = 5.0 + rng.uniform(0.01, 0.99, n)
    n = parts[2].shape[1]
    corner = rng.integers(0, 4, n)
    parts[2][0] = np.where(corner & 1, cfg.nz - 1.0 - rng.uniform(0.0, 0.5, n), rng.uniform(0.0, 0.5, n))
    parts[2][1] = np.where(corner & 2, cfg.nr - 1.0 - rng.uniform(0.0, 0.5, n), rng.uniform(0.0, 0.5, n))
    parts = [P[:, rng.permutation(P.shape[1])] for P in parts]
    a, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.5)
    a.set_sort_tile(tile)
    a.sort(-1)
    for s in range(cfg.ns):
        P = a.download(s)
        assert P.shape[1] == parts[s].shape[1] == a.count(s)
        np.testing.assert_array_equal(_hashes(P), _hashes(parts[s]), err_msg=f"species {s}")
        iz = np.clip(np.ceil(P[0]), 1, cfg.nz - 1) - 1
        ir = np.clip(np.ceil(P[1]), 1, cfg.nr - 1) - 1
        cell = iz + ir * (cfg.nz - 1)
        assert np.all(np.diff(cell) >= 0), f"species {s} not in cell order"
    a.close()


def test_sort_interval_in_graph_replays():
    """ipm_set_sort_interval: the sort every 7 steps replayed as a graph of
    its own inside ipm_step_graph equals explicit eager sorts at the same
    points (particle multisets and counts exactly, the deposit to 1e-12),
    keeps cell order, and survives a capacity change (both graphs are
    recaptured for the new tile count)."""
    cfg = W.make_config("c1")
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    rng = np.random.default_rng(cfg.seed)
    parts = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    a, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.5)
    b, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.5)
    a.set_sort_interval(7)
    for k in range(23):
        a.one_step_graph()
        b.one_step()
        if (k + 1) % 7 == 0:
            b.sort(-1)
    a.synchronize()
    for s in range(cfg.ns):
        pa, pb = a.download(s), b.download(s)
        assert a.count(s) == b.count(s) == pa.shape[1]
        np.testing.assert_array_equal(_hashes(pa), _hashes(pb), err_msg=f"species {s}")
    ra, _ = a.p2ces_all()
    rb, _ = b.p2ces_all()
    np.testing.assert_allclose(ra, rb, rtol=1e-12, atol=1e-12 * np.abs(rb).max())
    # one more replay reaches the fourth sort (step 28): cell order holds
    for _ in range(5):
        a.one_step_graph()
    a.synchronize()
    for s in range(cfg.ns):
        P = a.download(s)
        cell = (np.ceil(P[0]) - 1) + (np.maximum(np.ceil(P[1]), 1) - 1) * (cfg.nz - 1)
        assert np.all(np.diff(cell) >= 0), f"species {s} not in cell order after the graph sort"
    # a larger capacity for species 1 drops both graphs; the recaptured ones
    # cover the new tile count
    P1 = a.download(1)
    q, m = W.species_params(cfg, 1)
    a.set_species(1, q, m, int(P1.shape[1] * 2.5) + 4096)
    a.upload(1, P1)
    a.set_sort_interval(3)
    for _ in range(6):
        a.one_step_graph()
    a.synchronize()
    P = a.download(1)
    assert P.shape[1] == a.count(1) > 0.9 * P1.shape[1]
    cell = (np.ceil(P[0]) - 1) + (np.maximum(np.ceil(P[1]), 1) - 1) * (cfg.nz - 1)
    assert np.all(np.diff(cell) >= 0), "species 1 not in cell order after the recaptured sort"
    # the eager form sorts on the same schedule
    b.set_sort_interval(7)
    for _ in range(7):
        b.one_step()
    b.synchronize()
    for s in range(cfg.ns):
        P = b.download(s)
        cell = (np.ceil(P[0]) - 1) + (np.maximum(np.ceil(P[1]), 1) - 1) * (cfg.nz - 1)
        assert np.all(np.diff(cell) >= 0), f"species {s} not in cell order after the eager interval sort"
    a.close()
    b.close()
