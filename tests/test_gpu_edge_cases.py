"""Edge populations through the fused step, the cell sort and the graph
(65 x 33 nodes, e / CF3+ / F- with the synthetic tables):

* an empty species and a single particle beside a species one slot past a
  2048-slot tile;
* a species one slot short of a tile, one a slot past two tiles, and a
  species every particle of which leaves through the z = 0 boundary in the
  first step (the reference deletes Z <= 0, ParticleMaterial.F90:22-25; the
  attachment products of the same step are all that remains of it).

In each case the fused step equals the separate calls (move + absorption,
deposit, MCC) as particle sets, the captured graph equals the fused step bit
for bit, deposits agree to 1e-12 of their magnitude, the sort keeps every
particle, and a second step after the sort equals a second graph step taken
without one.
"""
import numpy as np
import pytest

from helpers import assert_close, sort_set
from ipm2d import HotPath
from ipm2d import workload as W

pytestmark = pytest.mark.gpu


def _hp(cfg, geom, tables, P):
    hp = HotPath(cfg.nz, cfg.nr, cfg.dx, cfg.dt, cfg.ns, 1)
    for s in range(cfg.ns):
        q, m = W.species_params(cfg, s)
        hp.set_species(s, q, m, 2 * P[s].shape[1] + 8192)
    hp.set_geometry(geom)
    for s in range(cfg.ns):
        hp.set_mcc(s, 0, tables.get((s, 0)))
        hp.upload(s, P[s])
    return hp


@pytest.mark.parametrize("counts, leave", [((2049, 1, 0), None), ((2047, 4097, 5000), 2)])
def test_edge_populations(rng, counts, leave):
    cfg = W.make_config("tiny", nz=65, nr=33, length=64e-3)
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    P = []
    for s, n in enumerate(counts):
        X = W.load_species(cfg, geom, s, rng)
        X = np.ascontiguousarray(X[:, rng.permutation(X.shape[1])[:n]])
        assert X.shape[1] == n
        if s == leave:
            X[0] = -0.25                     # every one of them leaves in the first step
        P.append(X)
    ez, er = W.fields(cfg, 1e-8)
    a, b, c = (_hp(cfg, geom, tables, P) for _ in range(3))
    for hp in (a, b, c):
        hp.set_fields(ez * 20, er * 20)
        hp.set_rng(7, 0)
    a.one_step()
    rho_a, _ = a.field_source_sum()
    b.move_absorb(-1)
    b.p2ces_all()
    rho_b, _ = b.field_source_sum()
    b.mcc()
    c.one_step_graph()
    rho_c, _ = c.field_source_sum()
    scale = max(np.abs(rho_a).max(), 1e-300)
    assert_close(rho_a, rho_b, rtol=1e-12, atol=1e-12 * scale, what="Rho fused vs separate")
    assert_close(rho_a, rho_c, rtol=1e-12, atol=1e-12 * scale, what="Rho fused vs graph")
    for s in range(cfg.ns):
        ga, gb, gc = a.download(s), b.download(s), c.download(s)
        assert ga.shape == gb.shape == gc.shape, f"species {s}"
        assert_close(sort_set(ga), sort_set(gb), rtol=1e-14, atol=1e-300, what=f"species {s} fused vs separate")
        np.testing.assert_array_equal(sort_set(ga), sort_set(gc), err_msg=f"species {s} fused vs graph")
    if leave is not None:
        # all of them left; what remains are products of this step's collisions
        left = a.download(leave)
        assert np.all(left[0] > 0.0) and left.shape[1] < counts[leave] // 10
    before = [a.download(s) for s in range(cfg.ns)]
    a.sort(-1)
    for s in range(cfg.ns):
        after = a.download(s)
        assert a.count(s) == before[s].shape[1]
        np.testing.assert_array_equal(sort_set(after), sort_set(before[s]), err_msg=f"species {s} sorted")
    a.one_step()
    c.one_step_graph()
    for s in range(cfg.ns):
        np.testing.assert_array_equal(sort_set(a.download(s)), sort_set(c.download(s)),
                                      err_msg=f"species {s} second step")
    for hp in (a, b, c):
        hp.close()
