"""The batched six-species configuration at its full size: 257x513 nodes,
e / CF3+ / CF2+ / CF+ / F- / CF3- in population ratios 1 : 0.8 : 0.1 : 0.05 :
0.6 : 0.1, 5.0e7 particles, null-collision MCC with the six synthetic tables
(elastic, ionization, dissociative attachment, charge exchange).

Three steps of the captured graph (one persistent launch over every species,
the split collision pass, the grid epilogue) against three steps of the
reference's call order with one launch per species and call (MoveES and
MaterialAborptionParticle per species, P2CES per species, FieldSource%Sum,
MCC, R/src/onestep.F90:141-177): every particle of every species identical
bit for bit (compared as multisets of 64-bit hashes of the 8 fields), the
populations equal, and Rho equal to 1e-12 of its magnitude (atomic summation
order).
"""
import numpy as np
import pytest

import ipm2d
from ipm2d import workload as W

pytestmark = pytest.mark.gpu

STEPS = 3


def _hashes(P):
    b = np.ascontiguousarray(P).view(np.uint64)
    h = np.full(b.shape[1], 0x9E3779B97F4A7C15, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for f in range(8):
            h ^= b[f] * np.uint64(0xBF58476D1CE4E5B9 + 2 * f)
            h = (h << np.uint64(27)) | (h >> np.uint64(37))
            h *= np.uint64(0x94D049BB133111EB)
    return np.sort(h)


def test_graph_steps_equal_per_species_calls():
    cfg = W.make_config("c3")
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    rng = np.random.default_rng(cfg.seed)
    parts = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    assert cfg.ns == 6 and sum(P.shape[1] for P in parts) > 4.5e7
    a, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.2)
    b, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.2)
    del parts
    for _ in range(STEPS):
        a.one_step_graph()
        for s in range(cfg.ns):
            b.move_es(s)
            b.material_absorption(s)
        for s in range(cfg.ns):
            b.p2ces(s, copy=False)
        b.field_source_sum(copy=False)
        b.mcc()
    rho_a, _ = a.field_source_sum()
    rho_b, _ = b.field_source_sum()
    np.testing.assert_allclose(rho_a, rho_b, rtol=1e-12, atol=1e-12 * np.abs(rho_b).max())
    for s in range(cfg.ns):
        assert a.count(s) == b.count(s), f"species {s}"
        ha, hb = _hashes(a.download(s)), _hashes(b.download(s))
        np.testing.assert_array_equal(ha, hb, err_msg=f"species {s}")
    a.close()
    b.close()
