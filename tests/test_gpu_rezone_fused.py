"""Split/merge folded into the fused step (ipm_set_step_rezone) against the
separate calls in OneStep's order (R/src/onestep.F90:141-177):
MoveES + MaterialAborptionParticle, P2CES, particle_split_merge_control's
split_particle_weight + merge_particle_random (Rezoning.F90:53-92), MCC.

c1's electrons are loaded with weights at the reference weight (WQ/ref
close to 1, as the reference's own load gives them); 3 % of them get 2.2-3.8
times their weight (one or two copies each) and another 3 % 0.05-0.3 times
(merge candidates), and several steps run so that copies and merged
survivors move and collide.  Bars: particle multisets of every species bit for bit
(64-bit hashes of the 8 fields), Rho to 1e-12 of its magnitude (atomic
order), event counts per (species, reaction) identical; the graph form
equals the eager form; switching the fold off returns to the plain step."""
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


def _setup():
    cfg = W.make_config("c1")
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    rng = np.random.default_rng(cfg.seed)
    parts = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    alpha, grid_r = 0.01, float(cfg.nr - 1)
    shape = (1.0 - alpha) * parts[0][1] / grid_r + alpha
    weight = float(np.median(parts[0][7] / shape))      # WQ/ref close to 1
    n = parts[0].shape[1]
    pick = rng.permutation(n)
    up, down = pick[: n * 3 // 100], pick[n * 3 // 100: n * 6 // 100]
    parts[0][7, up] *= rng.uniform(2.2, 3.8, up.size)
    parts[0][7, down] *= rng.uniform(0.05, 0.3, down.size)
    return cfg, geom, tables, parts, dict(weight=weight, alpha=alpha, grid_r=grid_r, merge_limit=0.34)


def _hp(cfg, geom, tables, parts):
    hp, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=4.0)
    hp.set_rng(2024, 3)
    hp.set_event_counting(True)
    hp.take_mcc_counts()
    return hp


def test_fused_rezone_equals_separate_calls():
    cfg, geom, tables, parts, rz = _setup()
    a = _hp(cfg, geom, tables, parts)
    b = _hp(cfg, geom, tables, parts)
    a.set_step_rezone(0, **rz)
    n0 = b.count(0)
    for k in range(3):
        a.one_step()
        rho_a, _ = a.field_source_sum()
        b.move_absorb(-1)
        b.p2ces_all()
        rho_b, _ = b.field_source_sum()
        st = b.split_merge(0, rz["weight"], rz["alpha"], rz["grid_r"], rz["merge_limit"])
        b.mcc()
        if k == 0:
            print(f"step 0: {n0} electrons, split copies {st[0]}, merge kept {st[1]}, deleted {st[2]}", flush=True)
            assert st[0] > 1000 and st[1] > 100 and st[2] > 100
        np.testing.assert_allclose(rho_a, rho_b, rtol=1e-12, atol=1e-12 * np.abs(rho_b).max())
        for s in range(cfg.ns):
            assert a.count(s) == b.count(s), (k, s)
            np.testing.assert_array_equal(_hashes(a.download(s)), _hashes(b.download(s)),
                                          err_msg=f"step {k} species {s}")
    np.testing.assert_array_equal(a.take_mcc_counts(), b.take_mcc_counts())
    a.close()
    b.close()


def test_fused_rezone_graph_equals_eager_and_switches_off():
    cfg, geom, tables, parts, rz = _setup()
    a = _hp(cfg, geom, tables, parts)
    b = _hp(cfg, geom, tables, parts)
    for hp in (a, b):
        hp.set_step_rezone(0, **rz)
    for _ in range(3):
        a.one_step_graph()
        b.one_step()
    # a new Weight takes effect without recapture; mode 0 returns to the plain step
    for hp in (a, b):
        hp.set_step_rezone(0, **dict(rz, weight=rz["weight"] * 1.01))
    a.one_step_graph()
    b.one_step()
    for hp in (a, b):
        hp.set_step_rezone(0, split=False, merge=False)
    a.one_step_graph()
    b.one_step()
    a.synchronize()
    b.synchronize()
    for s in range(cfg.ns):
        assert a.count(s) == b.count(s)
        np.testing.assert_array_equal(_hashes(a.download(s)), _hashes(b.download(s)), err_msg=f"species {s}")
    a.close()
    b.close()
