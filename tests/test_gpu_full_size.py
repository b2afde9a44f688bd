"""The headline workload (c2) and the largest one (c5) at their full sizes.

c2: 257x513 nodes, 49.7 M particles of e / CF3+ / F- (cos^2 profile, CF4
at 50 mTorr).
* Three captured-graph steps against three steps of the reference's call
  order with one launch per species and call (MoveES and
  MaterialAborptionParticle per species, P2CES per species, FieldSource%Sum,
  MCC, R/src/onestep.F90:141-177): every particle identical bit for bit
  (multisets of 64-bit hashes of the 8 fields), Rho to 1e-12 of its
  magnitude (atomic summation order).
* One fused step against the oracle: push and absorption of every particle
  (the survivors as a set, exactly), Rho of all of them (1e-12), and the
  MCC of every 40th survivor -- positions and weights exactly, velocities
  to 1e-10 of |v|, its products found among the device's.  The particles
  are independent in MCC (each draws from its own state-keyed stream), so a
  sample checks the same events the full run makes.

c5: 513x1025 nodes, 1.0e9 particles of six species (100 mTorr), made on the
GPU (workload.load_species_device) since a host copy of 64 GB would take
minutes to build.
* Charge conservation of the deposit per species after push and
  absorption: sum over nodes of RhoOne / point_inv_volume = Charge * sum WQ
  of the live particles (bilinear weights sum to one), to 1e-11; the nodes
  inside the electrodes (no weighting volume) receive nothing.
* Two graph steps against two steps of per-species calls from the same
  initial state (made twice from one seed): equal populations and equal
  order-free digests of every particle's 8 fields, bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import ipm2d
from helpers import sort_set
from ipm2d import geometry as geo
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


def _per_species_step(hp, ns):
    for s in range(ns):
        hp.move_es(s)
        hp.material_absorption(s)
    for s in range(ns):
        hp.p2ces(s, copy=False)
    hp.field_source_sum(copy=False)
    hp.mcc()


# ---------------------------------------------------------------- c2

@pytest.fixture(scope="module")
def c2():
    cfg = W.make_config("c2")
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    rng = np.random.default_rng(cfg.seed)
    parts = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    assert sum(P.shape[1] for P in parts) > 4.9e7
    return cfg, geom, tables, parts


def test_c2_graph_steps_equal_per_species_calls(c2):
    cfg, geom, tables, parts = c2
    ez, er = W.fields(cfg, 1e-8)
    a, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.1)
    b, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.1)
    for hp in (a, b):
        hp.set_fields(ez, er)
    for _ in range(3):
        a.one_step_graph()
        _per_species_step(b, cfg.ns)
    rho_a, _ = a.field_source_sum()
    rho_b, _ = b.field_source_sum()
    np.testing.assert_allclose(rho_a, rho_b, rtol=1e-12, atol=1e-12 * np.abs(rho_b).max())
    for s in range(cfg.ns):
        assert a.count(s) == b.count(s), f"species {s}"
        np.testing.assert_array_equal(_hashes(a.download(s)), _hashes(b.download(s)), err_msg=f"species {s}")
    a.close()
    b.close()


def _alive_after_absorption(geom, X):
    """Slots MaterialAborptionParticle keeps (ParticleMaterial.F90:22-41) and
    the kept particles with the axis reflection applied, in slot order."""
    nz, nr = geom.nz, geom.nr
    Z, R = X[0], X[1]
    dead = (Z <= 0.0) | (Z >= nz - 1) | (R >= nr - 1)
    inside = ~dead & (R > 0.0)
    iz = np.clip(np.ceil(Z[inside]).astype(np.int64), 1, nz - 1) - 1
    ir = np.clip(np.ceil(R[inside]).astype(np.int64), 1, nr - 1) - 1
    kind = geom.material_type[geom.cell_material[iz, ir] + geom.ndiel]
    d2 = np.zeros_like(dead)
    d2[np.nonzero(inside)[0]] = (kind == geo.MAT_METAL) | (kind == geo.MAT_DIEL)
    alive = ~(dead | d2)
    Y = X[:, alive].copy()
    refl = Y[1] <= 0.0
    Y[1, refl] = np.abs(Y[1, refl])
    Y[3, refl] = np.abs(Y[3, refl])
    Y[6, refl] = np.abs(Y[6, refl])
    return alive, Y


def test_c2_step_matches_oracle(c2, oracle):
    cfg, geom, tables, parts = c2
    nz, nr = cfg.nz, cfg.nr
    ez, er = W.fields(cfg, 1e-8)
    hp, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=parts, capacity_factor=1.1)
    hp.set_fields(ez, er)
    hp.set_rng(cfg.seed, 0)
    hp.one_step()
    rho, _ = hp.field_source_sum()
    species = [W.species_params(cfg, s) for s in range(cfg.ns)]
    setup = oracle.MccSetup(cfg.dx, cfg.dt, [q for q, _ in species], [m for _, m in species], tables)
    rho_ref = np.zeros((nz, nr))
    survivors, samples = [], []
    lost = 0
    for s, (q, m) in enumerate(species):
        X = oracle.push(nz, nr, cfg.dx, cfg.dt, q, m, ez, er, parts[s], fast=True)
        alive, Y = _alive_after_absorption(geom, X)
        Xa, _, _ = oracle.absorb(geom, q, X, fast=True)
        np.testing.assert_array_equal(_hashes(Y), _hashes(Xa), err_msg=f"species {s} survivors")
        lost += int((~alive).sum())
        r_s, _ = oracle.p2ces(nz, nr, cfg.dt, q, m, Y, geom.inv_volume)
        rho_ref = rho_ref + r_s
        survivors.append(Y)
        samples.append(np.arange(0, Y.shape[1], 40))
    assert lost > 100, "no particle was absorbed: the absorption check would be empty"
    np.testing.assert_allclose(rho, rho_ref, rtol=1e-12, atol=1e-12 * np.abs(rho_ref).max())
    ref, cnt = oracle.mcc(setup, cfg.seed, 0, [Y[:, k] for Y, k in zip(survivors, samples)], counts=True)
    assert cnt[0, 0, 1:].sum() > 1000
    for s in range(cfg.ns):
        G = hp.download(s)
        na, k = survivors[s].shape[1], samples[s]
        assert G.shape[1] >= na
        g, r = G[:, :na][:, k], ref[s][:, :k.size]
        np.testing.assert_array_equal(g[[0, 1, 7]], r[[0, 1, 7]], err_msg=f"species {s} positions/weights")
        np.testing.assert_allclose(g[5:7], r[5:7], rtol=1e-12, atol=1e-300)
        err = np.linalg.norm(g[2:5] - r[2:5], axis=0)
        assert np.all(err <= 1e-10 * np.linalg.norm(r[2:5], axis=0) + 1e-300), f"species {s} velocities"
        # the sample's products are among the device's (keyed by place and weight)
        prods = G[:, na:]
        index = {}
        for j in range(prods.shape[1]):
            index.setdefault(prods[[0, 1, 7], j].tobytes(), []).append(prods[2:5, j])
        mine = ref[s][:, k.size:]
        for j in range(mine.shape[1]):
            cands = index.get(mine[[0, 1, 7], j].tobytes(), [])
            v = mine[2:5, j]
            assert any(np.linalg.norm(c - v) <= 1e-10 * np.linalg.norm(v) for c in cands), \
                f"species {s}: product {j} of the sample not produced on the device"
    hp.close()


# ---------------------------------------------------------------- c5

def _hip():
    return C.CDLL("libamdhip64.so.7")     # the runtime the process has loaded


def _digest(hp, s, n):
    """Order-free digest of the first n slots of species s (a compacted
    species): the int64 sum of a mixing hash of each particle's 8 fields."""
    import torch
    ptrs = hp.device_ptrs(s)
    h = torch.full((n,), 0x1F83D9AB, dtype=torch.int64, device="cuda")
    t = torch.empty(n, dtype=torch.float64, device="cuda")
    hip = _hip()
    for f in range(8):
        torch.cuda.synchronize()
        assert hip.hipMemcpy(C.c_void_p(t.data_ptr()), C.c_void_p(ptrs[f]), C.c_size_t(8 * n), 3) == 0
        b = t.view(torch.int64)
        h = h ^ (b * (0x2545F4914F6CDD1D + 2 * f))
        h = h ^ (h >> 29)
        h = h * 0x5851F42D4C957F2D
    return int(h.sum().item())


def _live_wq_sum(hp, s, cap):
    """Sum of WQ over the live slots among the first cap slots (dead slots
    hold Z = -1e300; slots past the count are not read: cap is the count of
    slots in use)."""
    import torch
    ptrs = hp.device_ptrs(s)
    z = torch.empty(cap, dtype=torch.float64, device="cuda")
    w = torch.empty(cap, dtype=torch.float64, device="cuda")
    hip = _hip()
    torch.cuda.synchronize()
    assert hip.hipMemcpy(C.c_void_p(z.data_ptr()), C.c_void_p(ptrs[0]), C.c_size_t(8 * cap), 3) == 0
    assert hip.hipMemcpy(C.c_void_p(w.data_ptr()), C.c_void_p(ptrs[7]), C.c_size_t(8 * cap), 3) == 0
    return float(torch.where(z > -1e299, w, torch.zeros_like(w)).sum().item()), int((z > -1e299).sum().item())


def test_c5_full_size():
    torch = pytest.importorskip("torch")
    cfg = W.make_config("c5")
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    inner = geom.inv_volume > 0.0            # nodes inside the electrodes have no weighting volume
    digests = []
    for path in ("graph", "per-species"):
        hp, _, _, n0 = ipm2d.from_config_device(cfg, geom=geom, tables=tables, capacity_factor=1.04)
        assert sum(n0) > 0.99e9, f"only {sum(n0)} particles"
        ez, er = W.fields(cfg, 1e-8)
        hp.set_fields(ez, er)
        if path == "graph":
            # deposit after push + absorption conserves each species' charge
            hp.move_absorb(-1)
            rho_one, _ = hp.p2ces_all()
            for s in range(cfg.ns):
                q, _ = W.species_params(cfg, s)
                wq, live = _live_wq_sum(hp, s, n0[s])
                assert live == hp.count(s) and live > 0.99 * n0[s]
                assert not rho_one[s][~inner].any()
                got = float(np.sum(rho_one[s][inner] / geom.inv_volume[inner]))
                assert abs(got - q * wq) <= 1e-11 * abs(q * wq), f"species {s}: {got} vs {q * wq}"
            hp.close()
            hp, _, _, n1 = ipm2d.from_config_device(cfg, geom=geom, tables=tables, capacity_factor=1.04)
            assert n1 == n0
            hp.set_fields(ez, er)
            for _ in range(2):
                hp.one_step_graph()
            hp.synchronize()
        else:
            for _ in range(2):
                _per_species_step(hp, cfg.ns)
        counts = [hp.count(s) for s in range(cfg.ns)]
        assert all(c > 0.98 * n for c, n in zip(counts, n0))
        assert counts[0] > n0[0] * 0.999                  # ionization keeps the electrons
        hp.sort(-1)                                       # compacts: the first count slots are the particles
        digests.append((counts, [_digest(hp, s, counts[s]) for s in range(cfg.ns)]))
        hp.close()
        torch.cuda.empty_cache()
    assert digests[0] == digests[1]
