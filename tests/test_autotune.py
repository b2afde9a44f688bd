"""Autotuner: selection logic with a fake context (CPU) and, on the GPU, that
tuning leaves the simulated physics unchanged."""
import json

import numpy as np
import pytest

from ipm2d import autotune


class FakeHP:
    nz, nr, ns = 65, 129, 3

    def __init__(self):
        self.bpc = 4
        self.cbpc = 6
        self.calls = []

    def set_launch(self, b):
        self.bpc = b
        self.calls.append(("launch", b))

    def set_tile_shape(self, block, rows):
        self.shape = (block, rows)
        self.calls.append(("shape", block, rows))

    def set_collide_grid(self, b):
        self.cbpc = b
        self.calls.append(("collide", b))

    def set_sort_tile(self, slots):
        self.sort_tile = slots
        self.calls.append(("sort_tile", slots))

    def sort(self, s=-1):
        self.calls.append(("sort", s))

    def one_step(self):
        pass

    def synchronize(self):
        pass


def test_picks_fastest_launch_shape():
    hp = FakeHP()
    cost = {2: 3.0, 4: 2.0, 6: 1.5, 8: 1.7}
    best, res = autotune.tune_blocks_per_cu(hp, candidates=list(cost), timer=lambda: cost[hp.bpc])
    assert best == 6 and hp.bpc == 6 and res == cost


def test_picks_fastest_collide_grid():
    hp = FakeHP()
    cost = {2: 1.2, 4: 1.3, 6: 1.4}
    best, res = autotune.tune_collide_grid(hp, candidates=list(cost), timer=lambda: cost[hp.cbpc])
    assert best == 2 and hp.cbpc == 2 and res == cost


def test_picks_fastest_tile_shape():
    hp = FakeHP()
    cost = {(256, 24): 1.3, (128, 24): 1.35, (512, 24): 1.25, (256, 16): 1.4, (256, 32): 1.28}
    best, res = autotune.tune_tile_shape(hp, candidates=list(cost), lead=2, timer=lambda: cost[hp.shape])
    assert best == (512, 24) and hp.shape == (512, 24) and res == cost
    # every candidate is timed after a sort and `lead` steps
    assert hp.calls.count(("sort", -1)) == len(cost)


def test_picks_fastest_sort_interval():
    hp = FakeHP()
    # sorting costs 10 units, an unsorted step costs 1 + 0.05 * steps since sort
    def timer(si):
        return (10 + sum(1 + 0.05 * k for k in range(si))) / si
    best, res = autotune.tune_sort_interval(hp, candidates=(5, 10, 20, 40), timer=timer)
    assert best == min(res, key=res.get) == 20


def test_picks_fastest_sort_tile():
    hp = FakeHP()
    cost = {2048: 2.3e-3, 4096: 2.4e-3}
    best, res = autotune.tune_sort_tile(hp, 60, candidates=list(cost), timer=lambda tile: cost[tile])
    assert best == 2048 and hp.sort_tile == 2048 and res == cost
    # each candidate is timed from a sorted state; the winner is applied last
    assert hp.calls[-1] == ("sort_tile", 2048)
    assert hp.calls.count(("sort", -1)) == len(cost)


def test_cache_round_trip(tmp_path, monkeypatch):
    hp = FakeHP()
    path = str(tmp_path / "tune.json")
    monkeypatch.setattr(autotune, "tune_blocks_per_cu", lambda h, c, s: (5, {5: 1.0}))
    monkeypatch.setattr(autotune, "tune_sort_interval", lambda h, c: (10, {10: 1.0}))
    monkeypatch.setattr(autotune, "tune_collide_grid", lambda h, c, s: (2, {2: 1.0}))
    monkeypatch.setattr(autotune, "tune_tile_shape", lambda h, c, s, l: ((128, 24), {(128, 24): 1.0}))
    monkeypatch.setattr(autotune, "tune_sort_tile", lambda h, i, c: (4096, {4096: 1.0}))
    out = autotune.autotune(hp, 2.46e6, cache_path=path)
    assert out["blocks_per_cu"] == 5 and out["sort_interval"] == 10 and out["collide_blocks_per_cu"] == 2
    assert out["tile_shape"] == [128, 24]
    monkeypatch.setattr(autotune, "tune_blocks_per_cu", lambda *a: pytest.fail("should use the cache"))
    hp2 = FakeHP()
    out2 = autotune.autotune(hp2, 2.46e6, cache_path=path)
    assert out2["sort_tile"] == 4096 and hp2.sort_tile == 4096
    assert out2["blocks_per_cu"] == 5 and hp2.bpc == 5 and hp2.cbpc == 2 and hp2.shape == (128, 24)
    assert autotune.cache_key(hp, 2.46e6) in json.load(open(path))


@pytest.mark.gpu
def test_tuning_does_not_change_results(rng):
    from helpers import sort_set
    from ipm2d import workload as W
    from test_gpu_parity import _config_hp, _energetic, _six_species_setup
    cfg, tables, _ = _six_species_setup(pressure=300.0)
    geom = W.make_geometry(cfg)
    P = [_energetic(cfg, s, 20000, rng, 10.0 if s == 0 else 1.0) for s in range(cfg.ns)]
    ez, er = W.fields(cfg, 3e-8)
    a = _config_hp(cfg, geom, tables, P)
    b = _config_hp(cfg, geom, tables, P)
    for hp in (a, b):
        hp.set_fields(ez, er)
        hp.set_rng(3, 0)
    out = autotune.autotune(a, 1.2e5, bpc_candidates=(2, 6), sort_candidates=(3, 5), steps=2,
                            collide_candidates=(1, 6), tile_lead=1)
    # a advanced 4 x (1 untimed + 2 timed) + 5 x (1 + 2) + 3 + 5 = 35 steps (plus sorts),
    # then one interval per sort tile size; b takes as many plain steps
    assert out["sort_tile"] in (2048, 4096)
    for _ in range(35 + 2 * out["sort_interval"]):
        b.one_step()
    for s in range(cfg.ns):
        ga, gb = sort_set(a.download(s)), sort_set(b.download(s))
        assert ga.shape == gb.shape
        np.testing.assert_allclose(ga, gb, rtol=1e-12, atol=1e-300)
    assert out["blocks_per_cu"] in (2, 6) and out["sort_interval"] in (3, 5)
    assert out["collide_blocks_per_cu"] in (1, 6)
    from ipm2d import IPMError
    for bad in (0, 7):
        with pytest.raises(IPMError):
            a.set_collide_grid(bad)


@pytest.mark.gpu
def test_tile_shapes_give_identical_steps(rng):
    """Every (threads per block, window rows) variant of the fused pass
    leaves every particle bit for bit as the default does, and Rho within
    1e-12 (atomic summation order), over two graph steps from one state."""
    from test_gpu_batched_c3 import _hashes
    from ipm2d import workload as W
    from test_gpu_parity import _config_hp, _energetic, _six_species_setup
    cfg, tables, _ = _six_species_setup(pressure=500.0)
    geom = W.make_geometry(cfg)
    P = [_energetic(cfg, s, 30000, rng, 15.0 if s == 0 else 1.0) for s in range(cfg.ns)]
    ez, er = W.fields(cfg, 3e-8)
    ref = None
    for shape in autotune.TILE_SHAPES:
        hp = _config_hp(cfg, geom, tables, P)
        hp.set_fields(ez * 20, er * 20)
        hp.set_rng(8, 0)
        hp.set_tile_shape(*shape)
        assert hp.tile_shape == tuple(shape)
        for _ in range(2):
            hp.one_step_graph()
        rho, _ = hp.field_source_sum()
        out = (rho, [_hashes(hp.download(s)) for s in range(cfg.ns)])
        hp.close()
        if ref is None:
            ref = out
            continue
        np.testing.assert_allclose(out[0], ref[0], rtol=1e-12, atol=1e-12 * np.abs(ref[0]).max())
        for s in range(cfg.ns):
            np.testing.assert_array_equal(out[1][s], ref[1][s], err_msg=f"shape {shape} species {s}")
