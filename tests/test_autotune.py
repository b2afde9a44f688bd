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

    def set_collide_grid(self, b):
        self.cbpc = b
        self.calls.append(("collide", b))

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


def test_picks_fastest_sort_interval():
    hp = FakeHP()
    # sorting costs 10 units, an unsorted step costs 1 + 0.05 * steps since sort
    def timer(si):
        return (10 + sum(1 + 0.05 * k for k in range(si))) / si
    best, res = autotune.tune_sort_interval(hp, candidates=(5, 10, 20, 40), timer=timer)
    assert best == min(res, key=res.get) == 20


def test_cache_round_trip(tmp_path, monkeypatch):
    hp = FakeHP()
    path = str(tmp_path / "tune.json")
    monkeypatch.setattr(autotune, "tune_blocks_per_cu", lambda h, c, s: (5, {5: 1.0}))
    monkeypatch.setattr(autotune, "tune_sort_interval", lambda h, c: (10, {10: 1.0}))
    monkeypatch.setattr(autotune, "tune_collide_grid", lambda h, c, s: (2, {2: 1.0}))
    out = autotune.autotune(hp, 2.46e6, cache_path=path)
    assert out["blocks_per_cu"] == 5 and out["sort_interval"] == 10 and out["collide_blocks_per_cu"] == 2
    monkeypatch.setattr(autotune, "tune_blocks_per_cu", lambda *a: pytest.fail("should use the cache"))
    hp2 = FakeHP()
    out2 = autotune.autotune(hp2, 2.46e6, cache_path=path)
    assert out2["blocks_per_cu"] == 5 and hp2.bpc == 5 and hp2.cbpc == 2
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
                            collide_candidates=(1, 6))
    # a advanced 4 x (1 untimed + 2 timed) + 3 + 5 = 20 steps (plus sorts); b takes 20 plain steps
    for _ in range(20):
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
