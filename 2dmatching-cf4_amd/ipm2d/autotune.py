"""Online autotuner for the batched step: launch shape and cell-sort interval.

The tuned knobs change only how the work is scheduled, never what is
computed (the GPU parity tests show the step is invariant to the launch
shape, to sorting and to storage order), so the tuner measures them on the
live simulation: every candidate advances the real state by a few steps.

* ``blocks_per_cu``: 0 for one block per tile, else persistent blocks per
  compute unit of the fused particle kernel; chosen by timing a few steps each.
* ``tile_shape``: threads per block of the fused pass (128, 256 or 512, i.e.
  16, 8 or 4 particles per thread of a 2048-slot tile) and the height of its
  LDS deposit window (16, 24 or 32 node rows).
* ``collide_blocks_per_cu``: grid of the collision pass (blocks per compute
  unit); chosen by timing a few whole steps each, since it also changes
  where the pass leaves the caches for the next particle pass.
* ``sort_interval``: steps between cell sorts.  A sort costs about one
  particle pass but keeps the deposit's LDS window and wave reduction
  effective; the interval minimising the mean step time over a full
  interval (sort included) is kept.
* ``sort_tile``: slots per tile of the cell sort (2048 or 4096), chosen by
  timing one sort after a full interval of drift with each.

Results can be cached in a JSON file keyed by grid, species count and
particle count (rounded), so a production run tunes once.
"""
import json
import os
import time


def _timed(hp, steps, sort_interval=0, offset=0):
    hp.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        hp.one_step()
        if sort_interval and (offset + k + 1) % sort_interval == 0:
            hp.sort(-1)
    hp.synchronize()
    return (time.perf_counter() - t0) / steps


def tune_blocks_per_cu(hp, candidates=(0, 2, 4, 6, 8), steps=6, timer=None):
    """Return (best, {candidate: seconds per step})."""
    timer = timer or (lambda: _timed(hp, steps))
    res = {}
    hp.sort(-1)
    for b in candidates:
        hp.set_launch(b)
        hp.one_step()                     # first launch at a new shape is not timed
        res[b] = timer()
    best = min(res, key=res.get)
    hp.set_launch(best)
    return best, res


# (threads per block, deposit-window rows) of the fused particle pass
TILE_SHAPES = ((256, 24), (128, 24), (512, 24), (256, 16), (256, 32))


def tune_tile_shape(hp, candidates=TILE_SHAPES, steps=6, lead=10, timer=None):
    """Threads per block (16, 8 or 4 particles per thread of a 2048-slot
    tile) and deposit-window height of the fused pass.  Each candidate is
    timed over `steps` steps that start `lead` steps after a sort: the
    window's height matters once particles have drifted out of cell order.
    Returns (best, {(block, rows): seconds per step})."""
    timer = timer or (lambda: _timed(hp, steps))
    res = {}
    for shape in candidates:
        hp.set_tile_shape(*shape)
        hp.sort(-1)
        for _ in range(lead):
            hp.one_step()
        res[tuple(shape)] = timer()
    best = min(res, key=res.get)
    hp.set_tile_shape(*best)
    return best, res


def tune_collide_grid(hp, candidates=(2, 4, 6), steps=6, timer=None):
    """Return (best, {candidate: seconds per step})."""
    timer = timer or (lambda: _timed(hp, steps))
    res = {}
    hp.sort(-1)
    for b in candidates:
        hp.set_collide_grid(b)
        hp.one_step()
        res[b] = timer()
    best = min(res, key=res.get)
    hp.set_collide_grid(best)
    return best, res


def tune_sort_interval(hp, candidates=(5, 10, 20, 40), timer=None):
    """Mean step time over one full interval (its sort included) per candidate;
    returns (best, {interval: seconds per step})."""
    res = {}
    for si in candidates:
        hp.sort(-1)
        if timer is not None:
            res[si] = timer(si)
        else:
            res[si] = _timed(hp, si, sort_interval=si)
    best = min(res, key=res.get)
    return best, res


def tune_sort_tile(hp, interval, candidates=(2048, 4096), timer=None):
    """Time of one cell sort after `interval` steps of drift from a sorted
    state, per sort tile size in slots; returns (best, {slots: seconds})."""
    res = {}
    for tile in candidates:
        hp.set_sort_tile(tile)
        hp.sort(-1)
        if timer is not None:
            res[tile] = timer(tile)
        else:
            for _ in range(interval):
                hp.one_step()
            hp.synchronize()
            t0 = time.perf_counter()
            hp.sort(-1)
            hp.synchronize()
            res[tile] = time.perf_counter() - t0
    best = min(res, key=res.get)
    hp.set_sort_tile(best)
    return best, res


def cache_key(hp, nparticles):
    return f"{hp.nz}x{hp.nr}_ns{hp.ns}_n{int(round(nparticles, -5))}"


def autotune(hp, nparticles, cache_path=None, bpc_candidates=(0, 2, 4, 6, 8),
             sort_candidates=(5, 10, 20, 40, 60, 80), steps=6, collide_candidates=(2, 4, 6),
             tile_candidates=TILE_SHAPES, tile_lead=10, sort_tile_candidates=(2048, 4096)):
    """Tune (or load from cache) and apply; returns the chosen settings."""
    key = cache_key(hp, nparticles)
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            hp.set_launch(cache[key]["blocks_per_cu"])
            hp.set_tile_shape(*cache[key].get("tile_shape", (256, 24)))
            hp.set_collide_grid(cache[key].get("collide_blocks_per_cu", 6))
            hp.set_sort_tile(cache[key].get("sort_tile", 2048))
            return cache[key]
    bpc, tb = tune_blocks_per_cu(hp, bpc_candidates, steps)
    shape, tt = tune_tile_shape(hp, tile_candidates, steps, tile_lead) if tile_candidates else ((256, 24), {})
    cbpc, tc = tune_collide_grid(hp, collide_candidates, steps) if collide_candidates else (6, {})
    si, ts = tune_sort_interval(hp, sort_candidates)
    stile, tst = tune_sort_tile(hp, si, sort_tile_candidates) if (sort_tile_candidates and si) else (2048, {})
    out = {"blocks_per_cu": bpc, "tile_shape": list(shape), "collide_blocks_per_cu": cbpc, "sort_interval": si,
           "sort_tile": stile, "sort_seconds_by_sort_tile": {str(k): v for k, v in tst.items()},
           "step_seconds_by_tile_shape": {f"{k[0]}x{k[1]}": v for k, v in tt.items()},
           "step_seconds_by_blocks_per_cu": {str(k): v for k, v in tb.items()},
           "step_seconds_by_collide_blocks_per_cu": {str(k): v for k, v in tc.items()},
           "step_seconds_by_sort_interval": {str(k): v for k, v in ts.items()}}
    if cache_path:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        cache[key] = out
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, cache_path)
    return out
