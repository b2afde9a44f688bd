"""ipm_moments_noalloc (the moment diagnostics with their sums in the
library's static device scratch instead of an allocation) against
ipm_moments, through the C ABI: two calls in a row, the second on a smaller
grid with fewer particles (the scratch keeps the first call's contents and
must be cleared), then the first again; every output within 1e-12 of
ipm_moments (FP64 atomic summation order), skipped outputs untouched, and a
grid above the scratch's 2^20 nodes refused."""
import numpy as np
import pytest

from helpers import AMU, ME, QE, assert_close, random_fields, random_particles, small_geometry
from ipm2d import HotPath, IPMError

pytestmark = pytest.mark.gpu

DT = 5e-11
SPECIES = [(-QE, ME), (QE, 69.0059096 * AMU)]
KEYS = ("rho", "energy", "mp", "jz", "jr")


def _context(nz, nr, counts, rng):
    geom = small_geometry(nz=nz, nr=nr, dielectric=False)
    hp = HotPath(nz, nr, geom.dx, DT, 2, 1)
    for s in range(2):
        hp.set_species(s, *SPECIES[s], 60000)
    hp.set_geometry(geom)
    ez, er = random_fields(nz, nr, rng, scale=2e3)
    hp.set_fields(ez, er)
    for s in range(2):
        hp.upload(s, random_particles(nz, nr, counts[s], rng, vscale=0.1))
    return hp


def _noalloc(hp, s, implicit_j, skip=()):
    import torch
    n = hp.nz * hp.nr
    bufs = {k: torch.full((n,), -7.0, dtype=torch.float64, device="cuda") for k in KEYS}
    torch.cuda.synchronize()
    hp.moments_noalloc(s, implicit_j, [None if k in skip else bufs[k].data_ptr() for k in KEYS])
    hp.synchronize()
    return {k: v.cpu().numpy().reshape(hp.nr, hp.nz).T for k, v in bufs.items()}


def _check(hp, s, implicit_j, skip=()):
    got = _noalloc(hp, s, implicit_j, skip)
    ref = hp.moments(s, implicit_j=implicit_j)
    for k in KEYS:
        if k in skip:
            assert np.all(got[k] == -7.0), f"skipped output {k} was written"
            continue
        assert_close(got[k], ref[k], rtol=1e-12, atol=1e-12 * np.abs(ref[k]).max() + 1e-300, what=f"{k} species {s}")
    assert np.abs(ref["rho"]).max() > 0


def test_noalloc_moments_match_ipm_moments():
    pytest.importorskip("torch")
    rng = np.random.default_rng(31)
    big = _context(33, 17, (30000, 8000), rng)
    small = _context(17, 9, (4000, 1000), rng)
    _check(big, 0, True)
    _check(small, 1, False)                 # smaller grid and input: stale sums must be cleared
    _check(small, 0, True, skip=("energy", "jr"))
    _check(big, 1, True)
    big.close()
    small.close()


def test_noalloc_refuses_grids_above_the_scratch():
    pytest.importorskip("torch")
    hp = HotPath(1025, 1025, 1e-3, DT, 1, 1)    # 1,050,625 nodes > 2^20
    hp.set_species(0, -QE, ME, 10)
    geom = small_geometry(nz=1025, nr=1025, dielectric=False)
    hp.set_geometry(geom)
    with pytest.raises(IPMError, match="scratch"):
        hp.moments_noalloc(0, False, [None] * 5)
    hp.close()
