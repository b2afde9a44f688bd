"""Restart through the device: PB%Dump of resident particles in both of the
reference's formats (ParticleBundle.f90:117-174) and PB%Load into a fresh
context (lines 178-262) reproduce the particles bit for bit, and an unscaled
dump (SI positions and velocities) loads in grid units."""
import numpy as np
import pytest

from helpers import QE, ME, random_particles, small_geometry, sort_set
from ipm2d import HotPath, restart

pytestmark = pytest.mark.gpu

DT = 5e-11


def _hp(geom):
    hp = HotPath(geom.nz, geom.nr, geom.dx, DT, 1, 1)
    hp.set_species(0, -QE, ME, 50000)
    hp.set_geometry(geom)
    return hp


@pytest.mark.parametrize("ext", ["dat", "h5"])
def test_dump_and_load_round_trip(tmp_path, rng, ext):
    geom = small_geometry()
    P = random_particles(geom.nz, geom.nr, 3000, rng)
    a = _hp(geom)
    a.upload(0, P)
    d = a.dump_restart(0, str(tmp_path / f"e.{ext}"), weight=1.25e5)
    assert d.npar == 3000 and d.xfactor == geom.dx
    b = _hp(geom)
    e = b.load_restart(0, str(tmp_path / f"e.{ext}"))
    assert (e.charge, e.mass, e.weight) == (-QE, ME, 1.25e5)
    assert b.count(0) == 3000
    np.testing.assert_array_equal(sort_set(b.download(0)), sort_set(a.download(0)))
    a.close()
    b.close()


def test_unscaled_dump_loads_in_grid_units(tmp_path, rng):
    geom = small_geometry()
    P = random_particles(geom.nz, geom.nr, 500, rng)
    vf = geom.dx / DT
    Q = P.copy()
    Q[0:2] = Q[0:2] * geom.dx
    Q[2:5] = Q[2:5] * vf
    d = restart.BundleDump(500, 500, -QE, ME, 1.0, geom.dx, vf, False, False, Q)
    restart.write_dat(tmp_path / "si.dat", d)
    b = _hp(geom)
    b.load_restart(0, str(tmp_path / "si.dat"))
    got = sort_set(b.download(0))
    want = Q.copy()
    want[0:2] = want[0:2] * (1.0 / geom.dx)
    want[2:5] = want[2:5] * (1.0 / vf)
    np.testing.assert_array_equal(got, sort_set(want))
    np.testing.assert_allclose(got, sort_set(P), rtol=1e-15, atol=0)
    b.close()
