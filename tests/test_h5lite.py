"""HDF5 reading and bundle-dump writing without the HDF5 library (ipm2d.h5lite).

* The reference's own geometry input (input/geom.h5, written by h5py from
  R/script/geom/geometry.py:721-738) decodes to arrays that agree with each
  other the way the script built them: the node volume fractions equal the
  script's rule (update_volume, geometry.py:477-555) applied to the decoded
  cell materials, for exactly one choice of which materials are vacuum, bit
  for bit on all 201 x 201 nodes; every cell of a material has the same
  permittivity; shapes match the Nz, Nr attributes.  A wrong offset, type or
  transposition breaks the agreement.  (Skipped where the reference tree is
  absent; the file is not copied into this repository.)
* A bundle dump written by write_bundle reads back bit for bit, attributes
  included, with NPar below size.
* A geometry written by write_geometry reads back bit for bit, dtypes and
  the reversed (Fortran-reader) dimension order included; the reference's
  own geometry input survives a write and a read unchanged.
* Compatibility of write_bundle's and write_geometry's output with the HDF5
  library itself is parity unpinned here: no HDF5 library is installed to
  read it.
"""
import itertools
import os
import struct

import numpy as np
import pytest

from ipm2d import h5lite, restart

GEOM = "/root/reference/iPM2D-multi-EI/input/geom.h5"


def _volume_rule(cm, vacuum):
    """update_volume of the geometry script, R/script/geom/geometry.py:477-555."""
    nz, nr = cm.shape[0] + 1, cm.shape[1] + 1
    vac = np.isin(cm, list(vacuum))
    v = np.zeros((nz, nr))
    v[0, 0] = 0.5 if vac[0, 0] else 0
    v[nz - 1, 0] = 0.5 if vac[nz - 2, 0] else 0
    v[0, nr - 1] = 0.5 if vac[0, nr - 2] else 0
    v[nz - 1, nr - 1] = 0.5 if vac[nz - 2, nr - 2] else 0
    for z in range(1, nz - 1):
        v[z, 0] = (0.5 if vac[z - 1, 0] else 0) + (0.5 if vac[z, 0] else 0)
        v[z, nr - 1] = (0.5 if vac[z - 1, nr - 2] else 0) + (0.5 if vac[z, nr - 2] else 0)
    for r in range(1, nr - 1):
        lo, hi = 0.25 * (1 - 1 / 4 / r), 0.25 * (1 + 1 / 4 / r)
        v[0, r] = (lo if vac[0, r - 1] else 0) + (hi if vac[0, r] else 0)
        v[nz - 1, r] = (lo if vac[nz - 2, r - 1] else 0) + (hi if vac[nz - 2, r] else 0)
        for z in range(1, nz - 1):
            s = 0
            s = s + (lo if vac[z - 1, r - 1] else 0)
            s = s + (hi if vac[z - 1, r] else 0)
            s = s + (lo if vac[z, r - 1] else 0)
            s = s + (hi if vac[z, r] else 0)
            v[z, r] = s
    return v


@pytest.mark.skipif(not os.path.exists(GEOM), reason="reference tree absent")
def test_reference_geometry_decodes_consistently():
    g = h5lite.read_geometry(GEOM)
    nz, nr = g["Nz"], g["Nr"]
    cm = g["cell_material"]
    assert cm.shape == (nz - 1, nr - 1) and g["epsilon_r"].shape == (nz - 1, nr - 1)
    assert g["point_type"].shape == (nz, nr) and g["volume"].shape == (nz, nr)
    assert g["hedge_type"].shape == (nz, nr - 1) and g["vedge_type"].shape == (nz - 1, nr)
    assert set(np.unique(g["hedge_type"])) <= {0, 1, 2, 3, 4}
    assert set(np.unique(g["vedge_type"])) <= {0, 1, 2, 3, 4}
    # one permittivity per material
    for m in np.unique(cm):
        assert np.unique(g["epsilon_r"][cm == m]).size == 1
    # the volume fractions follow from the materials for exactly one vacuum set
    mats = sorted(int(m) for m in np.unique(cm))
    hits = []
    for k in range(1, len(mats) + 1):
        for vac in itertools.combinations(mats, k):
            if np.array_equal(_volume_rule(cm, vac), g["volume"]):
                hits.append(vac)
    assert len(hits) == 1, hits
    # a vacuum material has permittivity 1
    assert all(np.all(g["epsilon_r"][cm == m] == 1.0) for m in hits[0])


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_bundle_round_trip(tmp_path, rng):
    P = rng.standard_normal((8, 37)) * np.logspace(-300, 300, 37)
    d = restart.BundleDump(29, 37, -1.6022e-19, 9.1095e-31, 7.5e11, 7.8125e-4, 1.5625e7, True, False, P)
    p = tmp_path / "b.h5"
    h5lite.write_bundle(p, d)
    raw = p.read_bytes()
    assert raw[:8] == b"\x89HDF\r\n\x1a\n" and raw[8] == 0
    f = h5lite.File(p)
    assert sorted(f.keys()) == sorted(restart.FIELDS)
    assert f.attrs["NPar"] == 29 and f.attrs["size"] == 37
    assert f.attrs["LXScaled"] == 1 and f.attrs["LVScaled"] == 0
    e = h5lite.read_bundle(p)
    assert (e.npar, e.size, e.lxscaled, e.lvscaled) == (29, 37, True, False)
    for a, b in ((e.charge, d.charge), (e.mass, d.mass), (e.weight, d.weight), (e.xfactor, d.xfactor),
                 (e.vfactor, d.vfactor)):
        assert struct.pack("<d", a) == struct.pack("<d", b)
    np.testing.assert_array_equal(_bits(e.po), _bits(P))
    np.testing.assert_array_equal(f.read("Vt"), P[4])


def test_not_hdf5_rejected(tmp_path):
    p = tmp_path / "x.h5"
    p.write_bytes(b"not an hdf5 file at all" * 100)
    with pytest.raises(ValueError):
        h5lite.File(p)


def test_geometry_round_trip(tmp_path, rng):
    nz, nr = 9, 6
    g = {"Nz": nz, "Nr": nr, "cell_material": rng.integers(0, 5, (nz - 1, nr - 1)).astype(np.int32),
         "epsilon_r": rng.random((nz - 1, nr - 1)), "load": rng.integers(0, 2, (nz - 1, nr - 1)).astype(np.int32),
         "point_type": rng.integers(-2, 2, (nz, nr)).astype(np.int32), "volume": rng.random((nz, nr)),
         "hedge_type": rng.integers(0, 5, (nz, nr - 1)).astype(np.int32),
         "vedge_type": rng.integers(0, 5, (nz - 1, nr)).astype(np.int32)}
    h5lite.write_geometry(tmp_path / "g.h5", g)
    e = h5lite.read_geometry(tmp_path / "g.h5")
    assert (e["Nz"], e["Nr"]) == (nz, nr)
    for k in g:
        if k not in ("Nz", "Nr"):
            np.testing.assert_array_equal(e[k], g[k])
            assert e[k].dtype == g[k].dtype
    # stored with the dimensions reversed, as a Fortran reader expects
    assert h5lite.File(tmp_path / "g.h5").read("hedge_type").shape == (nr - 1, nz)
    if os.path.exists(GEOM):
        r = h5lite.read_geometry(GEOM)
        h5lite.write_geometry(tmp_path / "ref.h5", r)
        back = h5lite.read_geometry(tmp_path / "ref.h5")
        for k in r:
            np.testing.assert_array_equal(np.asarray(back[k]), np.asarray(r[k]))
