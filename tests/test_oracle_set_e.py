"""SetE (R/src/pic/field/FieldSource.f90:95-120) after linear ghost
extrapolation (R/src/pic/domain/PICCom.F90:516-578), CPU checks:

* a linear potential phi = a z + b r + c gives Ez = -a and Er = -b on every
  face, including the ghost-fed first and last faces (a wrong ghost rule,
  sign or stagger fails);
* a quadratic in z gives the centred difference -(phi(z_i+dx) - phi(z_i))/dx
  on interior faces exactly;
* the package's numpy restatement (geometry.set_e_from_phi) agrees bit for bit.
"""
import numpy as np

from ipm2d.geometry import set_e_from_phi


def _grid(nz=17, nr=9, dx=1e-3):
    z = np.arange(nz)[:, None] * dx
    r = np.arange(nr)[None, :] * dx
    return nz, nr, dx, z, r


def test_linear_potential_gives_uniform_field(oracle):
    nz, nr, dx, z, r = _grid()
    a, b = 1234.5, -678.25
    phi = a * z + b * r + 3.0
    ez, er = oracle.set_e(phi, dx)
    np.testing.assert_allclose(ez, -a, rtol=1e-9)
    np.testing.assert_allclose(er, -b, rtol=1e-9)


def test_quadratic_potential_interior_faces(oracle):
    nz, nr, dx, z, r = _grid()
    phi = 5e5 * z ** 2 + 0.0 * r
    ez, _ = oracle.set_e(phi, dx)
    ez = ez.reshape(nr, nz + 1).T                  # [i, j], i = 0..nz
    i = np.arange(1, nz)
    ref = -(5e5 * ((i * dx) ** 2 - ((i - 1) * dx) ** 2)) / dx
    np.testing.assert_allclose(ez[i, 0], ref, rtol=1e-12)


def test_numpy_restatement_bitwise(oracle, rng):
    nz, nr, dx, z, r = _grid(23, 11)
    phi = rng.standard_normal((nz, nr)) * 50.0
    a = oracle.set_e(phi, dx)
    b = set_e_from_phi(phi, dx)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
