"""Pins of the oracle's Poisson-system assembly (SetFieldSolverCoe,
SetFieldSolverCoeWithoutChi, SetFieldSolverSource, SetSolverBoundary;
R/src/pic/field/FieldSolver.f90:136-301) by what the mathematics fixes:

* uniform medium: the textbook cylindrical five-point Laplacian, A = C = 1,
  D = (j-1.5)/(j-1), E = (j-0.5)/(j-1), on the axis E = 4 (the reference's
  own commented-out forms of the same lines);
* the stencil applied to phi = a z^2 + b r^2 + c returns (2a + 4b) times
  (1 + chi) eps for a uniform medium (div grad of the quadratic);
* flux symmetry for arbitrary eps and chi: the coefficient a node gives its
  z neighbour equals the one it receives back, CoeA(i+1, j) = CoeC(i, j), and
  the r faces carry the radius, (j-1) CoeE(i, j) = j CoeD(i, j+1);
* hand-worked values on a grid whose permittivity names its cell, for the
  axis rows (which average the cells (i-1, 1), (i, 1), (i+1, 1) and (i, 2))
  and for the axis node next to z = L, where the reference's address of cell
  (Nz, 1) holds cell (1, 2);
* the Chi form with Chi = 0 equals the form without Chi bit for bit;
* source closed form and boundary merge.
"""
import numpy as np
import pytest

from oracle import ipm_oracle as O

EPS0 = 8.8542e-12


def _grid(nz, nr, rng=None, eps=None, chi=None, metal=None):
    eps = np.ones((nz - 1, nr - 1)) if eps is None else eps
    chi = np.zeros((nz + 2, nr + 2)) if chi is None else chi
    pt = np.zeros((nz, nr), dtype=np.int32) if metal is None else metal
    return chi, eps, pt


def test_uniform_medium_is_the_cylindrical_laplacian():
    nz, nr = 9, 7
    chi, eps, pt = _grid(nz, nr)
    c = O.solver_coe(chi, eps, pt, with_chi=True)
    for i in range(2, nz):
        for j in range(2, nr):
            a, b, cc, d, e = (c[k][i - 1, j - 1] for k in "ABCDE")
            assert (a, cc) == (1.0, 1.0)
            assert d == pytest.approx((j - 1.5) / (j - 1.0), rel=1e-15)
            assert e == pytest.approx((j - 0.5) / (j - 1.0), rel=1e-15)
            assert b == pytest.approx(-4.0, rel=1e-15)
        # axis
        assert [c[k][i - 1, 0] for k in "ABCDE"] == [1.0, -6.0, 1.0, 0.0, 4.0]
    # outer rows are left to the boundary values
    for k in "ABCDE":
        assert not c[k][0, :].any() and not c[k][-1, :].any() and not c[k][:, -1].any()


@pytest.mark.parametrize("chi0,eps0", [(0.0, 1.0), (0.5, 2.0), (3.25, 0.75)])
def test_stencil_of_a_quadratic_potential(chi0, eps0):
    nz, nr = 12, 10
    chi, eps, pt = _grid(nz, nr, eps=np.full((nz - 1, nr - 1), eps0), chi=np.full((nz + 2, nr + 2), chi0))
    c = O.solver_coe(chi, eps, pt)
    z = np.arange(nz, dtype=np.float64)[:, None]          # grid units, node 1 at z = 0
    r = np.arange(nr, dtype=np.float64)[None, :]          # node j at r = j - 1
    a, b, k = 0.7, -1.3, 5.0
    phi = a * z ** 2 + b * r ** 2 + k + 0 * (z + r)
    for i in range(2, nz):
        for j in range(1, nr):
            lap = (c["A"][i - 1, j - 1] * phi[i - 2, j - 1] + c["B"][i - 1, j - 1] * phi[i - 1, j - 1]
                   + c["C"][i - 1, j - 1] * phi[i, j - 1] + c["E"][i - 1, j - 1] * phi[i - 1, j])
            if j > 1:
                lap += c["D"][i - 1, j - 1] * phi[i - 1, j - 2]
            assert lap == pytest.approx((2 * a + 4 * b) * (1 + chi0) * eps0, rel=1e-12, abs=1e-12)


def test_flux_symmetry_for_any_medium():
    rng = np.random.default_rng(5)
    nz, nr = 11, 9
    chi, eps, pt = _grid(nz, nr, eps=rng.uniform(1, 10, (nz - 1, nr - 1)), chi=rng.uniform(0, 3, (nz + 2, nr + 2)))
    c = O.solver_coe(chi, eps, pt)
    for j in range(2, nr):
        for i in range(2, nz - 1):
            assert c["A"][i, j - 1] == c["C"][i - 1, j - 1]
    for j in range(2, nr - 1):
        for i in range(2, nz):
            assert (j - 1) * c["E"][i - 1, j - 1] == pytest.approx(j * c["D"][i - 1, j], rel=1e-14)
    # rows sum to zero off the metal
    tot = sum(c[k] for k in "ABCDE")
    assert np.abs(tot[1:-1, :-1]).max() < 1e-12 * np.abs(c["B"]).max()


def test_hand_worked_axis_and_aliased_cell():
    nz, nr = 6, 5
    ii, jj = np.meshgrid(np.arange(1, nz), np.arange(1, nr), indexing="ij")
    eps = ii + 10.0 * jj                                   # cell (a, b) holds a + 10 b
    chi = np.zeros((nz + 2, nr + 2))
    chi[3, 1], chi[2, 1], chi[4, 1], chi[3, 2] = 1.0, 0.5, 0.25, 2.0     # Chi(3,1), Chi(2,1), Chi(4,1), Chi(3,2)
    c = O.solver_coe(chi, eps, np.zeros((nz, nr), dtype=np.int32))
    # axis node (3, 1): cells (2,1)=12, (3,1)=13, (4,1)=14, (3,2)=23
    A = (1 + 0.5 * (1.0 + 0.5)) * 0.5 * (13 + 12)
    C = (1 + 0.5 * (1.0 + 0.25)) * 0.5 * (13 + 14)
    E = 4 * (1 + 0.5 * (1.0 + 2.0)) * 0.5 * (13 + 23)
    assert (c["A"][2, 0], c["C"][2, 0], c["D"][2, 0], c["E"][2, 0]) == (A, C, 0.0, E)
    assert c["B"][2, 0] == -(A + C + E)
    # axis node (Nz-1, 1) = (5, 1): CoeC names cell (6, 1), stored at (1, 2) = 21
    assert c["C"][4, 0] == 0.5 * (15 + 21)
    # interior node (3, 3): cells (2,2)=22, (2,3)=32, (3,2)=23, (3,3)=33; Chi(3,2) = 2 enters D only
    assert c["A"][2, 2] == 0.5 * (22 + 32)
    assert c["C"][2, 2] == 0.5 * (23 + 33)
    assert c["D"][2, 2] == (1 + 0.5 * (0.0 + 2.0)) * 0.5 * (22 + 23) * (3 - 1.5) / 2.0
    assert c["E"][2, 2] == 0.5 * (32 + 33) * (3 - 0.5) / 2.0


def test_metal_nodes_and_chi_free_form():
    rng = np.random.default_rng(9)
    nz, nr = 10, 8
    metal = np.zeros((nz, nr), dtype=np.int32)
    metal[0, :], metal[-1, :], metal[4, 3] = 1, 2, 1
    eps = rng.uniform(1, 5, (nz - 1, nr - 1))
    c = O.solver_coe(np.zeros((nz + 2, nr + 2)), eps, metal, with_chi=True)
    w = O.solver_coe(rng.uniform(0, 1, (nz + 2, nr + 2)), eps, metal, with_chi=False)
    for k in "ABCDE":
        assert np.array_equal(c[k], w[k])                 # Chi = 0 form == form without Chi, bit for bit
    m = metal > 0
    assert np.all(c["B"][m] == 1.0) and all(not c[k][m].any() for k in "ACDE")


def test_source_and_boundary_merge():
    rng = np.random.default_rng(2)
    nz, nr, dx = 7, 6, 2.5e-4
    rho = rng.normal(size=(nz, nr))
    iv = rng.uniform(0, 1e9, (nz, nr))
    sg = rng.normal(size=(nz, nr)) * 1e-9
    src = O.solver_source(dx, rho, iv)
    np.testing.assert_allclose(src, -dx * dx / EPS0 * rho, rtol=1e-15)
    src = O.solver_source(dx, rho, iv, sigma=sg)
    np.testing.assert_allclose(src, -dx * dx / EPS0 * (rho + sg * iv), rtol=1e-13)
    mask = np.zeros((nz, nr), dtype=np.int32)
    mask[0, :], mask[:, -1] = 1, 3
    bound = rng.normal(size=(nz, nr))
    merged = O.solver_source(dx, rho, iv, sigma=sg, bound=bound, bound_mask=mask)
    assert np.array_equal(merged[mask != 0], bound[mask != 0])
    assert np.array_equal(merged[mask == 0], src[mask == 0])
