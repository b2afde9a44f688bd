"""Pins of the oracle's push, deposit and node volumes against what the
mathematics fixes (not against a retyped formula):

* push, E = 0: the step is a straight Cartesian line; |v| and R*Vt conserved
* push, linear fields: bilinear gathers on the staggered Ez/Er grids are exact
  for linear fields, so the implicit update has a closed form at each
  particle -- a wrong stagger, index or weight fails it
* push, random fields: a particle at rest gathers a value between the
  field values at the two face rows and the two node columns that bracket
  it (Ez at z = i - 1/2, r = j - 1; Er at z = i - 1, r = j - 1/2), which a
  linear field cannot show: extrapolating from the wrong faces is exact for
  it
* push, uniform Ez: Z' = Z + Vz + 5/2 E, Vz' = Vz + 3/2 E (Az = 0)
* push, axis branch: R' < MinReal gives Vt = MinReal, Vr = |(Vx, Vy)|
* deposit: total weight, first moments in z and r (bilinear weights keep
  linear moments), a particle on a node, uniform density against volumes
* node volumes: the weighting volumes of an all-vacuum cylinder tile it
  exactly (sum = pi R^2 L) and match the analytic shell integrals
"""
import math

import numpy as np
import pytest

from helpers import AMU, ME, QE, random_particles
from ipm2d import geometry as geo

NZ, NR, DX, DT = 21, 13, 1e-3, 5e-11


def efactor(q, m):
    return q / m * DT / (DX / DT) * 0.5


def test_push_zero_field_is_a_straight_line(oracle, rng):
    P = random_particles(NZ, NR, 5000, rng, vscale=0.2, axis_fraction=0.0)
    P[5:7] = 0.0
    ez = np.zeros((NZ + 1) * NR)
    er = np.zeros(NZ * (NR + 1))
    Q = oracle.push(NZ, NR, DX, DT, -QE, ME, ez, er, P)
    X, Y = P[1] + P[3], P[4]
    np.testing.assert_allclose(Q[1], np.hypot(X, Y), rtol=1e-15)
    np.testing.assert_allclose(Q[0], P[0] + P[2], rtol=1e-15)
    np.testing.assert_array_equal(Q[2], P[2])
    np.testing.assert_allclose(Q[3] ** 2 + Q[4] ** 2, P[3] ** 2 + P[4] ** 2, rtol=1e-14)
    np.testing.assert_allclose(Q[1] * Q[4], P[1] * P[4], rtol=1e-12, atol=1e-15)


def _linear_fields(a, b, c, a2, b2, c2):
    # Ez(i, j), i = 0..NZ, j = 1..NR sits at z = i - 0.5, r = j - 1
    iz = np.arange(0, NZ + 1)[:, None] - 0.5
    jr = np.arange(1, NR + 1)[None, :] - 1.0
    ez = a + b * iz + c * jr
    # Er(i, j), i = 1..NZ, j = 0..NR sits at z = i - 1, r = j - 0.5
    iz2 = np.arange(1, NZ + 1)[:, None] - 1.0
    jr2 = np.arange(0, NR + 1)[None, :] - 0.5
    er = a2 + b2 * iz2 + c2 * jr2
    return geo.fortran(ez), geo.fortran(er)


def test_push_gather_stays_between_bracketing_faces(oracle, rng):
    # charge/mass chosen so that EFactor = 1: a particle at rest with
    # Az = Ar = Vt = 0 leaves with Vz = 3/2 Ez and Vr = 3/2 Er (gathered)
    dx, dt = 1.0, 1.0
    q, m = 2.0, 1.0
    # |E| <= 0.3 keeps the rotated radius X = R + 5/2 Er positive (R >= 2),
    # so the rotation leaves Vr = 3/2 Er
    ez = rng.uniform(-0.3, 0.3, (NZ + 1) * NR)
    er = rng.uniform(-0.3, 0.3, NZ * (NR + 1))
    n = 20000
    P = np.zeros((8, n))
    P[0] = rng.uniform(2.0, NZ - 3.0, n)
    P[1] = rng.uniform(2.0, NR - 3.0, n)
    P[7] = 1.0
    Q = oracle.push(NZ, NR, dx, dt, q, m, ez, er, P)
    gz, gr = Q[2] / 1.5, Q[3] / 1.5
    EZ = ez.reshape(NR, NZ + 1).T            # Ez(i, j) at [i, j - 1]
    ER = er.reshape(NR + 1, NZ).T            # Er(i, j) at [i - 1, j]
    z, r = P[0], P[1]
    iz = np.floor(z + 0.5).astype(int)       # faces i (z = i - 1/2) and i + 1 bracket z
    jz = np.ceil(r).astype(int)              # nodes j (r = j - 1) and j + 1 bracket r
    cz = np.stack([EZ[iz, jz - 1], EZ[iz + 1, jz - 1], EZ[iz, jz], EZ[iz + 1, jz]])
    ir = np.ceil(z).astype(int)              # nodes i (z = i - 1) and i + 1 bracket z
    jr = np.floor(r + 0.5).astype(int)       # faces j (r = j - 1/2) and j + 1 bracket r
    cr = np.stack([ER[ir - 1, jr], ER[ir, jr], ER[ir - 1, jr + 1], ER[ir, jr + 1]])
    tol = 1e-13
    assert np.all(gz >= cz.min(axis=0) - tol) and np.all(gz <= cz.max(axis=0) + tol)
    assert np.all(gr >= cr.min(axis=0) - tol) and np.all(gr <= cr.max(axis=0) + tol)


def test_push_linear_field_closed_form(oracle, rng):
    q, m = QE, 19.0 * AMU
    coef = (3.0e4, 2.0e3, -1.5e3, -2.0e4, 1.0e3, 2.5e3)
    ez, er = _linear_fields(*coef)
    n = 4000
    P = np.zeros((8, n))
    P[0] = rng.uniform(0.01, NZ - 1.01, n)
    P[1] = rng.uniform(0.01, NR - 1.01, n)
    P[7] = 1.0
    Q = oracle.push(NZ, NR, DX, DT, q, m, ez, er, P)
    ef = efactor(q, m)
    E2 = (coef[0] + coef[1] * P[0] + coef[2] * P[1]) * ef
    E1 = (coef[3] + coef[4] * P[0] + coef[5] * P[1]) * ef
    np.testing.assert_allclose(Q[2], 1.5 * E2, rtol=1e-9, atol=1e-20)
    np.testing.assert_allclose(Q[0], P[0] + 2.5 * E2, rtol=1e-12, atol=1e-13)
    X = P[1] + 2.5 * E1
    np.testing.assert_allclose(Q[1], np.abs(X), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(Q[3], 1.5 * E1 * np.sign(X), rtol=1e-9, atol=1e-20)
    np.testing.assert_allclose(Q[5], 0.5 * E2, rtol=1e-9, atol=1e-20)
    np.testing.assert_allclose(Q[6], 0.5 * E1 * np.sign(X), rtol=1e-9, atol=1e-20)


def test_push_uniform_ez_with_history(oracle, rng):
    q, m = -QE, ME
    E0 = 1234.5
    ez = np.full((NZ + 1) * NR, E0)
    er = np.zeros(NZ * (NR + 1))
    P = random_particles(NZ, NR, 3000, rng, axis_fraction=0.0)
    Q = oracle.push(NZ, NR, DX, DT, q, m, ez, er, P)
    E = E0 * efactor(q, m)
    az = (P[5] + E) * 0.5
    np.testing.assert_allclose(Q[5], az, rtol=1e-14)
    np.testing.assert_allclose(Q[2], P[2] + E + az, rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(Q[0], P[0] + E + P[2] + E + az, rtol=1e-13)


def test_push_axis_branch(oracle):
    P = np.zeros((8, 2))
    P[0] = [3.3, 4.4]
    P[1] = [0.0, 0.25]
    P[3] = [0.0, -0.25]
    P[4] = [0.0, 0.0]
    P[6] = [-0.125, 0.0]
    ez = np.zeros((NZ + 1) * NR)
    er = np.zeros(NZ * (NR + 1))
    Q = oracle.push(NZ, NR, DX, DT, QE, AMU, ez, er, P)
    # first: X = 0 + (0 + -0.0625) -> R = 0.0625, regular branch with CosA = -1
    assert Q[1, 0] == pytest.approx(0.0625)
    assert Q[4, 1] == 1e-15 and Q[1, 1] == 0.0          # exactly onto the axis
    assert Q[3, 1] == pytest.approx(0.25)
    assert Q[6, 0] == pytest.approx(0.0625)             # Ar*CosA with CosA = -1


def test_deposit_conserves_weight_and_linear_moments(oracle, rng):
    P = random_particles(NZ, NR, 20000, rng, axis_fraction=0.0)
    raw = oracle.deposit_raw(NZ, NR, P)
    assert raw.sum() == pytest.approx(P[7].sum(), rel=1e-13)
    zpos = np.arange(NZ)[:, None] + 0.0 * np.arange(NR)[None, :]
    rpos = 0.0 * np.arange(NZ)[:, None] + np.arange(NR)[None, :]
    assert (raw * zpos).sum() == pytest.approx((P[7] * P[0]).sum(), rel=1e-12)
    assert (raw * rpos).sum() == pytest.approx((P[7] * P[1]).sum(), rel=1e-12)


def test_deposit_particle_on_node(oracle):
    P = np.zeros((8, 3))
    P[0] = [3.0, 7.0, 0.5]
    P[1] = [2.0, 5.0, 0.0]
    P[7] = [2.0, 3.0, 4.0]
    raw = oracle.deposit_raw(NZ, NR, P)
    assert raw[3, 2] == 2.0 and raw[7, 5] == 3.0
    assert raw[0, 0] == 2.0 and raw[1, 0] == 2.0        # R = 0 exactly: all on the axis row
    assert raw.sum() == 9.0


def test_node_volumes_tile_the_cylinder(oracle):
    nz, nr, dx = 17, 11, 2e-3
    cm = np.zeros((nz - 1, nr - 1), dtype=np.int32)
    g = geo.build_geometry(nz, nr, dx, cm, nmetal=0, ndiel=0)
    vol, inv = oracle.node_volumes(nz, nr, dx, g.volume_factor)
    np.testing.assert_array_equal(vol, g.volume)
    np.testing.assert_array_equal(inv, g.inv_volume)
    R, L = dx * (nr - 1), dx * (nz - 1)
    assert vol.sum() == pytest.approx(math.pi * R * R * L, rel=1e-13)
    # interior node j (0-based r index k >= 1): ring volume with hat weight = 2 pi r dx * dx
    k = np.arange(1, nr - 1)
    np.testing.assert_allclose(vol[5, 1:-1], 2 * math.pi * (k * dx) * dx * dx, rtol=1e-13)
    assert vol[5, 0] == pytest.approx(math.pi * dx ** 3 / 3, rel=1e-14)


def test_uniform_density_deposit_recovers_density(oracle):
    from ipm2d import workload as W
    cfg = W.make_config("tiny", nz=33, nr=17, length=32e-3)
    cfg.species[0].ppc = 400
    geom = W.make_geometry(cfg)
    P = W.load_species(cfg, geom, 0, np.random.default_rng(1))
    q, m = W.species_params(cfg, 0)
    rho, _ = oracle.p2ces(cfg.nz, cfg.nr, cfg.dt, q, m, P, geom.inv_volume)
    dens = rho / q
    inner = dens[3:-3, 2:-3] / cfg.n0
    assert abs(inner.mean() - 1.0) < 0.01
    assert inner.std() < 0.12
