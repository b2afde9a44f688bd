"""Shared builders for the parity tests: small grids, particles that hit every
branch of the hot path, and sorted comparisons of particle sets."""
import numpy as np

from ipm2d import geometry as geo
from ipm2d import workload as W
from ipm2d import xsec

QE = xsec.ELECTRON_CHARGE
ME = xsec.ELECTRON_MASS
AMU = xsec.ATOMIC_MASS


def small_geometry(nz=33, nr=17, dx=1e-3, dielectric=True):
    """Electrodes at both z ends plus (optionally) a dielectric block in the
    interior and a dielectric ring at the outer radius, so every absorption
    rule (delete, reflect, metal, 4 dielectric edge cases) is reachable."""
    cm = np.zeros((nz - 1, nr - 1), dtype=np.int32)
    cm[0, :] = 1
    cm[nz - 2, :] = 2
    ndiel = 0
    if dielectric:
        cm[10:14, 5:8] = -1
        cm[:, nr - 3:] = -1
        cm[0, :] = 1
        cm[nz - 2, :] = 2
        ndiel = 1
    return geo.build_geometry(nz, nr, dx, cm, nmetal=2, ndiel=ndiel, permittivity={-1: 2.1} if ndiel else None)


def random_fields(nz, nr, rng, scale=2e4):
    ez = rng.standard_normal((nz + 1) * nr) * scale
    er = rng.standard_normal(nz * (nr + 1)) * scale
    return ez, er


def random_particles(nz, nr, n, rng, vscale=0.3, inside=True, axis_fraction=0.02):
    """(8, n) particles strictly inside the domain (0 < Z < nz-1, 0 < R < nr-1)."""
    P = np.zeros((8, n))
    P[0] = rng.uniform(1e-6, nz - 1 - 1e-6, n)
    P[1] = rng.uniform(1e-6, nr - 1 - 1e-6, n)
    k = int(n * axis_fraction)
    if k:
        P[1, :k] = rng.uniform(0.0, 0.6, k)       # near the axis: N2H = 0 rows, the R < MinReal branch
        P[1, 0] = 0.0                              # exactly on the axis
    P[2:5] = rng.standard_normal((3, n)) * vscale
    P[5:7] = rng.standard_normal((2, n)) * 0.01
    P[7] = rng.uniform(0.5, 2.0, n) * 1e6
    return P


def sort_set(P):
    """Particles (8, n) sorted lexicographically (Z, R, Vz, ...): order-free comparison."""
    P = np.asarray(P)
    if P.shape[1] == 0:
        return P
    idx = np.lexsort(P[::-1])
    return P[:, idx]


def assert_close(a, b, rtol=1e-12, atol=0.0, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    scale = np.maximum(np.abs(a), np.abs(b))
    bad = np.abs(a - b) > (rtol * scale + atol)
    if bad.any():
        i = np.argwhere(bad)[0]
        raise AssertionError(f"{what}: {bad.sum()} mismatches; first at {tuple(i)}: {a[tuple(i)]!r} vs {b[tuple(i)]!r}")


def tiny_config():
    return W.make_config("tiny")
