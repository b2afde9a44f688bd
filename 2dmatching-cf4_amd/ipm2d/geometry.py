"""Geometry arrays the hot path reads, built the way the reference builds them.

* cell materials, material types and edge types follow the generator
  (R/script/geom/geometry.py:434-467) and the index mapping of
  initGeometryAndMaterialsJson (R/src/base/init.F90:216-435): vacuum 0,
  metals 1..nmetal, dielectrics -1..-ndiel.
* node volume factors follow update_volume (R/script/geom/geometry.py:477-564)
* node weighting volumes follow init.F90:402-424

All arrays are numpy arrays indexed [iz, ir] (0-based); ``fortran(a)``
flattens them z-fastest, the layout the C ABI takes.
"""
from dataclasses import dataclass

import numpy as np

PI = 3.141592653589793238

MAT_VACUUM, MAT_METAL, MAT_DIEL = 0, 1, 2
EDGE_V2M, EDGE_V2D, EDGE_M2V, EDGE_D2V, EDGE_OTHER = 0, 1, 2, 3, 4


def fortran(a):
    return np.ascontiguousarray(np.asarray(a).ravel(order="F"))


@dataclass
class Geometry:
    nz: int
    nr: int
    dx: float
    cell_material: np.ndarray     # (nz-1, nr-1) material index in [-ndiel, nmetal]
    material_type: np.ndarray     # (ndiel+nmetal+1,) type of material m at [m + ndiel]
    ndiel: int
    nmetal: int
    hedge_type: np.ndarray        # (nz, nr-1)
    vedge_type: np.ndarray        # (nz-1, nr)
    volume_factor: np.ndarray     # (nz, nr)
    volume: np.ndarray            # (nz, nr) point_volume_weighting [m^3]
    inv_volume: np.ndarray        # (nz, nr) point_inv_volume_weighting
    cell_load: np.ndarray         # (nz-1, nr-1)
    cell_epsilon: np.ndarray      # (nz-1, nr-1)
    point_type: np.ndarray        # (nz, nr): metal index > 0 on metal nodes, else 0

    def mtype(self, m):
        return self.material_type[m + self.ndiel]


def node_volumes(nz, nr, dx, factor):
    """point_volume_diag / point_inv_volume_diag, R/src/base/init.F90:402-424."""
    j = np.arange(1, nr + 1, dtype=np.float64)
    g = np.empty(nr)
    g[0] = 1.0
    g[1:-1] = j[1:-1] * ((j[1:-1] - 1.0) + j[1:-1]) - (j[1:-1] - 2.0) * ((j[1:-1] - 2.0) + (j[1:-1] - 1.0))
    g[-1] = (j[-1] - 2.0) + 2.0 * (j[-1] - 1.0)
    f = np.asarray(factor, dtype=np.float64)
    vol = f * 1.0 / 3.0 * g[None, :] * PI * dx * dx * dx
    inv = np.where(f > 1.0e-6, 1.0 / np.where(vol != 0.0, vol, 1.0), 0.0)
    return vol, inv


def _volume_factor(nz, nr, vac):
    """update_volume, R/script/geom/geometry.py:477-564 (vac: (nz-1, nr-1) bool)."""
    pv = np.zeros((nz, nr))
    pv[0, 0] = 0.5 if vac[0, 0] else 0.0
    pv[nz - 1, 0] = 0.5 if vac[nz - 2, 0] else 0.0
    pv[0, nr - 1] = 0.5 if vac[0, nr - 2] else 0.0
    pv[nz - 1, nr - 1] = 0.5 if vac[nz - 2, nr - 2] else 0.0
    z = np.arange(1, nz - 1)
    pv[z, 0] = 0.5 * vac[z - 1, 0] + 0.5 * vac[z, 0]
    pv[z, nr - 1] = 0.5 * vac[z - 1, nr - 2] + 0.5 * vac[z, nr - 2]
    r = np.arange(1, nr - 1, dtype=np.float64)
    ri = np.arange(1, nr - 1)
    lo = 0.25 * (1 - 1 / 4 / r)
    hi = 0.25 * (1 + 1 / 4 / r)
    pv[0, ri] = np.where(vac[0, ri - 1], lo, 0.0) + np.where(vac[0, ri], hi, 0.0)
    pv[nz - 1, ri] = np.where(vac[nz - 2, ri - 1], lo, 0.0) + np.where(vac[nz - 2, ri], hi, 0.0)
    zz = np.arange(1, nz - 1)[:, None]
    rr = ri[None, :]
    s1 = np.where(vac[zz - 1, rr - 1], lo[None, :], 0.0)
    s2 = np.where(vac[zz - 1, rr], hi[None, :], 0.0)
    s3 = np.where(vac[zz, rr - 1], lo[None, :], 0.0)
    s4 = np.where(vac[zz, rr], hi[None, :], 0.0)
    pv[1:nz - 1, 1:nr - 1] = s1 + s2 + s3 + s4
    return pv


def _edge_types(ctype):
    """update_edge_type, R/script/geom/geometry.py:434-467 (ctype: cell material type)."""
    nzc, nrc = ctype.shape
    nz, nr = nzc + 1, nrc + 1
    he = np.full((nz, nr - 1), EDGE_OTHER, dtype=np.int32)
    top, bot = ctype[1:, :], ctype[:-1, :]
    h = he[1:nz - 1, :]
    h[(bot == MAT_VACUUM) & (top == MAT_METAL)] = EDGE_V2M
    h[(bot == MAT_VACUUM) & (top == MAT_DIEL)] = EDGE_V2D
    h[(bot == MAT_METAL) & (top == MAT_VACUUM)] = EDGE_M2V
    h[(bot == MAT_DIEL) & (top == MAT_VACUUM)] = EDGE_D2V
    ve = np.full((nz - 1, nr), EDGE_OTHER, dtype=np.int32)
    right, left = ctype[:, 1:], ctype[:, :-1]
    v = ve[:, 1:nr - 1]
    v[(left == MAT_VACUUM) & (right == MAT_METAL)] = EDGE_V2M
    v[(left == MAT_VACUUM) & (right == MAT_DIEL)] = EDGE_V2D
    v[(left == MAT_METAL) & (right == MAT_VACUUM)] = EDGE_M2V
    v[(left == MAT_DIEL) & (right == MAT_VACUUM)] = EDGE_D2V
    return he, ve


def build_geometry(nz, nr, dx, cell_material, nmetal, ndiel, permittivity=None):
    """Assemble every derived array from a (nz-1, nr-1) map of material indices."""
    cm = np.asarray(cell_material, dtype=np.int32)
    assert cm.shape == (nz - 1, nr - 1)
    mt = np.zeros(ndiel + nmetal + 1, dtype=np.int32)
    mt[:ndiel] = MAT_DIEL
    mt[ndiel] = MAT_VACUUM
    mt[ndiel + 1:] = MAT_METAL
    ctype = mt[cm + ndiel]
    vac = ctype == MAT_VACUUM
    vf = _volume_factor(nz, nr, vac)
    vol, inv = node_volumes(nz, nr, dx, vf)
    he, ve = _edge_types(ctype)
    eps = np.ones((nz - 1, nr - 1))
    if permittivity is not None:
        for m, e in permittivity.items():
            eps[cm == m] = e
    pt = metal_point_types(np.zeros((nz, nr), dtype=np.int32), cm, ctype == MAT_METAL)
    return Geometry(nz=nz, nr=nr, dx=dx, cell_material=cm, material_type=mt, ndiel=ndiel, nmetal=nmetal,
                    hedge_type=he, vedge_type=ve, volume_factor=vf, volume=vol, inv_volume=inv,
                    cell_load=vac.astype(np.int32), cell_epsilon=eps, point_type=pt)


def metal_point_types(point_type, cell_material, metal):
    """Each metal cell marks its four nodes with its material index; cells
    are visited i-outer, j-inner and a later cell overwrites an earlier one
    (R/src/base/init.F90:388-398), so a node shared by two metals takes the
    index of the cell that comes last in that order."""
    pt = np.array(point_type, dtype=np.int32, copy=True)
    nzc, nrc = cell_material.shape
    for a, b in ((1, 1), (1, 0), (0, 1), (0, 0)):    # lowest priority first
        sl = pt[a:a + nzc, b:b + nrc]
        sl[metal] = cell_material[metal]
    return pt


def ccp_geometry(nz, nr, dx, electrode_layers=1, dielectric_ring=0):
    """Capacitively coupled cylinder: metal 1 (powered) at z = 0, metal 2
    (grounded) at z = L, each ``electrode_layers`` cells thick; optionally a
    dielectric (-1) ring of ``dielectric_ring`` cells at the outer radius."""
    cm = np.zeros((nz - 1, nr - 1), dtype=np.int32)
    ndiel = 0
    if dielectric_ring > 0:
        cm[:, nr - 1 - dielectric_ring:] = -1
        ndiel = 1
    cm[:electrode_layers, :] = 1
    cm[nz - 1 - electrode_layers:, :] = 2
    return build_geometry(nz, nr, dx, cm, nmetal=2, ndiel=ndiel, permittivity={-1: 2.1} if ndiel else None)


def set_e_from_phi(phi, dx):
    """SetE (R/src/pic/field/FieldSource.f90:95-120) after the linear ghost
    extrapolation ghost = 2*edge - next (R/src/pic/domain/PICCom.F90:474-590).

    phi: (nz, nr) node potential.  Returns Fortran-layout flat Ez (nz+1)*nr
    (Ez(0:nz, 1:nr)) and Er nz*(nr+1) (Er(1:nz, 0:nr)).
    """
    nz, nr = phi.shape
    p = np.zeros((nz + 2, nr + 2))
    p[1:-1, 1:-1] = phi
    p[0, 1:-1] = 2.0 * phi[0, :] - phi[1, :]
    p[-1, 1:-1] = 2.0 * phi[-1, :] - phi[-2, :]
    p[1:-1, 0] = 2.0 * phi[:, 0] - phi[:, 1]
    p[1:-1, -1] = 2.0 * phi[:, -1] - phi[:, -2]
    ez = (p[0:nz + 1, 1:nr + 1] - p[1:nz + 2, 1:nr + 1]) / dx      # i = 0..nz
    er = (p[1:nz + 1, 0:nr + 1] - p[1:nz + 1, 1:nr + 2]) / dx      # j = 0..nr
    return fortran(ez), fortran(er)
