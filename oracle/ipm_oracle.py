"""ctypes loader of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline; the product
(2dmatching-cf4_amd/) never does.  Grids use numpy (nz, nr) arrays indexed
[iz-1, ir-1]; particles are (8, n) float64 SoA arrays Z R Vz Vr Vt Az Ar WQ.
"""
import contextlib
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libipm_oracle.so")
LIB_FAST = os.path.join(HERE, "build", "libipm_oracle_fast.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)


class OraGeom(C.Structure):
    _fields_ = [("nz", C.c_int), ("nr", C.c_int), ("dx", C.c_double), ("dt", C.c_double),
                ("cell_material", _ip), ("hedge_type", _ip), ("vedge_type", _ip), ("material_type", _ip),
                ("ndiel", C.c_int), ("nmetal", C.c_int)]


class OraTable(C.Structure):
    _fields_ = [("model", C.c_int), ("nreaction", C.c_int), ("nsigma", C.c_int),
                ("emin", C.c_double), ("einterval", C.c_double), ("emax", C.c_double),
                ("collision_ratio", C.c_double), ("rtype", _ip), ("resultant", _ip), ("threshold", _dp),
                ("prob", _dp), ("gas_mass", C.c_double), ("gas_temperature", C.c_double),
                ("beta_max", C.c_double)]


class OraSetup(C.Structure):
    _fields_ = [("ns", C.c_int), ("ng", C.c_int), ("dx", C.c_double), ("dt", C.c_double),
                ("charge", _dp), ("mass", _dp), ("tab", C.POINTER(OraTable))]


_libs = {}
_timed_path = None
_override_path = None


def use_timed_library(path):
    """Point lib(fast=True) -- the build the CPU baseline times -- at another
    build of the same source (bench.py compiles one for the host it runs on)."""
    global _timed_path
    _timed_path = path


@contextlib.contextmanager
def library_override(path):
    """Within the block every oracle function runs against another build of
    the source (the mutant tests compile copies with a planted mistake)."""
    global _override_path
    old = _override_path
    _override_path = path
    try:
        yield
    finally:
        _override_path = old


def lib(fast=False):
    path = _override_path or ((_timed_path or LIB_FAST) if fast else LIB)
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"oracle not built: {path} (run oracle/build_oracle.py)")
    L = C.CDLL(path)
    L.ora_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.ora_uniform.restype = C.c_double
    L.ora_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
    L.ora_state_key.argtypes = [C.c_double] * 5 + [C.POINTER(C.c_uint32)]
    L.ora_node_volumes.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp]
    L.ora_push.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp,
                           C.c_int64] + [_dp] * 7
    L.ora_absorb.argtypes = [C.POINTER(OraGeom), C.c_double, _i64p, C.POINTER(_dp), _dp, _dp]
    L.ora_deposit_raw.argtypes = [C.c_int, C.c_int, C.c_int64, _dp, _dp, _dp, _dp]
    L.ora_deposit_finish.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
    L.ora_set_e.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp]
    L.ora_solver_coe.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _ip, _dp]
    L.ora_solver_coe.restype = None
    L.ora_solver_source.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _ip, _dp]
    L.ora_solver_source.restype = None
    L.ora_split_merge.argtypes = [_i64p, C.POINTER(_dp), C.c_int64, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_uint64, C.c_uint32, C.c_int, _i64p]
    L.ora_split_merge.restype = C.c_int
    L.ora_recomb_density.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int64,
                                     C.POINTER(_dp), _dp, _dp]
    L.ora_recombine.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _i64p, C.POINTER(_dp), _dp,
                                C.c_int, _ip, _dp, C.c_double, C.c_uint64, C.c_uint32, _i64p]
    L.ora_recombine.restype = C.c_int
    L.ora_moments.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp,
                              C.c_int, C.c_int64, C.POINTER(_dp)] + [_dp] * 6
    L.ora_mcc.restype = C.c_int
    L.ora_mcc.argtypes = [C.POINTER(OraSetup), C.c_uint64, C.c_uint32, _i64p, _i64p, C.POINTER(_dp * 8), _i64p]
    L.ora_collide_given.restype = C.c_int
    L.ora_collide_given.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                                    _dp, C.c_uint64, C.c_uint32, _dp, _dp, _dp, _ip]
    L.ora_vel_maxwell.restype = C.c_uint32
    L.ora_vel_maxwell.argtypes = [C.c_double, C.c_double, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32, _dp, _dp, _dp]
    L.ora_step.restype = C.c_int
    L.ora_step.argtypes = [C.POINTER(OraGeom), C.POINTER(OraSetup), _dp, _dp, C.c_uint64, C.c_uint32, _i64p,
                           _i64p, C.c_void_p, _dp, _dp, _dp, _i64p]
    L.ora_cos_chi.restype = C.c_double
    L.ora_cos_chi.argtypes = [C.c_int, C.c_double, C.c_uint64, C.c_uint32, C.c_uint32]
    L.ora_secondary_energy.restype = C.c_double
    L.ora_secondary_energy.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_uint32]
    L.ora_locate.restype = C.c_int
    L.ora_locate.argtypes = [_dp, C.c_int, C.c_double]
    L.ora_sigma_normalize.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
    L.ora_sigma_normalize_reactive.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, C.c_double,
                                               C.c_double, _dp]
    L.ora_probability_nonreactive.restype = C.c_double
    L.ora_probability_nonreactive.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                              C.c_double, _dp, _dp]
    L.ora_probability_reactive.restype = C.c_double
    L.ora_probability_reactive.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_double, _dp, _dp]
    L.ora_probability_hard_sphere.restype = C.c_double
    L.ora_probability_hard_sphere.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                              C.c_double, C.c_double, C.c_double, _dp]
    _libs[path] = L
    return L


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def F(a, dtype=np.float64):
    """numpy [iz, ir] array -> contiguous z-fastest (Fortran) flat copy."""
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))


def unF(flat, nz, nr):
    return np.asarray(flat).reshape(nr, nz).T.copy()


# ------------------------------------------------------------------ RNG
def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().ora_philox4x32_10(c, k, o)
    return [o[i] for i in range(4)]


def uniform(seed, step, tag, c0, c1, c2, draw):
    return lib().ora_uniform(seed, step, tag, c0, c1, c2, draw)


def state_key(z, r, vz, vr, vt):
    o = (C.c_uint32 * 3)()
    lib().ora_state_key(z, r, vz, vr, vt, o)
    return [o[0], o[1], o[2]]


# ------------------------------------------------------------- kernels
def node_volumes(nz, nr, dx, factor):
    f = F(factor)
    vol = np.zeros(nz * nr)
    inv = np.zeros(nz * nr)
    lib().ora_node_volumes(nz, nr, dx, _d(f), _d(vol), _d(inv))
    return unF(vol, nz, nr), unF(inv, nz, nr)


def push(nz, nr, dx, dt, charge, mass, ez_flat, er_flat, P, fast=False):
    """MoveESIm on a copy of P (8, n); returns the pushed copy."""
    P = np.array(P, dtype=np.float64, copy=True, order="C")
    ez = np.ascontiguousarray(ez_flat, dtype=np.float64)
    er = np.ascontiguousarray(er_flat, dtype=np.float64)
    n = P.shape[1]
    rows = [P[f] for f in range(7)]
    lib(fast).ora_push(nz, nr, dx, dt, charge, mass, _d(ez), _d(er), n, *[_d(r) for r in rows])
    return P


class _GeomRef:
    def __init__(self, geom):
        self.cm = F(geom.cell_material, np.int32)
        self.he = F(geom.hedge_type, np.int32)
        self.ve = F(geom.vedge_type, np.int32)
        self.mt = np.ascontiguousarray(geom.material_type, dtype=np.int32)
        self.s = OraGeom(geom.nz, geom.nr, geom.dx, 0.0, _i(self.cm), _i(self.he), _i(self.ve), _i(self.mt),
                         geom.ndiel, geom.nmetal)


def absorb(geom, charge, P, fast=False):
    """MaterialAborptionParticle on a copy of P with the reference's
    swap-with-last deletion order.  Returns (P_out, metal_charge, sigma[nz, nr])."""
    P = np.array(P, dtype=np.float64, copy=True, order="C")
    g = _GeomRef(geom)
    n = C.c_int64(P.shape[1])
    ptrs = (_dp * 8)(*[_d(P[f]) for f in range(8)])
    mc = np.zeros(max(geom.nmetal, 1))
    sg = np.zeros(geom.nz * geom.nr)
    lib(fast).ora_absorb(C.byref(g.s), charge, C.byref(n), ptrs, _d(mc), _d(sg))
    return P[:, :n.value].copy(), mc[:geom.nmetal], unF(sg, geom.nz, geom.nr)


def recomb_density(nz, nr, dx, dt, mass, P, want_t=True, fast=False):
    """Cell number density and inverse mean energy (1/eV) of one bundle
    (Recombination.F90:175-206), each (nz-1, nr-1)."""
    P = np.ascontiguousarray(P, dtype=np.float64)
    ncell = (nz - 1) * (nr - 1)
    d = np.zeros(ncell)
    t = np.zeros(ncell)
    ptrs = (_dp * 8)(*[_d(P[f]) for f in range(8)])
    lib(fast).ora_recomb_density(nz, nr, dx, dt, mass, 1 if want_t else 0, P.shape[1], ptrs, _d(d), _d(t))
    return d.reshape(nr - 1, nz - 1).T, t.reshape(nr - 1, nz - 1).T


def recombine(nz, nr, dx, dt, species, masses, reactions, rates, time_step, seed, step, fast=False):
    """executingRecombination2D (Recombination.F90:164-243) on copies of the
    bundles.  reactions: (neg, pos, type) triples with 0-based species.
    Returns (bundles_out, deleted per species)."""
    ns = len(species)
    Ps = [np.array(P, dtype=np.float64, copy=True, order="C") for P in species]
    n = np.array([P.shape[1] for P in Ps], dtype=np.int64)
    ptrs = (_dp * (8 * ns))(*[_d(Ps[s][f]) for s in range(ns) for f in range(8)])
    rl = np.asfortranarray(np.asarray(reactions, dtype=np.int32).reshape(-1, 3))
    flat = np.ascontiguousarray(rl.T.reshape(-1))          # Fortran (nreaction, 3) order
    rt = np.ascontiguousarray(np.asarray(rates, dtype=np.float64))
    ms = np.ascontiguousarray(np.asarray(masses, dtype=np.float64))
    deleted = np.zeros(ns, dtype=np.int64)
    rc = lib(fast).ora_recombine(nz, nr, dx, dt, ns, n.ctypes.data_as(_i64p), ptrs, _d(ms), rl.shape[0], _i(flat),
                                 _d(rt), time_step, seed, step, deleted.ctypes.data_as(_i64p))
    if rc:
        raise ValueError("invalid recombination reaction")
    return [Ps[s][:, :n[s]].copy() for s in range(ns)], deleted


def split_merge(P, weight, alpha, grid_r, merge_limit, seed, step, s, mode=3, fast=False):
    """split_particle_weight then merge_particle_random (Rezoning.F90:53-92)
    on a copy of P.  Returns (P_out, stats[copies, kept, deleted])."""
    P = np.asarray(P, dtype=np.float64)
    n0 = P.shape[1]
    extra = 0
    if mode & 1 and n0:
        ref = weight * ((1.0 - alpha) * P[1] / grid_r + alpha)
        with np.errstate(divide="ignore", invalid="ignore"):
            ratio = np.where(ref > 0, P[7] / np.where(ref > 0, ref, 1.0), 0.0)
        extra = int(np.floor(ratio[ratio >= 2.0]).sum() - np.count_nonzero(ratio >= 2.0)) + 1
    cap = n0 + extra
    Q = np.zeros((8, max(cap, 1)))
    Q[:, :n0] = P
    n = C.c_int64(n0)
    st = np.zeros(3, dtype=np.int64)
    ptrs = (_dp * 8)(*[_d(Q[f]) for f in range(8)])
    rc = lib(fast).ora_split_merge(C.byref(n), ptrs, cap, mode, weight, alpha, grid_r, merge_limit, seed, step, s,
                                   st.ctypes.data_as(_i64p))
    if rc:
        raise ValueError("split copies exceed the capacity")
    return Q[:, :n.value].copy(), st


def deposit_raw(nz, nr, P, fast=False):
    z = np.ascontiguousarray(P[0], dtype=np.float64)
    r = np.ascontiguousarray(P[1], dtype=np.float64)
    w = np.ascontiguousarray(P[7], dtype=np.float64)
    raw = np.zeros(nz * nr)
    lib(fast).ora_deposit_raw(nz, nr, z.size, _d(z), _d(r), _d(w), _d(raw))
    return unF(raw, nz, nr)


def deposit_finish(nz, nr, dt, charge, mass, raw, inv_vol):
    rf, iv = F(raw), F(inv_vol)
    rho = np.zeros(nz * nr)
    chi = np.zeros(nz * nr)
    lib().ora_deposit_finish(nz, nr, dt, charge, mass, _d(rf), _d(iv), _d(rho), _d(chi))
    return unF(rho, nz, nr), unF(chi, nz, nr)


def p2ces(nz, nr, dt, charge, mass, P, inv_vol):
    return deposit_finish(nz, nr, dt, charge, mass, deposit_raw(nz, nr, P), inv_vol)


def set_e(phi, dx):
    """SetE from a (nz, nr) potential: Fortran-layout flat Ez, Er."""
    nz, nr = phi.shape
    ph = F(phi)
    ez = np.zeros((nz + 1) * nr)
    er = np.zeros(nz * (nr + 1))
    lib().ora_set_e(nz, nr, dx, _d(ph), _d(ez), _d(er))
    return ez, er


def solver_coe(chi, cell_epsilon, point_type, with_chi=True):
    """SetFieldSolverCoe / SetFieldSolverCoeWithoutChi: chi (nz+2, nr+2),
    cell_epsilon (nz-1, nr-1), point_type (nz, nr) -> {"A".."E": (nz, nr)}."""
    nz, nr = point_type.shape
    assert chi.shape == (nz + 2, nr + 2) and cell_epsilon.shape == (nz - 1, nr - 1)
    ch, ep = F(chi), F(cell_epsilon)
    pt = np.ascontiguousarray(np.asarray(point_type).T.reshape(-1), dtype=np.int32)
    out = np.zeros(5 * nz * nr)
    lib().ora_solver_coe(nz, nr, 1 if with_chi else 0, _d(ch), _d(ep), pt.ctypes.data_as(_ip), _d(out))
    return dict(zip("ABCDE", out.reshape(5, nr, nz).transpose(0, 2, 1)))


def solver_source(dx, rho, inv_vol, sigma=None, bound=None, bound_mask=None):
    """SetFieldSolverSource + SetSolverBoundary on (nz, nr) grids."""
    nz, nr = rho.shape
    opt = [None if a is None else F(a) for a in (sigma, bound)]
    mk = None if bound_mask is None else np.ascontiguousarray(np.asarray(bound_mask).T.reshape(-1), dtype=np.int32)
    out = np.zeros(nz * nr)
    lib().ora_solver_source(nz, nr, dx, _d(F(rho)), None if opt[0] is None else _d(opt[0]), _d(F(inv_vol)),
                            None if opt[1] is None else _d(opt[1]), None if mk is None else mk.ctypes.data_as(_ip),
                            _d(out))
    return out.reshape(nr, nz).T


def moments(nz, nr, dx, dt, charge, mass, ez_flat, er_flat, P, inv_vol, implicit_j=True):
    """WeightingParticleField of one species: dict of (nz, nr) arrays
    rho, energy (eV), mp, jz, jr."""
    P = np.array(P, dtype=np.float64, copy=True, order="C")
    ez = np.ascontiguousarray(ez_flat, dtype=np.float64)
    er = np.ascontiguousarray(er_flat, dtype=np.float64)
    iv = F(inv_vol)
    outs = {k: np.zeros(nz * nr) for k in ("rho", "energy", "mp", "jz", "jr")}
    ptrs = (_dp * 8)(*[_d(P[f]) for f in range(8)])
    lib().ora_moments(nz, nr, dx, dt, charge, mass, _d(ez), _d(er), 1 if implicit_j else 0, P.shape[1], ptrs,
                      _d(iv), *[_d(outs[k]) for k in ("rho", "energy", "mp", "jz", "jr")])
    return {k: unF(v, nz, nr) for k, v in outs.items()}


def window_plain(geom, species, e0, e1, coef, mask, weights, dt, metal=1):
    """The collisionless window by the plain method: for each step n the
    field E0 + coef[n] E1 (formed on the host), MoveESIm and
    MaterialAborptionParticle per species, then the loss
    l[n] = dt * sum_s weights[s] * sum_nodes mask * JzOne_s (explicit current).
    species: list of (charge, mass, P).  Returns (l, electrode charge
    collected per step by metal `metal`, final particle arrays)."""
    nz, nr, dx = geom.nz, geom.nr, geom.dx
    Ps = [np.array(P, dtype=np.float64, copy=True) for _, _, P in species]
    ez0, er0 = (np.asarray(x, dtype=np.float64).ravel() for x in e0)
    ez1, er1 = (np.asarray(x, dtype=np.float64).ravel() for x in e1)
    loss, qel = [], []
    for n, c in enumerate(coef):
        ez = ez0 + c * ez1
        er = er0 + c * er1
        ln, qn = 0.0, 0.0
        for s, (q, m, _) in enumerate(species):
            P = push(nz, nr, dx, dt, q, m, ez, er, Ps[s])
            P, mc, _ = absorb(geom, q, P)
            Ps[s] = P
            if geom.nmetal >= metal:
                qn += mc[metal - 1]
            if weights[s] != 0.0:
                mo = moments(nz, nr, dx, dt, q, m, ez, er, P, geom.inv_volume, implicit_j=False)
                ln += weights[s] * dt * float(np.sum(mask * mo["jz"]))
        loss.append(ln)
        qel.append(qn)
    return np.array(loss), np.array(qel), Ps


class MccSetup:
    """Keeps the ctypes structures (and the arrays they point to) alive."""

    def __init__(self, dx, dt, charges, masses, tables, ng=1):
        self.ns = len(charges)
        self.ng = ng
        self.charge = np.ascontiguousarray(charges, dtype=np.float64)
        self.mass = np.ascontiguousarray(masses, dtype=np.float64)
        self.keep = []
        arr = (OraTable * (self.ns * ng))()
        for s in range(self.ns):
            for g in range(ng):
                b = tables.get((s, g))
                t = arr[s * ng + g]
                if b is None or b.model == 0:
                    t.model = 0
                    continue
                rt = np.ascontiguousarray(b.rtype, dtype=np.int32)
                rs = np.ascontiguousarray(b.resultant, dtype=np.int32)
                th = np.ascontiguousarray(b.threshold, dtype=np.float64)
                pr = np.ascontiguousarray(b.prob, dtype=np.float64)
                self.keep += [rt, rs, th, pr]
                t.model, t.nreaction, t.nsigma = int(b.model), int(b.nreaction), int(b.nsigma)
                t.emin, t.einterval, t.emax = float(b.emin), float(b.einterval), float(b.emax)
                t.collision_ratio = float(b.collision_ratio)
                t.rtype, t.resultant, t.threshold, t.prob = _i(rt), _i(rs), _d(th), _d(pr)
                t.gas_mass, t.gas_temperature, t.beta_max = float(b.gas_mass), float(b.gas_temperature), \
                    float(b.beta_max)
        self.tab = arr
        self.s = OraSetup(self.ns, ng, dx, dt, _d(self.charge), _d(self.mass), arr)


def mcc(setup, seed, step, species, extra_capacity=None, counts=False, fast=False):
    """MCC over every species (list of (8, n) arrays, copied).  Returns
    (list of (8, n') arrays, counts[ns, ng, 128] or None)."""
    ns = setup.ns
    caps = []
    bufs = []
    total = sum(np.asarray(P).shape[1] for P in species)
    for s in range(ns):
        P = np.asarray(species[s], dtype=np.float64)
        n = P.shape[1]
        cap = n + (extra_capacity if extra_capacity is not None else 2 * total + 1024)
        B = np.zeros((8, cap))
        B[:, :n] = P
        bufs.append(B)
        caps.append(cap)
    nvec = np.array([np.asarray(species[s]).shape[1] for s in range(ns)], dtype=np.int64)
    capv = np.array(caps, dtype=np.int64)
    ptrs = ((_dp * 8) * ns)()
    for s in range(ns):
        for f in range(8):
            ptrs[s][f] = _d(bufs[s][f])
    cnt = np.zeros(ns * setup.ng * 128, dtype=np.int64) if counts else None
    rc = lib(fast).ora_mcc(C.byref(setup.s), seed, step, nvec.ctypes.data_as(_i64p), capv.ctypes.data_as(_i64p),
                           ptrs, cnt.ctypes.data_as(_i64p) if counts else None)
    if rc != 0:
        raise RuntimeError("oracle MCC capacity overflow")
    out = [bufs[s][:, :nvec[s]].copy() for s in range(ns)]
    return out, (cnt.reshape(ns, setup.ng, 128) if counts else None)


def collide_given(model, species, rtype, resultant, threshold, species_mass, gas_mass, poi, pot_v, seed=1, c0=0):
    """One reaction against a given target velocity (ora_collide_given).
    poi: 8 values Z R Vz Vr Vt Az Ar WQ (velocities in m/s); pot_v: Vz Vr Vt.
    Returns (particle afterwards, energy afterwards, [(species, product), ...])."""
    a = np.ascontiguousarray(poi, dtype=np.float64)
    t = np.ascontiguousarray(pot_v, dtype=np.float64)
    out = np.zeros(8)
    en = C.c_double()
    prod = np.zeros(16)
    ps = np.zeros(2, dtype=np.int32)
    k = lib().ora_collide_given(model, species, rtype, resultant, threshold, species_mass, gas_mass, _d(a), _d(t),
                                seed, c0, _d(out), C.byref(en), _d(prod), _i(ps))
    return out, en.value, [(int(ps[j]), prod[8 * j:8 * j + 8].copy()) for j in range(k)]


def vel_maxwell(mass, temperature, seed, slot, step, draw, tag):
    vz, vr, vt = C.c_double(), C.c_double(), C.c_double()
    k = lib().ora_vel_maxwell(mass, temperature, seed, slot, step, draw, tag, C.byref(vz), C.byref(vr), C.byref(vt))
    return vz.value, vr.value, vt.value, k


# ---------------------------------------------------------------- tables
def locate(xx, x):
    xx = np.ascontiguousarray(xx, dtype=np.float64)
    return lib().ora_locate(_d(xx), xx.size, x)


def sigma_normalize(reactions, emin, eint, nsigma):
    nr = len(reactions)
    e = np.zeros((nr, 500))
    s = np.zeros((nr, 500))
    th = np.array([r.threshold for r in reactions], dtype=np.float64)
    for j, r in enumerate(reactions):
        e[j, :len(r.energy)] = r.energy
        s[j, :len(r.sigma)] = r.sigma
    out = np.zeros(nsigma * nr)
    lib().ora_sigma_normalize(nr, nsigma, emin, eint, _d(th), _d(e), _d(s), _d(out))
    return out.reshape(nsigma, nr)


def sigma_normalize_reactive(thresholds, natom_s, natom_g, emin, eint, nsigma):
    th = np.ascontiguousarray(thresholds, dtype=np.float64)
    out = np.zeros(nsigma * th.size)
    lib().ora_sigma_normalize_reactive(th.size, nsigma, emin, eint, _d(th), natom_s, natom_g, _d(out))
    return out.reshape(nsigma, th.size)


def probability_nonreactive(sigma, emin, eint, ms, mg, ng):
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    nsigma, nr = sig.shape
    out = np.zeros(nsigma * nr)
    mx = lib().ora_probability_nonreactive(nr, nsigma, emin, eint, ms, mg, ng, _d(sig), _d(out))
    return out.reshape(nsigma, nr), mx


def probability_reactive(sigma, ms, mg, ng, beta_max, pol, dt):
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    nsigma, nr = sig.shape
    out = np.zeros(nsigma * nr)
    ratio = lib().ora_probability_reactive(nr, nsigma, ms, mg, ng, beta_max, pol, dt, _d(sig), _d(out))
    return out.reshape(nsigma, nr), ratio


def probability_hard_sphere(nreaction, emin, eint, nsigma, ms, mg, ng, r_s, r_g):
    out = np.zeros(nsigma * nreaction)
    mx = lib().ora_probability_hard_sphere(nreaction, nsigma, emin, eint, ms, mg, ng, r_s, r_g, _d(out))
    return out.reshape(nsigma, nreaction), mx


class StepState:
    """Persistent in-place buffers for ora_step (timed CPU baseline)."""

    def __init__(self, geom, setup, species, ez, er, extra=0.05, fast=True):
        self.geom = _GeomRef(geom)
        self.setup = setup
        self.fast = fast
        ns = setup.ns
        self.nz, self.nr = geom.nz, geom.nr
        self.n = np.array([np.asarray(P).shape[1] for P in species], dtype=np.int64)
        self.cap = np.array([int(n * (1 + extra)) + 1024 for n in self.n], dtype=np.int64)
        self.bufs = []
        self.ptrs = ((_dp * 8) * ns)()
        for s in range(ns):
            B = np.zeros((8, int(self.cap[s])))
            B[:, :self.n[s]] = species[s]
            self.bufs.append(B)
            for f in range(8):
                self.ptrs[s][f] = _d(B[f])
        self.ez = np.ascontiguousarray(ez, dtype=np.float64)
        self.er = np.ascontiguousarray(er, dtype=np.float64)
        self.rho = np.zeros(ns * geom.nz * geom.nr)
        self.mc = np.zeros(max(geom.nmetal, 1))
        self.sg = np.zeros(geom.nz * geom.nr)

    def step(self, seed, step, counts=None):
        """One step; counts (int64, ns*ng*128) accumulates MCC events if given."""
        self.rho[:] = 0.0
        cp = counts.ctypes.data_as(_i64p) if counts is not None else None
        rc = lib(self.fast).ora_step(C.byref(self.geom.s), C.byref(self.setup.s), _d(self.ez), _d(self.er), seed,
                                     step, self.n.ctypes.data_as(_i64p), self.cap.ctypes.data_as(_i64p),
                                     C.cast(self.ptrs, C.c_void_p), _d(self.rho), _d(self.mc), _d(self.sg), cp)
        if rc != 0:
            raise RuntimeError("oracle step capacity overflow")
        return int(self.n.sum())

    def species(self, s):
        """Live particles of species s as (8, n) (a view)."""
        return self.bufs[s][:, :self.n[s]]
