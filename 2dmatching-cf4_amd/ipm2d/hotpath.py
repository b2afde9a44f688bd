"""Python host mirror of the reference's per-step particle calls.

``HotPath`` owns one device context (include/ipm_hip.h).  Its methods carry
the reference's names and argument meaning:

    move_es(s)                     PB(s)%MoveES(FG)            ParticleBundle.f90:649
    material_absorption(s)         MaterialAborptionParticle   ParticleMaterial.F90:10
    p2ces(s) -> RhoOne, ChiOne     FO(s)%P2CES(PB(s), Geom)    FieldOne.f90:132
    field_source_sum() -> Rho, Chi FT%Sum(FO)                  FieldSource.f90:72
    solver_system() -> Coe, Source FS%SetCoe/%SetSrc/%SetBound FieldSolver.f90:136-301
    mcc()                          MCC(Ns, Ng, PB, ...)        MCCPublic.F90:38
    one_step()                     OneStep's particle calls    onestep.F90:141-177
    moments(s)                     WeightingParticleField      diagParticleField.F90:173
    split_merge(s, ...)            split_particle_weight +     Rezoning.F90:53-92
                                   merge_particle_random
    recombine(reactions, ...)      RecombinationLocal%recb     Recombination.F90:164-273

Errors raise ``IPMError`` (``CapacityError`` where the reference would
``stop`` on a full particle bundle).  Grids are returned as numpy arrays
shaped (nz, nr) (Rho(iz, ir) at [iz-1, ir-1]); Chi keeps its ghost layers,
(nz+2, nr+2).
"""
import ctypes as C

import numpy as np

from . import _native as N
from .geometry import fortran

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class HotPath:
    def __init__(self, nz, nr, dx, dt, ns, ng=1):
        N.load()
        self.nz, self.nr, self.dx, self.dt, self.ns, self.ng = nz, nr, dx, dt, ns, ng
        h = C.c_void_p()
        N.call("ipm_create", C.byref(h), nz, nr, float(dx), float(dt), ns, ng)
        self._h = h
        self.nmetal = 0
        self._block, self._bpc, self._winr = 256, 0, 24
        self.capacity = [0] * ns
        self.species_params = [None] * ns     # (charge, mass) per species

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.load().ipm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --------------------------------------------------------------- set-up
    def set_species(self, s, charge, mass, capacity):
        N.call("ipm_set_species", self._h, s, float(charge), float(mass), int(capacity))
        self.capacity[s] = int(capacity)
        self.species_params[s] = (float(charge), float(mass))

    def set_geometry(self, geom):
        inv = fortran(geom.inv_volume).astype(np.float64)
        cm = fortran(geom.cell_material).astype(np.int32)
        he = fortran(geom.hedge_type).astype(np.int32)
        ve = fortran(geom.vedge_type).astype(np.int32)
        mt = np.ascontiguousarray(geom.material_type, dtype=np.int32)
        N.call("ipm_set_geometry", self._h, _d(inv), _i(cm), _i(he), _i(ve), geom.ndiel, geom.nmetal, _i(mt))
        self.nmetal = geom.nmetal
        if self.nz >= 3 and self.nr >= 3:
            self.set_solver_geometry(geom.cell_epsilon, geom.point_type)

    def set_solver_geometry(self, cell_epsilon, point_type):
        """Geom%cell_epsilon (nz-1, nr-1) and Geom%point_type (nz, nr) of the field solve."""
        eps = np.ascontiguousarray(fortran(np.asarray(cell_epsilon, dtype=np.float64)))
        pt = np.ascontiguousarray(fortran(np.asarray(point_type)).astype(np.int32))
        assert eps.size == (self.nz - 1) * (self.nr - 1) and pt.size == self.nz * self.nr
        N.call("ipm_set_solver_geometry", self._h, _d(eps), _i(pt))

    def solver_system(self, with_chi=True, sigma=None, bound=None, bound_mask=None):
        """FS%SetCoe (or %SetCoeWN when with_chi is False), %SetSrc and
        %SetBound (FieldSolver.f90:136-301) on the device, from the Rho and
        Chi of the last field_source_sum / one_step.  sigma = MT%sigma,
        bound / bound_mask = boundary%bound / %bound_mask, each (nz, nr) or
        None.  Returns ({"A".."E": (nz, nr)}, source (nz, nr))."""
        nn = self.nz * self.nr

        def host(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(fortran(np.asarray(a)).astype(dt))
            assert a.size == nn
            return a
        sg, bd, mk = host(sigma, np.float64), host(bound, np.float64), host(bound_mask, np.int32)
        coe = np.zeros(5 * nn)
        src = np.zeros(nn)
        N.call("ipm_solver_system", self._h, 1 if with_chi else 0, None if sg is None else _d(sg),
               None if bd is None else _d(bd), None if mk is None else _i(mk), _d(coe), _d(src))
        grids = coe.reshape(5, self.nr, self.nz).transpose(0, 2, 1)
        return dict(zip("ABCDE", grids)), src.reshape(self.nr, self.nz).T

    def set_mcc(self, s, g, b):
        """Upload one MCCBundle (ipm2d.xsec.MCCBundle or model 0 = none)."""
        if b is None or b.model == 0:
            z = np.zeros(1, dtype=np.int32)
            zd = np.zeros(1)
            N.call("ipm_set_mcc", self._h, s, g, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, _i(z), _i(z), _d(zd), _d(zd),
                   0.0, 0.0, 0.0)
            return
        rt = np.ascontiguousarray(b.rtype, dtype=np.int32)
        rs = np.ascontiguousarray(b.resultant, dtype=np.int32)
        th = np.ascontiguousarray(b.threshold, dtype=np.float64)
        pr = np.ascontiguousarray(b.prob, dtype=np.float64)        # (nsigma, nreaction) == Probility(j,i)
        assert pr.shape == (b.nsigma, b.nreaction)
        N.call("ipm_set_mcc", self._h, s, g, int(b.model), int(b.nreaction), int(b.nsigma), float(b.emin),
               float(b.einterval), float(b.emax), float(b.collision_ratio), _i(rt), _i(rs), _d(th), _d(pr),
               float(b.gas_mass), float(b.gas_temperature), float(b.beta_max))

    def set_rng(self, seed, step=0):
        N.call("ipm_set_rng", self._h, C.c_uint64(int(seed)), C.c_uint32(int(step)))

    def set_event_counting(self, on=True):
        N.call("ipm_set_event_counting", self._h, 1 if on else 0)

    def set_null_screen(self, on=True):
        """Null-event screen of the collision pass (on by default; exact)."""
        N.call("ipm_set_null_screen", self._h, 1 if on else 0)

    def set_launch(self, blocks_per_cu=0, block=None):
        """Grid of the fused pass (0: one block per tile; else persistent
        blocks per CU) and, if given, its threads per block (128/256/512)."""
        block = self._block if block is None else int(block)
        N.call("ipm_set_launch", self._h, block, int(blocks_per_cu))
        self._block, self._bpc = block, int(blocks_per_cu)

    def set_deposit_window(self, rows=24):
        N.call("ipm_set_deposit_window", self._h, int(rows))
        self._winr = int(rows)

    def set_tile_shape(self, block, rows):
        """Threads per block of the fused pass and its deposit window height
        (a pair ipm_set_launch and ipm_set_deposit_window accept together)."""
        self.set_deposit_window(24)
        self.set_launch(self._bpc, block)
        self.set_deposit_window(rows)

    @property
    def tile_shape(self):
        return self._block, self._winr

    def set_sort_tile(self, slots=2048):
        """Slots per tile of the cell sort (2048 or 4096; results unchanged)."""
        N.call("ipm_set_sort_tile", self._h, int(slots))
        self._sort_tile = int(slots)

    @property
    def sort_tile(self):
        return getattr(self, "_sort_tile", 2048)

    def set_collide_grid(self, blocks_per_cu=6):
        N.call("ipm_set_collide_grid", self._h, int(blocks_per_cu))

    # ------------------------------------------------------------ particles
    def upload(self, s, soa):
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        assert soa.ndim == 2 and soa.shape[0] == 8
        n = soa.shape[1]
        N.call("ipm_upload_species", self._h, s, n, _d(soa), N.LAYOUT_SOA)

    def upload_aos(self, s, aos):
        aos = np.ascontiguousarray(aos, dtype=np.float64)
        assert aos.ndim == 2 and aos.shape[1] == 8
        N.call("ipm_upload_species", self._h, s, aos.shape[0], _d(aos), N.LAYOUT_AOS)

    def upload_device(self, s, offset, n, ptr, ld):
        """Slots [offset, offset + n) of species s from device memory: 8 SoA
        arrays of leading dimension ld at device address ptr (a float64
        tensor's data_ptr(); the producer's stream must be synchronized)."""
        N.call("ipm_upload_species_device", self._h, s, int(offset), int(n), C.c_void_p(int(ptr)), int(ld))

    def count(self, s):
        n = C.c_int64()
        N.call("ipm_count", self._h, s, C.byref(n))
        return n.value

    def download(self, s):
        n = self.count(s)
        out = np.empty((8, max(n, 1)))
        got = C.c_int64()
        N.call("ipm_download_species", self._h, s, C.byref(got), _d(out), max(n, 1), N.LAYOUT_SOA)
        assert got.value == n
        return np.ascontiguousarray(out.reshape(-1)[:8 * n].reshape(8, n))

    def download_aos(self, s):
        n = self.count(s)
        out = np.empty((max(n, 1), 8))
        got = C.c_int64()
        N.call("ipm_download_species", self._h, s, C.byref(got), _d(out), max(n, 1), N.LAYOUT_AOS)
        return out[:got.value].copy()

    # --------------------------------------------------------------- fields
    def set_fields(self, ez, er):
        ez = np.ascontiguousarray(ez, dtype=np.float64).ravel()
        er = np.ascontiguousarray(er, dtype=np.float64).ravel()
        assert ez.size == (self.nz + 1) * self.nr and er.size == self.nz * (self.nr + 1)
        N.call("ipm_set_fields", self._h, _d(ez), _d(er))

    def set_potential(self, phi):
        """FT%SetE(FG) on the device from a (nz, nr) node potential."""
        ph = np.ascontiguousarray(fortran(np.asarray(phi, dtype=np.float64)))
        assert ph.size == self.nz * self.nr
        N.call("ipm_set_potential", self._h, _d(ph))

    # ------------------------------------------------- reference-named calls
    def move_es(self, s=-1):
        N.call("ipm_move_es", self._h, s)

    def material_absorption(self, s=-1):
        N.call("ipm_absorb", self._h, s)

    def move_absorb(self, s=-1):
        N.call("ipm_move_absorb", self._h, s)

    def take_tallies(self):
        mc = np.zeros(max(self.nmetal, 1))
        sg = np.zeros(self.nz * self.nr)
        N.call("ipm_take_tallies", self._h, _d(mc), _d(sg))
        return mc[:self.nmetal], sg.reshape(self.nr, self.nz).T

    def p2ces(self, s, copy=True):
        if not copy:
            N.call("ipm_p2c", self._h, s, None, None)
            return None
        rho = np.zeros(self.nz * self.nr)
        chi = np.zeros(self.nz * self.nr)
        N.call("ipm_p2c", self._h, s, _d(rho), _d(chi))
        return rho.reshape(self.nr, self.nz).T, chi.reshape(self.nr, self.nz).T

    def p2ces_all(self):
        rho = np.zeros(self.ns * self.nz * self.nr)
        chi = np.zeros(self.ns * self.nz * self.nr)
        N.call("ipm_p2c", self._h, -1, _d(rho), _d(chi))
        sh = (self.ns, self.nr, self.nz)
        return rho.reshape(sh).transpose(0, 2, 1), chi.reshape(sh).transpose(0, 2, 1)

    def field_source_sum(self, copy=True):
        if not copy:
            N.call("ipm_field_sum", self._h, None, None)
            return None
        rho = np.zeros(self.nz * self.nr)
        chi = np.zeros((self.nz + 2) * (self.nr + 2))
        N.call("ipm_field_sum", self._h, _d(rho), _d(chi))
        return rho.reshape(self.nr, self.nz).T, chi.reshape(self.nr + 2, self.nz + 2).T

    def moments(self, s, implicit_j=True):
        """WeightingParticleField for species s: dict of (nz, nr) arrays
        rho (RhoOne), energy (TOne, eV), mp (MPOne), jz, jr (JzOne, JrOne)."""
        out = {k: np.zeros(self.nz * self.nr) for k in ("rho", "energy", "mp", "jz", "jr")}
        N.call("ipm_moments", self._h, s, 1 if implicit_j else 0, *(_d(out[k]) for k in
                                                                    ("rho", "energy", "mp", "jz", "jr")))
        return {k: v.reshape(self.nr, self.nz).T for k, v in out.items()}

    def moments_noalloc(self, s, implicit_j, device_outputs):
        """ipm_moments_noalloc: the five moment grids of species s into device
        arrays (addresses of Nz*Nr float64 buffers in the order rho, energy,
        mp, jz, jr; None to skip).  Only enqueues: synchronize before reading."""
        ptrs = [C.c_void_p(int(p)) if p else None for p in device_outputs]
        N.call("ipm_moments_noalloc", self._h, s, 1 if implicit_j else 0, *ptrs)

    def split_merge(self, s, weight, alpha, grid_r, merge_limit=0.34, split=True, merge=True):
        """split_particle_weight then merge_particle_random on species s
        (Rezoning.F90:53-92) with CalRW's reference weight
        weight * ((1 - alpha) * R / grid_r + alpha).  Returns the counts
        (copies appended, merge candidates kept, merge candidates deleted)."""
        st = np.zeros(3, dtype=np.int64)
        mode = (1 if split else 0) | (2 if merge else 0)
        N.call("ipm_split_merge", self._h, s, mode, float(weight), float(alpha), float(grid_r), float(merge_limit),
               st.ctypes.data_as(C.POINTER(C.c_int64)))
        return tuple(int(x) for x in st)

    def set_step_rezone(self, s, weight=1.0, alpha=0.0, grid_r=1.0, merge_limit=0.34, split=True, merge=True):
        """split_merge of species s folded into every later one_step /
        one_step_graph (after the deposit, before MCC, as OneStep orders
        them); split=merge=False switches it off."""
        mode = (1 if split else 0) | (2 if merge else 0)
        N.call("ipm_set_step_rezone", self._h, s, mode, float(weight), float(alpha), float(grid_r),
               float(merge_limit))

    def recombine(self, reactions, rates, time_step):
        """executingRecombination2D (Recombination.F90:164-273).  reactions:
        (first, second, type) triples with 0-based species, type 0 constant
        rate, 1 temperature form; rates in m^3/s; time_step =
        recombination_time_step_number * dt.  Returns the macro-particles
        deleted per species."""
        rl = np.asarray(reactions, dtype=np.int32).reshape(-1, 3)
        flat = np.ascontiguousarray(rl.T.reshape(-1))       # recb_reaction(n, 3), Fortran order
        rt = np.ascontiguousarray(np.asarray(rates, dtype=np.float64).reshape(-1))
        if rt.size != rl.shape[0]:
            raise ValueError("one rate per reaction")
        out = np.zeros(self.ns, dtype=np.int64)
        N.call("ipm_recombine", self._h, rl.shape[0], _i(flat), _d(rt), float(time_step),
               out.ctypes.data_as(C.POINTER(C.c_int64)))
        return out

    def load_restart(self, s, path):
        """PB(s)%Load (ParticleBundle.f90:86-113) from a text (.dat) or HDF5
        (.h5) bundle dump: positions and velocities are brought to grid
        units if the dump is not scaled, the NPar live particles replace
        species s.  Returns the dump (header fields and particles)."""
        from . import restart
        if str(path).endswith(".h5"):
            from . import h5lite
            d = h5lite.read_bundle(path)
        else:
            d = restart.read_dat(path)
        d = d.to_grid_units()
        self.upload(s, d.particles())
        return d

    def dump_restart(self, s, path, weight):
        """PB(s)%Dump (ParticleBundle.f90:117-174) of the live particles of
        species s, in grid units, with PB%Weight = weight; .h5 writes HDF5,
        anything else the text form."""
        from . import restart
        charge, mass = self.species_params[s]
        d = restart.dump_from_particles(self.download(s), charge, mass, weight, self.dx, self.dt)
        if str(path).endswith(".h5"):
            from . import h5lite
            h5lite.write_bundle(path, d)
        else:
            restart.write_dat(path, d)
        return d

    def mcc(self):
        N.call("ipm_mcc", self._h)

    def take_mcc_counts(self):
        c = np.zeros(self.ns * self.ng * 128, dtype=np.int64)
        N.call("ipm_take_mcc_counts", self._h, c.ctypes.data_as(C.POINTER(C.c_int64)))
        return c.reshape(self.ns, self.ng, 128)

    def one_step(self):
        N.call("ipm_step", self._h)

    def one_step_graph(self):
        N.call("ipm_step_graph", self._h)

    def set_timing(self, on=True):
        N.call("ipm_set_timing", self._h, 1 if on else 0)

    def take_timing(self):
        """{'particle'|'collide'|'grid': (total_ms, launches)} since the last call."""
        ms = np.zeros(3)
        n = np.zeros(3, dtype=np.int64)
        N.call("ipm_take_timing", self._h, _d(ms), n.ctypes.data_as(C.POINTER(C.c_int64)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(("particle", "collide", "grid"))}

    def sort(self, s=-1):
        N.call("ipm_sort", self._h, s)

    def set_sort_interval(self, steps):
        """Sort every species after each `steps` one_step / one_step_graph
        calls (0: never); the graph form replays the sort as a graph too."""
        N.call("ipm_set_sort_interval", self._h, int(steps))

    def synchronize(self):
        N.call("ipm_synchronize", self._h)

    def device_ptrs(self, s):
        arr = (_dp * 8)()
        N.call("ipm_species_device_ptrs", self._h, s, arr)
        return [C.cast(arr[f], C.c_void_p).value for f in range(8)]


def from_config(cfg, geom=None, tables=None, capacity_factor=1.5, rng=None, particles=None, dielectric_ring=0):
    """Build a HotPath for a workload config: species, geometry, tables,
    particles (loaded with numpy's generator seeded by cfg.seed) and fields."""
    from . import workload as W
    geom = geom if geom is not None else W.make_geometry(cfg, dielectric_ring=dielectric_ring)
    tables = tables if tables is not None else W.build_tables(cfg)
    rng = rng if rng is not None else np.random.default_rng(cfg.seed)
    hp = HotPath(cfg.nz, cfg.nr, cfg.dx, cfg.dt, cfg.ns, 1)
    if particles is None:
        particles = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    for s in range(cfg.ns):
        q, m = W.species_params(cfg, s)
        cap = int(particles[s].shape[1] * capacity_factor) + 4096
        hp.set_species(s, q, m, cap)
    hp.set_geometry(geom)
    for s in range(cfg.ns):
        hp.set_mcc(s, 0, tables.get((s, 0)))
    for s in range(cfg.ns):
        hp.upload(s, particles[s])
    ez, er = W.fields(cfg, 0.0)
    hp.set_fields(ez, er)
    hp.set_rng(cfg.seed, 0)
    return hp, geom, tables, particles


def from_config_device(cfg, geom=None, tables=None, capacity_factor=1.05, seed=None):
    """from_config with the particles made on the GPU (workload.load_species_device,
    species s seeded by seed + s) and copied in device to device: for
    workloads too large to build on the host.  Capacities are the expected
    counts times capacity_factor plus a margin, at most IPM_CAPACITY_MAX.
    Returns (hp, geom, tables, counts)."""
    from . import workload as W
    geom = geom if geom is not None else W.make_geometry(cfg)
    tables = tables if tables is not None else W.build_tables(cfg)
    seed = cfg.seed if seed is None else seed
    hp = HotPath(cfg.nz, cfg.nr, cfg.dx, cfg.dt, cfg.ns, 1)
    for s in range(cfg.ns):
        q, m = W.species_params(cfg, s)
        mean = W.expected_count(cfg, geom, s)
        hp.set_species(s, q, m, min(int(mean * capacity_factor + 8.0 * np.sqrt(mean + 1.0)) + 4096, N.CAPACITY_MAX))
    hp.set_geometry(geom)
    for s in range(cfg.ns):
        hp.set_mcc(s, 0, tables.get((s, 0)))
    counts = [W.load_species_device(cfg, geom, s, hp, seed * 1000 + s) for s in range(cfg.ns)]
    import torch
    torch.cuda.empty_cache()          # the staging chunks
    ez, er = W.fields(cfg, 0.0)
    hp.set_fields(ez, er)
    hp.set_rng(cfg.seed, 0)
    return hp, geom, tables, counts
