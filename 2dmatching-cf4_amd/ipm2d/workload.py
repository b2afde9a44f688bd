"""Synthetic workloads for the particle hot path (CF4 capacitive discharges).

Each configuration names a grid, species, gas, drive and seed.  Particles are
loaded cell by cell in cell-sorted order (z fastest) with a fixed count per
cell and the cell-volume weight WQ = n * pi * dx^3 * (2j-1) / ppc, the
weighting AllInit uses (R/src/pic/particle/ParticleBundle.f90:352-384,
Wq proportional to j), so every species has a uniform (or profiled) density.
Velocities are Maxwellian at the species temperature; Az = Ar = 0.
"""
import math
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from . import geometry as geo
from . import xsec

AMU = xsec.ATOMIC_MASS
QE = xsec.ELECTRON_CHARGE
ME = xsec.ELECTRON_MASS
KB = xsec.KB
EV_TO_K = 11605.0          # R/src/base/Constants.F90:13


@dataclass
class SpeciesSpec:
    name: str
    charge: int                  # in units of e
    mass_amu: float
    natom: int
    ppc: float                   # macro-particles per cell
    density: float               # relative density (x n0)
    temperature_ev: float


@dataclass
class Config:
    name: str
    nz: int
    nr: int
    length: float                # L [m] (z extent); dx = L/(nz-1)
    dt: float
    species: List[SpeciesSpec]
    n0: float = 1.0e15           # reference density [m^-3]
    pressure_mtorr: float = 50.0
    gas_temperature: float = 300.0
    gas_mass_amu: float = 88.0043128     # CF4 (input/gas/lxcat/CF4/CF4.txt)
    gas_natom: int = 5
    beta_max: float = 3.0
    drive: List[Tuple[float, float]] = field(default_factory=lambda: [(13.56e6, 200.0)])
    profile: str = "uniform"     # or "cos2"
    steps: int = 100
    seed: int = 0
    sort_interval: int = 0       # 0: never

    @property
    def dx(self):
        return self.length / (self.nz - 1)

    @property
    def radius(self):
        return self.dx * (self.nr - 1)

    @property
    def ns(self):
        return len(self.species)


def _three(ppc_e, ppc_p, ppc_n, alpha):
    return [SpeciesSpec("e", -1, ME / AMU, 1, ppc_e, 1.0, 3.0),
            SpeciesSpec("CF3+", 1, 69.0059096, 4, ppc_p, 1.0 + alpha, 0.026),
            SpeciesSpec("F-", -1, 18.9984030, 1, ppc_n, alpha, 0.026)]


def _six(ppc_total_per_cell):
    ratios = [1.0, 0.8, 0.1, 0.05, 0.6, 0.1]
    tot = sum(ratios)
    ppc = [ppc_total_per_cell * r / tot for r in ratios]
    # densities chosen quasi-neutral: n+ = n_e + n_-
    return [SpeciesSpec("e", -1, ME / AMU, 1, ppc[0], 1.0, 3.0),
            SpeciesSpec("CF3+", 1, 69.0059096, 4, ppc[1], 1.3, 0.026),
            SpeciesSpec("CF2+", 1, 50.0075, 3, ppc[2], 0.3, 0.026),
            SpeciesSpec("CF+", 1, 31.0091, 2, ppc[3], 0.1, 0.026),
            SpeciesSpec("F-", -1, 18.9984030, 1, ppc[4], 0.6, 0.026),
            SpeciesSpec("CF3-", -1, 69.0059096, 4, ppc[5], 0.1, 0.026)]


def make_config(name, **over):
    """The five synthetic configurations of the measurement plan.

    c1: 65x129 nodes, L = 8 cm, 100 ppc x 3 species, uniform, 1000 steps, seed 0
    c2: 257x513 nodes, L = 20 cm, e 200 / CF3+ 110 / F- 72 ppc (~50.3 M), cos^2, seed 1
    c3: 257x513, six species 1:0.8:0.1:0.05:0.6:0.1, 5.0e7 total, seed 2
    c4: 129x257, 4 M particles (collisionless window), seed 3
    c5: 513x1025, L = 30 cm, six species, 1.0e9 total, 100 mTorr, dual frequency, seed 4
    """
    name = name.lower()
    if name == "c1":
        cfg = Config("c1", 65, 129, 0.08, 5e-11, _three(100, 100, 100, 0.5), steps=1000, seed=0)
    elif name == "c2":
        cfg = Config("c2", 257, 513, 0.20, 5e-11, _three(200, 110, 72, 0.5), profile="cos2", steps=200, seed=1,
                     sort_interval=20)
    elif name == "c3":
        cfg = Config("c3", 257, 513, 0.20, 5e-11, _six(5.0e7 / (256 * 512)), profile="cos2", steps=10000, seed=2,
                     sort_interval=20)
    elif name == "c4":
        ppc = 4.0e6 / 3 / (128 * 256)
        cfg = Config("c4", 129, 257, 0.20, 5e-11, _three(ppc, ppc, ppc, 0.5), profile="cos2", steps=200, seed=3)
    elif name == "c5":
        cfg = Config("c5", 513, 1025, 0.30, 5e-11, _six(1.0e9 / (512 * 1024)), profile="cos2", steps=2000, seed=4,
                     pressure_mtorr=100.0, drive=[(13.56e6, 300.0), (2.0e6, 500.0)], sort_interval=20)
    elif name == "tiny":
        cfg = Config("tiny", 17, 33, 0.02, 5e-11, _three(8, 8, 8, 0.5), steps=5, seed=7)
    else:
        raise ValueError(f"unknown config {name}")
    for k, v in over.items():
        setattr(cfg, k, v)
    return cfg


# ------------------------------------------------------------------ geometry
def make_geometry(cfg, dielectric_ring=0):
    return geo.ccp_geometry(cfg.nz, cfg.nr, cfg.dx, electrode_layers=1, dielectric_ring=dielectric_ring)


def _profile(cfg, zc, rc):
    if cfg.profile == "uniform":
        return np.ones_like(zc)
    L, R = cfg.length, cfg.radius
    return np.cos(math.pi * rc / (2.0 * R)) ** 2 * np.cos(math.pi * (zc / L - 0.5)) ** 2


def load_species(cfg, geom, s, rng, max_particles=None):
    """Particles of species s as an (8, n) SoA array in grid units, cell-sorted.

    Every vacuum cell (cell_load == 1) gets round(ppc) particles (the
    fractional part is placed with probability equal to it), positions
    uniform in z and linear in r within the cell (uniform in volume), weight
    WQ = n_s(cell centre) * pi * dx^3 * (2j-1) / ppc.
    """
    sp = cfg.species[s]
    dx = cfg.dx
    nzc, nrc = cfg.nz - 1, cfg.nr - 1
    iz, ir = np.meshgrid(np.arange(nzc), np.arange(nrc), indexing="ij")
    loadable = geom.cell_load.astype(bool)
    # cell order z fastest (matches the cell sort key)
    izf = iz.ravel(order="F")[loadable.ravel(order="F")]
    irf = ir.ravel(order="F")[loadable.ravel(order="F")]
    base = int(math.floor(sp.ppc))
    frac = sp.ppc - base
    counts = np.full(izf.size, base, dtype=np.int64)
    if frac > 0:
        counts += (rng.random(izf.size) < frac).astype(np.int64)
    n = int(counts.sum())
    if max_particles is not None and n > max_particles:
        raise ValueError(f"species {sp.name}: {n} particles exceed {max_particles}")
    cz = np.repeat(izf, counts)
    cr = np.repeat(irf, counts)
    z = cz + rng.random(n)
    # uniform in volume within [j-1, j]: r = sqrt(u*(j^2-(j-1)^2) + (j-1)^2)
    j = cr + 1.0
    r = np.sqrt(rng.random(n) * (j * j - (j - 1.0) ** 2) + (j - 1.0) ** 2)
    r = np.minimum(np.maximum(r, np.nextafter(cr.astype(np.float64), np.inf)), j)
    zc = (izf + 0.5) * dx
    rc = (irf + 0.5) * dx
    dens = cfg.n0 * sp.density * _profile(cfg, zc, rc)
    wcell = dens * math.pi * dx ** 3 * (2.0 * (irf + 1.0) - 1.0) / max(sp.ppc, 1e-30)
    wq = np.repeat(wcell, counts)
    mass = sp.mass_amu * AMU
    vth = math.sqrt(KB * sp.temperature_ev * EV_TO_K / mass)
    vf = dx / cfg.dt
    v = rng.standard_normal((3, n)) * (vth / vf)
    out = np.empty((8, n))
    out[0] = z
    out[1] = r
    out[2:5] = v
    out[5] = 0.0
    out[6] = 0.0
    out[7] = wq
    return out


def expected_count(cfg, geom, s):
    """Mean particle count load_species gives species s (for capacities)."""
    return int(geom.cell_load.astype(bool).sum()) * cfg.species[s].ppc


def load_species_device(cfg, geom, s, hp, seed, chunk_cells=1 << 15):
    """Species s loaded on the GPU straight into hp's device arrays, drawn
    from the distribution load_species uses (cell-sorted with z fastest,
    floor(ppc) particles per loadable cell plus one more with probability
    frac(ppc), uniform in volume inside the cell, the cell weight
    n_s(centre) pi dx^3 (2j-1)/ppc, Maxwellian velocities, Az = Ar = 0) with
    torch's device generator seeded by `seed`: the way to set up workloads
    too large to build on the host (1.0e9 particles).  Cells are filled in
    chunks of chunk_cells, each staged in a device tensor and copied in by
    ipm_upload_species_device.  Returns the particle count."""
    import torch
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    sp = cfg.species[s]
    dx = cfg.dx
    nzc = cfg.nz - 1
    load = torch.as_tensor(geom.cell_load.astype(bool).ravel(order="F"), device=dev)
    cells = torch.nonzero(load).squeeze(1)                 # z-fastest cell index
    base = int(math.floor(sp.ppc))
    frac = sp.ppc - base
    mass = sp.mass_amu * AMU
    vth = math.sqrt(KB * sp.temperature_ev * EV_TO_K / mass) / (dx / cfg.dt)
    L, R = cfg.length, cfg.radius
    offset = 0
    for c0 in range(0, cells.numel(), chunk_cells):
        ci = cells[c0:c0 + chunk_cells]
        iz = ci % nzc
        ir = ci // nzc
        counts = torch.full((ci.numel(),), base, dtype=torch.int64, device=dev)
        if frac > 0:
            counts += (torch.rand(ci.numel(), generator=gen, device=dev, dtype=torch.float64) < frac).to(torch.int64)
        n = int(counts.sum().item())
        if n == 0:
            continue
        cz = torch.repeat_interleave(iz, counts).to(torch.float64)
        cr = torch.repeat_interleave(ir, counts).to(torch.float64)
        zc = (iz.to(torch.float64) + 0.5) * dx
        rc = (ir.to(torch.float64) + 0.5) * dx
        if cfg.profile == "uniform":
            prof = torch.ones_like(zc)
        else:
            prof = torch.cos(math.pi * rc / (2.0 * R)) ** 2 * torch.cos(math.pi * (zc / L - 0.5)) ** 2
        wcell = cfg.n0 * sp.density * prof * math.pi * dx ** 3 * (2.0 * (ir.to(torch.float64) + 1.0) - 1.0) \
            / max(sp.ppc, 1e-30)
        buf = torch.empty((8, n), dtype=torch.float64, device=dev)
        buf[0] = cz + torch.rand(n, generator=gen, device=dev, dtype=torch.float64)
        j = cr + 1.0
        r = torch.sqrt(torch.rand(n, generator=gen, device=dev, dtype=torch.float64) * (j * j - (j - 1.0) ** 2)
                       + (j - 1.0) ** 2)
        buf[1] = torch.minimum(torch.maximum(r, torch.nextafter(cr, torch.full_like(cr, math.inf))), j)
        buf[2:5] = torch.randn((3, n), generator=gen, device=dev, dtype=torch.float64) * vth
        buf[5:7] = 0.0
        buf[7] = torch.repeat_interleave(wcell, counts)
        torch.cuda.synchronize()
        hp.upload_device(s, offset, n, buf.data_ptr(), n)
        offset += n
        del buf, cz, cr, r, j
    return offset


def species_params(cfg, s):
    sp = cfg.species[s]
    return sp.charge * QE, sp.mass_amu * AMU


# --------------------------------------------------------------------- field
def potential(cfg, t, plasma_potential=20.0):
    """Synthetic node potential: the electrode drive sum V_k sin(2 pi f_k t) on
    z = 0 falling linearly to the grounded z = L, plus a cos^2 plasma bump."""
    z = np.arange(cfg.nz)[:, None] * cfg.dx
    r = np.arange(cfg.nr)[None, :] * cfg.dx
    v = sum(V * math.sin(2.0 * math.pi * f * t) for f, V in cfg.drive)
    phi = v * (1.0 - z / cfg.length) + plasma_potential * (
        np.cos(math.pi * r / (2.0 * cfg.radius)) ** 2 * np.sin(math.pi * z / cfg.length) ** 2)
    return phi + 0.0 * r


def fields(cfg, t=0.0):
    return geo.set_e_from_phi(potential(cfg, t), cfg.dx)


# -------------------------------------------------------------------- tables
def gas_dict(cfg):
    return dict(mass=cfg.gas_mass_amu * AMU, temperature=cfg.gas_temperature,
                density=xsec.gas_density(cfg.pressure_mtorr, cfg.gas_temperature),
                natom=float(cfg.gas_natom), beta_max=cfg.beta_max)


def _ion_model2(elastic, cx, kind=103, n=xsec.NSIGMA_RAW_MAX):
    e = xsec.log_grid(0.01, 100.0, n)
    return [xsec.Reaction(0, 101, 0, 0.0, e.copy(), np.full(n, elastic)),
            xsec.Reaction(0, kind, 0, 0.0, e.copy(), cx * (1.0 + 0.0 * e))]


def sigma_files(cfg):
    """Per-species parsed cross-section files (SigmaFile) for the config."""
    names = [sp.name for sp in cfg.species]
    idx = {nm: k for k, nm in enumerate(names)}
    out = {}
    if len(names) == 3:
        out[0] = xsec.SigmaFile(1, xsec.synthetic_electron_reactions(ion_product=idx["CF3+"], att_product=idx["F-"]))
        # reactive (Langevin) ion sets with thresholds of the CF4 ion files
        out[1] = xsec.SigmaFile(3, [xsec.Reaction(1, 101, 1, 0.0), xsec.Reaction(1, 111, 1, 5.621),
                                    xsec.Reaction(1, 100, 0, 5.843)])
        out[2] = xsec.SigmaFile(3, [xsec.Reaction(2, 101, 2, 0.0), xsec.Reaction(2, 113, 0, 3.521),
                                    xsec.Reaction(2, 111, 2, 5.621)])
    else:
        e = xsec.synthetic_electron_reactions(ion_product=idx["CF3+"], att_product=idx["F-"],
                                              excitation=(7.0, 2e-20),
                                              dissociation=(12.0, 1e-20, idx["F-"] * 10 + idx["CF3+"]))
        # dissociative attachment into CF3- (Gaussian at 7 eV)
        ee = xsec.log_grid()
        e.append(xsec.Reaction(0, 14, idx["CF3-"], 0.0, ee.copy(), 2e-21 * np.exp(-0.5 * ((ee - 7.0) / 1.0) ** 2)))
        out[0] = xsec.SigmaFile(1, e)
        out[idx["CF3+"]] = xsec.SigmaFile(2, _ion_model2(4e-19, 2e-19, 103))
        out[idx["CF2+"]] = xsec.SigmaFile(2, _ion_model2(4e-19, 1e-19, 103))
        out[idx["CF+"]] = xsec.SigmaFile(2, _ion_model2(4e-19, 1e-19, 102))
        out[idx["F-"]] = xsec.SigmaFile(3, [xsec.Reaction(0, 101, 0, 0.0), xsec.Reaction(0, 113, 0, 3.521),
                                            xsec.Reaction(0, 111, idx["F-"], 5.621)])
        out[idx["CF3-"]] = xsec.SigmaFile(3, [xsec.Reaction(0, 101, 0, 0.0), xsec.Reaction(0, 112, idx["F-"], 1.927),
                                              xsec.Reaction(0, 113, 0, 5.448)])
        for k in (idx["CF3+"], idx["CF2+"], idx["CF+"]):
            for r in out[k].reactions:
                r.resultant = k
    return out


def build_tables(cfg):
    """MCCBundle per (species, gas=0), as MCCBundleInit builds them."""
    gas = gas_dict(cfg)
    files = sigma_files(cfg)
    tabs = {}
    for s, sf in files.items():
        sp = cfg.species[s]
        tabs[(s, 0)] = xsec.build_bundle(sf, sp.mass_amu * AMU, float(sp.natom), gas, cfg.dt)
    return tabs
