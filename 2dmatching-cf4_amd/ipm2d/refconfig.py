"""The reference's run configuration -- config.json and the geometry and
materials inputs it names -- read into this package's host-side structures.

The reference loads ./config.json with json-fortran (R/src/onestep.F90:65-70)
and reads it in the Init*Json routines of R/src/base/init.F90.  This module
reads the same keys under the same rules:

* every key the reference reads must be there, or the load fails
  (CHECK_FOUND / FoundError, init.F90:24 and 880-890);
* a key is read only where the reference reads it: grid.specified_deps_in_*
  only with deps_type 1 (init.F90:172-193), gas.recb.* only with
  is_open_recombination (init.F90:832-878), a circuit's parameters only for
  its type, frequency lists only with n_freq > 0, pulse segments only with
  seg > 1, and only the first circuits_number entries of circuits_list
  (init.F90:574-800);
* paths are json-fortran's: dots between keys, 1-based [i] into arrays,
  case-sensitive keys, the first of duplicate keys;
* values convert between JSON types as json-fortran does without strict type
  checking (R/lib/json/json_value_module.F90: integers 8134-8179, reals
  8346-8400, logicals 8838-8874; a string becomes an integer through an Iw
  read, R/lib/json/json_string_utilities.F90:116-142): a real truncates to an
  integer, a logical is 1 or 0, a number is true when > 0, a string is true
  only as "true", null is NaN as a real and missing otherwise; a failed
  conversion counts as a missing key.

Derived quantities follow the reference:

* Nz, Nr are the geometry file's attributes; dz = dr = z_length / (Nz - 1)
  and r_length = dz (Nr - 1) (init.F90:105-135);
* with diag_freq > 0 the time step gives way to the diagnostic period:
  Period = diag_preiod, then dt = 1 / diag_freq / Period when Period > 0,
  else Period = int(1 / diag_freq / dt); with diag_freq <= 0 a zero Period is
  an error (init.F90:152-167);
* a bundle's base weight and array sizes (AllInitializationParticleBundle,
  R/src/pic/particle/ParticleBundle.f90:343-357);
* the geometry's material indices: the file holds 0-based indices into the
  materials list; metals become 1..nmetal and dielectrics -1..-ndiel in list
  order, vacuum 0, a dielectric's cells take its permittivity, and each
  metal cell marks its four nodes with its index, later cells in i-then-j
  order overwriting earlier ones (init.F90:316-398); node volumes follow
  init.F90:400-424 from the file's volume fractions;
* the species and gases of an lxcat gas set (GasInitPegasus,
  R/src/mc/ReactionOnePegasus.f90:521-600): the electron first, then each
  gas's species in its file's order, masses in amu times AtomicMass, charge
  numbers times ElectronCharge, gas density (p mTorrtoPa) / (kB T) times the
  gas's ratio; and each (species, gas) collision table from
  input/gas/lxcat/<gas>/<species><gas>.txt (MCCBundleInit,
  R/src/mc/MCCInit.F90:15-120).
"""
import dataclasses
import json
import math
import os
import re
from typing import List, Optional, Tuple

import numpy as np

from . import geometry as G
from .rezoning import RezoningLimits

INT_MAX = 2 ** 31 - 1

CIRCUIT_GROUND, CIRCUIT_VOLTAGE, CIRCUIT_PULSE, CIRCUIT_RLC, CIRCUIT_IMN = 1, 2, 3, 4, 5
DEPS_UNIFORM, DEPS_SPECIFY = 0, 1


class ConfigError(ValueError):
    """A key the reference reads is missing or does not convert (the
    reference stops with "The key is not exit in josn file.")."""


# ------------------------------------------------------------ json-fortran
_MISSING = object()
_PATH_TOKEN = re.compile(r"([^.\[\]]+)|\[(\d+)\]")


def _first_key_object(pairs):
    out = {}
    for k, v in pairs:
        if k not in out:
            out[k] = v
    return out


def _lookup(doc, path):
    node = doc
    pos = 0
    while pos < len(path):
        if path[pos] == ".":
            pos += 1
            continue
        m = _PATH_TOKEN.match(path, pos)
        if not m:
            raise ConfigError(f"bad path {path!r}")
        pos = m.end()
        if m.group(1) is not None:
            if not isinstance(node, dict) or m.group(1) not in node:
                return _MISSING
            node = node[m.group(1)]
        else:
            i = int(m.group(2))
            if not isinstance(node, list) or not 1 <= i <= len(node):
                return _MISSING
            node = node[i - 1]
    return node


def _to_int(v):
    if isinstance(v, bool):
        return 1 if v else 0
    if isinstance(v, int):
        return v
    if isinstance(v, float):
        if not math.isfinite(v) or abs(v) >= 2.0 ** 31:
            return _MISSING
        return int(v)                       # int(dbl_value, IK): toward zero
    if isinstance(v, str):
        t = v.replace(" ", "")               # an Iw read skips blanks
        if not re.fullmatch(r"[+-]?\d+", t):
            return _MISSING
        x = int(t)
        return x if -INT_MAX - 1 <= x <= INT_MAX else _MISSING
    return _MISSING


def _to_real(v):
    if isinstance(v, bool):
        return 1.0 if v else 0.0
    if isinstance(v, (int, float)):
        return float(v)
    if v is None:
        return float("nan")                  # null_to_real_mode 2, quiet NaN
    if isinstance(v, str):
        from .restart import _value
        try:
            return _value(v)
        except ValueError:
            return _MISSING
    return _MISSING


def _to_bool(v):
    if isinstance(v, bool):
        return v
    if isinstance(v, (int, float)):
        return v > 0
    if isinstance(v, str):
        return v == "true"
    return _MISSING


def _to_str(v):
    return v if isinstance(v, str) else _MISSING


class _Doc:
    def __init__(self, doc, source):
        self.doc, self.source = doc, source

    def _get(self, path, conv):
        v = _lookup(self.doc, path)
        if v is _MISSING:
            raise ConfigError(f"{self.source}: key {path!r} not found")
        x = conv(v)
        if x is _MISSING:
            raise ConfigError(f"{self.source}: key {path!r} does not convert ({v!r})")
        if conv is _to_int and not -INT_MAX - 1 <= x <= INT_MAX:
            raise ConfigError(f"{self.source}: key {path!r} exceeds a 4-byte integer")
        return x

    def int(self, path):
        return self._get(path, _to_int)

    def real(self, path):
        return self._get(path, _to_real)

    def bool(self, path):
        return self._get(path, _to_bool)

    def str(self, path):
        return self._get(path, _to_str)


def _load_json(path):
    with open(path) as f:
        return _Doc(json.load(f, object_pairs_hook=_first_key_object), str(path))


# ------------------------------------------------------------ structures
@dataclasses.dataclass
class Diag:
    diag_type: int
    cycles: int
    period: int              # diag_preiod as read
    freq: float
    nt: int


@dataclasses.dataclass
class Grid:
    z_length: float
    geometry: str            # path as resolved against the configuration's folder
    materials: str
    deps_z: int
    deps_r: int
    deps_type: int
    specified_z: Optional[List[float]] = None
    specified_r: Optional[List[float]] = None


@dataclasses.dataclass
class Plasma:
    init_density: float
    init_electron_temperature: float
    ppg: int
    alpha: float
    load_type: int
    particle_number_max_ratio: float
    particle_limit_max_velocity: float
    rezoning: RezoningLimits


@dataclasses.dataclass
class Gas:
    ng: int
    pressure: float
    temperature: float
    names: List[str]
    ratios: List[float]
    collision_section_type: int


@dataclasses.dataclass
class RecombinationSpec:
    enabled: bool
    time_step_number: int = 0
    reactions: List[Tuple[int, int, int]] = dataclasses.field(default_factory=list)
    rates: List[float] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class Circuit:
    """One entry of circuit.circuits_list.  ``index`` is its 1-based rank
    among the circuits of its type (EC%circuits_index, init.F90:619-622)."""
    type: int
    metal_index: int
    name: str
    index: int
    vdc: float = 0.0
    frequency: List[float] = dataclasses.field(default_factory=list)
    amplitude: List[float] = dataclasses.field(default_factory=list)
    phase: List[float] = dataclasses.field(default_factory=list)
    time: List[float] = dataclasses.field(default_factory=list)
    params: dict = dataclasses.field(default_factory=dict)


@dataclasses.dataclass
class Material:
    type: int
    name: str
    permittivity: Optional[float] = None
    permeability: Optional[float] = None
    conductivity: Optional[float] = None


@dataclasses.dataclass
class RunConfig:
    source: str
    nrun: int
    dt: float                # after the diagnostic-period rule
    dt_read: float           # control.dt as written
    period: int              # CF%Period
    is_restart: bool
    pic_type: int
    diag: Diag
    grid: Grid
    plasma: Plasma
    gas: Gas
    recombination: RecombinationSpec
    circuits: List[Circuit]
    nz: int
    nr: int

    @property
    def dz(self):
        return self.grid.z_length / float(self.nz - 1)

    @property
    def r_length(self):
        return self.grid.z_length / float(self.nz - 1) * float(self.nr - 1)

    def base_weight(self, ppg=None):
        """PB%Weight at load, ParticleBundle.f90:352."""
        ppg = self.plasma.ppg if ppg is None else int(ppg)
        nz, nr, d = float(self.nz - 1), float(self.nr - 1), self.dz
        return (self.plasma.init_density * G.PI * (nr * d) * (nr * d) * (nz * d)) / (nr * nz * float(ppg))

    def bundle_sizes(self, ppg=None):
        """(initial size, npar_max) of a bundle, ParticleBundle.f90:354-357."""
        ppg = self.plasma.ppg if ppg is None else int(ppg)
        cells = (self.nz - 1) * (self.nr - 1)
        return ppg * cells, int(math.ceil(self.plasma.particle_number_max_ratio * ppg * (self.nz - 1) * (self.nr - 1)))

    def cell_load_weight(self, ir, ppg=None):
        """Wq of a particle loaded in radial cell ir (1-based j), ParticleBundle.f90:368."""
        ppg = self.plasma.ppg if ppg is None else int(ppg)
        d = self.dz
        return self.plasma.init_density / float(ppg) * 2.0 * G.PI * d * d * d * float(ir)

    def rezoning(self, hp, ppg=None, fused_electrons=False):
        """A Rezoning control for the species of ``hp`` with the base weight
        every bundle starts from (ppg_dif(k) = ppg for every species,
        init.F90:552-556); fused_electrons folds the electrons' every-step
        split/merge into the fused step (Rezoning's docstring)."""
        from .rezoning import Rezoning
        ppg = self.plasma.ppg if ppg is None else int(ppg)
        w = self.base_weight(ppg)
        return Rezoning(hp, self.nz, self.nr, [w] * hp.ns, [ppg] * hp.ns, alpha=self.plasma.alpha,
                        period=self.period, limits=dataclasses.replace(self.plasma.rezoning),
                        fused_electrons=fused_electrons)

    def recombination_control(self, hp):
        from .recombination import Recombination
        r = self.recombination
        if not r.enabled:
            return Recombination(hp, np.zeros((0, 3), dtype=np.int64), [], 1, self.dt, enabled=False)
        return Recombination(hp, r.reactions, r.rates, r.time_step_number, self.dt, enabled=True)

    def imn(self, **kw):
        """circuit.IMN from the first IMN entry (its single frequency; the
        electrode capacitance and induced charge come from the field solve,
        keyword arguments)."""
        from .circuit import IMN
        for c in self.circuits:
            if c.type != CIRCUIT_IMN:
                continue
            if len(c.frequency) != 1:
                raise ValueError("the integrator drives one frequency; this IMN entry has "
                                 f"{len(c.frequency)}")
            p = c.params
            return IMN(Rs=p["Rs"], Cm1=p["Cm1"], Cm2=p["Cm2"], Lm=p["Lm"], Rm=p["Rm"],
                       stray=p["is_open_capacitance"], C_stray=p["C_stray"], L_stray=p["L_stray"],
                       R_stray=p["R_stray"], frequency=c.frequency[0], amplitude=c.amplitude[0],
                       phase_deg=c.phase[0], **kw)
        raise ValueError("no IMN circuit among the first circuits_number entries")


# ------------------------------------------------------------ reading
def _circuits(J):
    n = J.int("circuit.circuits_number")
    out = []
    counts = {}
    for i in range(1, n + 1):
        J.int(f"circuit.circuits_list[{i}].type")   # every type is read first, init.F90:590-598
    for i in range(1, n + 1):
        p = f"circuit.circuits_list[{i}]"
        t = J.int(p + ".type")
        metal = J.int(p + ".metal_index")
        name = J.str(p + ".name")
        if t not in (CIRCUIT_GROUND, CIRCUIT_VOLTAGE, CIRCUIT_PULSE, CIRCUIT_RLC, CIRCUIT_IMN):
            continue                                 # no case of the select matches
        counts[t] = counts.get(t, 0) + 1
        c = Circuit(t, metal, name, counts[t])
        if t == CIRCUIT_VOLTAGE:
            nf = J.int(p + ".n_freq")
            c.vdc = J.real(p + ".Vdc")
            for j in range(1, nf + 1):
                c.frequency.append(J.real(f"{p}.frequency[{j}]"))
                c.amplitude.append(J.real(f"{p}.amplitude[{j}]"))
                c.phase.append(J.real(f"{p}.phase[{j}]"))
        elif t == CIRCUIT_PULSE:
            seg = J.int(p + ".seg")
            if seg > 1:
                for j in range(1, seg + 1):
                    c.time.append(J.real(f"{p}.time[{j}]"))
                    c.amplitude.append(J.real(f"{p}.amplitude[{j}]"))
        elif t in (CIRCUIT_RLC, CIRCUIT_IMN):
            nf = J.int(p + ".n_freq")
            if nf > 0:
                for j in range(1, nf + 1):
                    c.frequency.append(J.real(f"{p}.frequency[{j}]"))
                    c.amplitude.append(J.real(f"{p}.amplitude[{j}]"))
                    c.phase.append(J.real(f"{p}.phase[{j}]"))
                if t == CIRCUIT_RLC:
                    keys = (("Rs", J.real), ("Cs", J.real), ("Ls", J.real))
                else:
                    keys = (("Rs", J.real), ("Cm1", J.real), ("Cm2", J.real), ("Lm", J.real), ("Rm", J.real),
                            ("is_open_capacitance", J.bool), ("C_stray", J.real), ("L_stray", J.real),
                            ("R_stray", J.real))
                for k, get in keys:
                    c.params[k] = get(f"{p}.{k}")
        out.append(c)
    return out


def read_materials(path):
    """The materials list of grid.materials (init.F90:316-385)."""
    J = _load_json(path)
    n = J.int("material_amount")
    if n <= 0:
        raise ConfigError(f"{path}: number of materials less than 0")
    out = []
    for i in range(1, n + 1):
        p = f"materials[{i}]"
        t = J.int(p + ".type")
        m = Material(t, J.str(p + ".name"))
        if t == G.MAT_DIEL:
            m.permittivity = J.real(p + ".permittivity")
            m.permeability = J.real(p + ".permeability")
            m.conductivity = J.real(p + ".conductivity")
        out.append(m)
    return out


def _resolve(base, p):
    return p if os.path.isabs(p) else os.path.normpath(os.path.join(base, p))


def load(path, nz=None, nr=None):
    """Read a config.json.  Relative paths in it are taken from the file's
    folder (the reference runs with that folder as its working directory).
    Nz, Nr come from the geometry file unless given."""
    J = _load_json(path)
    base = os.path.dirname(os.path.abspath(path))
    diag = Diag(J.int("diag.diag_type"), J.int("diag.diag_cycles"), J.int("diag.diag_preiod"),
                J.real("diag.diag_freq"), J.int("diag.diag_Nt"))
    nrun = J.int("control.nrun")
    dt = J.real("control.dt")
    z_length = J.real("grid.z_length")
    is_restart = J.bool("control.is_restart")
    load_type = J.int("plasma.load_type")
    pic_type = J.int("control.pic_type")
    geom_path = _resolve(base, J.str("grid.geometry"))
    if nz is None or nr is None:
        from . import h5lite
        a = h5lite.File(geom_path).attrs
        nz, nr = int(a["Nz"]), int(a["Nr"])
    deps = (J.int("grid.deps_in_z"), J.int("grid.deps_in_r"), J.int("grid.deps_type"))
    dt_read = dt
    period = diag.period
    if diag.freq > 0.0:
        if diag.period == 0:
            period = int(1.0 / diag.freq / dt)
        else:
            dt = 1.0 / diag.freq / float(period)
    if diag.freq <= 0.0 and period == 0:
        raise ConfigError(f"{path}: invalid diag period")
    grid = Grid(z_length, geom_path, "", *deps)
    if deps[2] == DEPS_SPECIFY:
        grid.specified_z = [J.real(f"grid.specified_deps_in_z[{i}]") for i in range(1, deps[0] + 1)]
        grid.specified_r = [J.real(f"grid.specified_deps_in_r[{i}]") for i in range(1, deps[1] + 1)]
    limits = RezoningLimits(J.bool("plasma.is_open_split_merge"), J.real("plasma.paritlc_upper_limit"),
                            J.real("plasma.paritlc_lower_limit"), J.real("plasma.min_mp_to_split_k"),
                            J.real("plasma.merge_limit"))
    grid.materials = _resolve(base, J.str("grid.materials"))
    ng = J.int("gas.ng")
    pressure, temperature = J.real("gas.pressure"), J.real("gas.temperature")
    te = J.real("plasma.init_electron_temperature")
    cst = J.int("gas.collision_section_type")
    if ng <= 0:
        raise ConfigError(f"{path}: the amount of gas must be greater than 0")
    names = [J.str(f"gas.name[{i}]") for i in range(1, ng + 1)]
    ratios = []
    for i in range(1, ng + 1):
        ratios.append(J.real(f"gas.ratio[{i}]"))
    gas = Gas(ng, pressure, temperature, names, ratios, cst)
    density, ppg = J.real("plasma.init_density"), J.int("plasma.ppg")
    max_ratio, max_v = J.real("plasma.particle_number_max_ratio"), J.real("plasma.particle_limit_max_velocity")
    plasma = Plasma(density, te, ppg, J.real("plasma.alpha"), load_type, max_ratio, max_v, limits)
    circuits = _circuits(J)
    recb = RecombinationSpec(J.bool("gas.is_open_recombination"))
    if recb.enabled:
        recb.time_step_number = J.int("gas.recb.recb_time_step_num")
        n = J.int("gas.recb.recb_reaction_num")
        for i in range(1, n + 1):
            r = tuple(J.int(f"gas.recb.recb_list[{i}][{k}]") for k in (1, 2, 3))
            recb.reactions.append(r)
            recb.rates.append(J.real(f"gas.recb.recb_rate[{i}]"))
    return RunConfig(str(path), nrun, dt, dt_read, period, is_restart, pic_type, diag, grid, plasma, gas,
                     recb, circuits, int(nz), int(nr))


def geometry_from_files(run, materials=None):
    """The reference's Geometry for ``run``: its geometry file and materials
    list, indices remapped and volumes formed as initGeometryAndMaterialsJson
    does (init.F90:216-430)."""
    from . import h5lite
    mats = read_materials(run.grid.materials) if materials is None else materials
    g = h5lite.read_geometry(run.grid.geometry)
    nz, nr = g["Nz"], g["Nr"]
    if (nz, nr) != (run.nz, run.nr):
        raise ValueError(f"geometry is {nz} x {nr}, configuration says {run.nz} x {run.nr}")
    raw = np.asarray(g["cell_material"], dtype=np.int64)
    remap = np.zeros(len(mats), dtype=np.int32)
    eps = np.array(g["epsilon_r"], dtype=np.float64, copy=True)
    ndiel = nmetal = 0
    for i, m in enumerate(mats):
        if m.type == G.MAT_DIEL:
            ndiel += 1
            remap[i] = -ndiel
            eps[raw == i] = m.permittivity
        elif m.type == G.MAT_METAL:
            nmetal += 1
            remap[i] = nmetal
    if raw.min() < 0 or raw.max() >= len(mats):
        raise ValueError("the geometry names a material the materials list does not hold")
    cm = remap[raw]
    mt = np.zeros(ndiel + nmetal + 1, dtype=np.int32)
    mt[:ndiel] = G.MAT_DIEL
    mt[ndiel] = G.MAT_VACUUM
    mt[ndiel + 1:] = G.MAT_METAL
    factor = np.asarray(g["volume"], dtype=np.float64)
    vol, inv = G.node_volumes(nz, nr, run.dz, factor)
    metal = mt[cm + ndiel] == G.MAT_METAL
    pt = G.metal_point_types(g["point_type"], cm, metal)
    return G.Geometry(nz=nz, nr=nr, dx=run.dz, cell_material=cm, material_type=mt, ndiel=ndiel, nmetal=nmetal,
                      hedge_type=np.asarray(g["hedge_type"], dtype=np.int32),
                      vedge_type=np.asarray(g["vedge_type"], dtype=np.int32), volume_factor=factor,
                      volume=vol, inv_volume=inv, cell_load=np.asarray(g["load"], dtype=np.int32),
                      cell_epsilon=eps, point_type=pt)


# ------------------------------------------------------------ gases and species
@dataclasses.dataclass
class SpeciesDef:
    """A species as SpecyOnePhysics2SpecyOne fills it (R/src/base/SpecyOne.F90:167-178)."""
    name: str
    natom: int
    charge: float            # C
    mass: float              # kg
    radius: float
    gas: int                 # 0 for the electron, else the 1-based gas


@dataclasses.dataclass
class GasDef:
    """A gas as GasPhysics2Gas and GasInitPegasus fill it (SpecyOne.F90:181-196;
    ReactionOnePegasus.f90:565-585)."""
    name: str
    mc_model: int
    natom: int
    ns: int
    mass: float              # kg
    radius: float
    beta_max: float
    temperature: float
    density: float           # m^-3
    index_start: int         # species of this gas are index_start + 1 .. index_start + ns
    cs_model: int = 1        # 1 lxcat, 2 Pegasus (GO%CSModel)
    sop_index: Optional[List[int]] = None   # Pegasus: the Index of each species line


_NML_ITEM = re.compile(r"GP%(\w+)(?:\((\d+)\))?\s*=", re.IGNORECASE)


def read_gas_physics(path):
    """The GasPhysicsfile namelist of input/gas/lxcat/<gas>/<gas>.txt, as
    LoadGasPhysicalLxCat reads it (ReactionOnePegasus.f90:497-517): a dict of
    GP%Name, MCModel, NAtom, Ns, Mass [amu], Radius, BetaMax, and 'SP', the
    species lines GP%SP(j) = name, natom, charge number, mass [amu], radius."""
    from .restart import _value
    with open(path, encoding="latin-1") as f:
        lines = [ln.split("!", 1)[0] for ln in f.read().splitlines()]
    text = " ".join(lines)
    a = text.lower().find("&gasphysicsfile")
    if a < 0:
        raise ConfigError(f"{path}: no &GasPhysicsfile group")
    text = text[a + len("&gasphysicsfile"):]
    end = text.find("/")                     # the group ends at the first '/'
    body = text[:end] if end >= 0 else text
    items = list(_NML_ITEM.finditer(body))
    out = {"SP": {}}
    for k, m in enumerate(items):
        stop = items[k + 1].start() if k + 1 < len(items) else len(body)
        toks = [t.strip() for t in body[m.end():stop].split(",")]
        toks = [t for t in toks if t]
        key = m.group(1).lower()
        if m.group(2) is not None:
            if key != "sp" or len(toks) < 5:
                raise ConfigError(f"{path}: bad GP%{m.group(1)}({m.group(2)}) entry")
            out["SP"][int(m.group(2))] = (toks[0], int(toks[1]), int(toks[2]), _value(toks[3]), _value(toks[4]))
        elif key == "name":
            out["Name"] = toks[0]
        elif key in ("mcmodel", "natom", "ns"):
            out[{"mcmodel": "MCModel", "natom": "NAtom", "ns": "Ns"}[key]] = int(toks[0])
        elif key in ("mass", "radius", "betamax"):
            out[{"mass": "Mass", "radius": "Radius", "betamax": "BetaMax"}[key]] = _value(toks[0])
    for k in ("Name", "NAtom", "Ns", "Mass"):
        if k not in out:
            raise ConfigError(f"{path}: GP%{k} missing")
    if sorted(out["SP"]) != list(range(1, out["Ns"] + 1)):
        raise ConfigError(f"{path}: GP%SP(1..{out['Ns']}) expected, found {sorted(out['SP'])}")
    return out


# ------------------------------------------------------------ Pegasus gas sets
class _Records:
    """Sequential formatted reads of a text file as Fortran does them: each
    READ starts on a new line, a list-directed READ takes its items from as
    many lines as it needs (blanks and commas separate them, '/' ends the
    list) and the rest of its last line is skipped."""

    def __init__(self, text, path):
        self.lines = text.splitlines()
        self.pos = 0
        self.path = path

    def skip_comments(self):
        """skip_comments (R/src/base/numerical/Numrical.F90:5-15): lines whose
        first character is '#'."""
        while self.pos < len(self.lines) and self.lines[self.pos][:1] == "#":
            self.pos += 1

    def read(self, n):
        toks = []
        while len(toks) < n:
            if self.pos >= len(self.lines):
                raise ConfigError(f"{self.path}: end of file after {len(toks)} of {n} items")
            line = self.lines[self.pos].split("/", 1)[0]
            self.pos += 1
            toks += [t for t in re.split(r"[\s,]+", line) if t]
        return toks[:n]


def read_gas_physics_pegasus(path):
    """input/gas/pegasus/<gas>/<gas>.txt as LoadGasPhysicalPegasus reads it
    (R/src/mc/ReactionOnePegasus.f90:450-486): name, MCModel, NAtom, the
    number of particle lines NSOP, then NSOP lines 'Index Mass Charge Radius
    Polarizability Name' (LoadSpecyOnePegasus, R/src/base/SpecyOne.F90:287-294),
    the first the gas itself.  The gas takes that first line's name, mass
    and radius (Angstrom, times 1e-10: ConvertSpecyOnePegasus2GasPhysical),
    Ns = NSOP - 1; species j its line's name, charge, mass and radius
    (NAtom keeps its default 1, SpecyOnePhysical); BetaMax is not read (0).
    Returns the dict read_gas_physics returns, plus 'SOP_index', the Index of
    each species line (the numbers the reaction files name particles by)."""
    with open(path, encoding="latin-1") as f:
        r = _Records(f.read(), path)
    name = r.read(1)[0]
    mc_model = int(r.read(1)[0])
    natom = int(r.read(1)[0])
    nsop = int(r.read(1)[0])
    sop = []
    for _ in range(nsop):
        t = r.read(6)
        sop.append((int(t[0]), _num(t[1]), _num(t[2]), _num(t[3]), _num(t[4]), t[5]))
    if nsop < 1:
        raise ConfigError(f"{path}: no particle lines")
    gas = sop[0]
    out = {"Name": gas[5], "FileName": name, "MCModel": mc_model, "NAtom": natom, "Ns": nsop - 1,
           "Mass": gas[1], "Radius": gas[3] * 1e-10, "BetaMax": 0.0, "SP": {},
           "SOP_index": [x[0] for x in sop[1:]]}
    for j, x in enumerate(sop[1:], start=1):
        out["SP"][j] = (x[5], 1, int(round(x[2])), x[1], x[3] * 1e-10)
    return out


def _num(tok):
    from .restart import _value
    return _value(tok)


def read_pegasus_reaction(path):
    """One reaction file of a Pegasus set as LoadReactionOnePegasus reads it
    (R/src/mc/ReactionOnePegasus.f90:105-145, LoadReactionOneBase 93-103):
    comment lines, Index, Name, comment lines, Threshold, NSpecy, SpecyIndex(1:NSpecy),
    NumberChange(1:NSpecy), ReactionType, Ncrosssection, then that many
    (energy [eV], cross section [cm^2]) lines.  Returns a dict."""
    with open(path, encoding="latin-1") as f:
        r = _Records(f.read(), path)
    r.skip_comments()
    index = int(r.read(1)[0])
    name = r.read(1)[0]
    r.skip_comments()
    th = _num(r.read(1)[0])
    nsp = int(r.read(1)[0])
    spidx = [int(t) for t in r.read(nsp)]
    nchg = [int(t) for t in r.read(nsp)]
    rtype = int(r.read(1)[0])
    ncs = int(r.read(1)[0])
    if ncs > 500:
        raise ConfigError(f"{path}: {ncs} cross-section points exceed NcrosssectionMax = 500")
    cs = np.array([[_num(t) for t in r.read(2)] for _ in range(ncs)], dtype=np.float64).reshape(ncs, 2)
    return {"index": index, "name": name, "threshold": th, "species": spidx, "change": nchg, "type": rtype,
            "energy": cs[:, 0].copy(), "sigma_cm2": cs[:, 1].copy()}


def pegasus_reaction_code(rx, sop_index, reactant):
    """(ReactionType, Resultant) of ConvertReactionOneBase2ReactionOne
    (R/src/mc/ReactionOnePegasus.f90:241-404) for a reaction read by
    read_pegasus_reaction, among the gas's species whose Index numbers are
    sop_index (species j = position j, 1-based; particle number 1 is the
    electron, 0; unknown numbers -1, ConvertReactionOneIndex 424-442).

    The source sets its electron flag in a loop over the first two
    particles that keeps only the second one's test; that is kept.  Where
    the source leaves Resultant unassigned (ion types 1 and 6, whose
    collisions make no product), the reactant's own index is used; reaction
    codes the source leaves unassigned are refused."""
    idx = []
    for t in rx["species"]:
        if t == 1:
            idx.append(0)
        else:
            idx.append(sop_index.index(t) + 1 if t in sop_index else -1)
    is_electron = len(idx) >= 2 and idx[1] == 0
    products = [k for k in idx[2:] if k > 0]
    n = len(products)
    rt = rx["type"]
    where = f"reaction {rx['index']} ({rx['name']})"
    if is_electron:
        if rt == 1:
            return 11, 0
        if rt == 100:
            return 12, 0
        if rt == 2:
            if n < 1:
                raise ConfigError(f"{where}: ionization without a product")
            return 13, products[0]
        if rt in (3, 4, 255):
            if n == 0:
                return 12, 0
            if n == 2:
                return 15, products[0] * 10 + products[1]
            if n == 1:
                if rt == 3:
                    return 14, products[0]
                ch = [rx["change"][k] for k in range(2) if idx[k] == 0]
                if not ch:
                    raise ConfigError(f"{where}: no electron among the first two particles")
                c = ch[-1]
                if rt == 255 and c == 1:
                    return 13, products[0]
                if c == 0:
                    return 12, (products[0] if rt == 4 else 0)
                if c == -1:
                    return 14, products[0]
        raise ConfigError(f"{where}: electron reaction type {rt} with {n} products is not converted by the source")
    if rt == 1:
        return 101, reactant
    if rt == 6:
        return 103, reactant
    if rt == 255 and n == 1:
        return 112, products[0]
    raise ConfigError(f"{where}: ion reaction type {rt} with {n} products is not converted here")


def pegasus_sigma_file(root, gas, species_name, species_index):
    """The SigmaFile SigmaNormalizationPegasus builds for one (species, gas)
    (R/src/mc/ReactionOnePegasus.f90:36-88): input/gas/pegasus/<gas>/<species><gas>.txt
    holds Model and NReaction, then one line per reaction naming its file
    (the type and threshold on that line are not used); cross sections are
    converted from cm^2 to m^2 (ConvertReactionOnePegasus2SigmaRaw, 443-449).
    None if the species file is missing."""
    from . import xsec
    d = os.path.join(root, "input", "gas", "pegasus", gas.name)
    path = os.path.join(d, f"{species_name}{gas.name}.txt")
    if not os.path.exists(path):
        return None
    with open(path, encoding="latin-1") as f:
        r = _Records(f.read(), path)
    model, nreaction = (int(t) for t in r.read(2))
    reactions = []
    for _ in range(nreaction):
        cs_index = r.read(3)[0]
        rx = read_pegasus_reaction(os.path.join(d, cs_index))
        rtype, res = pegasus_reaction_code(rx, gas.sop_index, species_index)
        reactions.append(xsec.Reaction(species_index, rtype, res, rx["threshold"], rx["energy"],
                                       rx["sigma_cm2"] / 1.0e4))
    return xsec.SigmaFile(model, reactions)


def _input_root(run):
    return os.path.dirname(os.path.abspath(run.source))


def gases_and_species(run, root=None):
    """(species, gases) of the run's gas list, GasInitPegasus order
    (R/src/mc/ReactionOnePegasus.f90:521-600): lxcat sets
    (collision_section_type 1) from input/gas/lxcat/<gas>/<gas>.txt,
    Pegasus sets (2) from input/gas/pegasus/<gas>/<gas>.txt."""
    from . import xsec
    if run.gas.collision_section_type not in (1, 2):
        raise NotImplementedError(f"collision_section_type {run.gas.collision_section_type}: only lxcat (1) and "
                                  f"Pegasus (2) gas sets exist")
    pegasus = run.gas.collision_section_type == 2
    root = _input_root(run) if root is None else root
    species = [SpeciesDef("Electron", 1, float(-1) * xsec.ELECTRON_CHARGE,
                          (xsec.ELECTRON_MASS / xsec.ATOMIC_MASS) * xsec.ATOMIC_MASS, 0.0, 0)]
    gases = []
    index = 0
    for g, (name, ratio) in enumerate(zip(run.gas.names, run.gas.ratios), start=1):
        if pegasus:
            gp = read_gas_physics_pegasus(os.path.join(root, "input", "gas", "pegasus", name, f"{name}.txt"))
        else:
            gp = read_gas_physics(os.path.join(root, "input", "gas", "lxcat", name, f"{name}.txt"))
        gases.append(GasDef(gp["Name"], gp.get("MCModel", 0), gp["NAtom"], gp["Ns"], gp["Mass"] * xsec.ATOMIC_MASS,
                            gp.get("Radius", 0.0), gp.get("BetaMax", 0.0), run.gas.temperature,
                            xsec.gas_density(run.gas.pressure, run.gas.temperature, ratio), index,
                            2 if pegasus else 1, gp.get("SOP_index")))
        for j in range(1, gp["Ns"] + 1):
            nm, natom, z, m_amu, radius = gp["SP"][j]
            species.append(SpeciesDef(nm, natom, float(z) * xsec.ELECTRON_CHARGE, m_amu * xsec.ATOMIC_MASS,
                                      radius, g))
            index += 1
    return species, gases


def mcc_tables(run, species, gases, root=None):
    """{(s, g): MCCBundle} for every species s and 0-based gas g, from
    input/gas/lxcat/<gas>/<species><gas>.txt, as MCCBundleElelctronInit and
    MCCBundleIonInit build them (R/src/mc/MCCInit.F90:43-120).  An ion
    against a gas it does not belong to, or without a file for its own gas,
    gets the hard-sphere table (ProbilityHardShpere, MCCInit.F90:107-116,
    172-270; xsec.hard_sphere_bundle).  The electron has no such fallback in
    the reference (its table would be read from an unset SigmaNormalized):
    a missing electron file is refused."""
    from . import xsec
    root = _input_root(run) if root is None else root
    out = {}
    for s, sp in enumerate(species):
        for g, gas in enumerate(gases):
            gd = {"mass": gas.mass, "temperature": gas.temperature, "density": gas.density, "natom": gas.natom,
                  "beta_max": gas.beta_max, "radius": gas.radius, "ns": gas.ns, "index_start": gas.index_start}
            if gas.cs_model == 2:
                # electron species index 0; reactions name products by
                # their position among the gas's species
                sf = pegasus_sigma_file(root, gas, sp.name, s - gas.index_start if s > 0 else 0) \
                    if (s == 0 or sp.gas == g + 1) else None
                path = os.path.join(root, "input", "gas", "pegasus", gas.name, f"{sp.name}{gas.name}.txt")
            else:
                path = os.path.join(root, "input", "gas", "lxcat", gas.name, f"{sp.name}{gas.name}.txt")
                sf = None
                if os.path.exists(path) and (s == 0 or sp.gas == g + 1):
                    with open(path, encoding="latin-1") as f:
                        sf = xsec.parse_sigma_file(f.read())
            if s > 0 and sf is None:
                out[(s, g)] = xsec.hard_sphere_bundle(s, sp.mass, sp.radius, gd, run.dt)
                continue
            if sf is None:
                raise NotImplementedError(f"{path} missing: the reference has no fallback table for the electron")
            out[(s, g)] = xsec.build_bundle(sf, sp.mass, sp.natom, gd, run.dt, index_start=gas.index_start)
    return out


# ------------------------------------------------------------ particles at load
C_EXP = 2.71828          # the constant `exp` of R/src/base/Constants.F90


def initial_particles(run, species, s, geom, rng, ppg=None):
    """Species s at load, (8, n) in grid units: the steps of
    AllInitializationParticleBundle (R/src/pic/particle/ParticleBundle.f90:
    362-385) with numpy's generator in place of RANDOM_NUMBER.

    Every loadable cell (i, j), 1-based, gets ppg candidates with R uniform in
    [j-1, j) and Z in [i-1, i) (PositionRandomInitializationParticleOne,
    ParticleOne.f90:35-48); a candidate is kept when a uniform is below R / j
    and then gets WQ = n0 / ppg * 2 pi dx^3 j, a speed from
    VelocityMaxwellianInitializationParticleOne (ParticleOne.f90:75-94: draw
    R1, R2 until R1^2 <= -e R2 ln R2, v = sqrt(-3 ln R2 kT / m)), an isotropic
    direction (ParticleOne.f90:107-121; the conversion to cylindrical
    components is the identity there), velocities times dt/dx and
    Az = Ar = 0.  The electron takes plasma.init_electron_temperature, every
    other species gas.temperature (GasInitPegasus)."""
    from . import xsec
    ppg = run.plasma.ppg if ppg is None else int(ppg)
    sp = species[s]
    temperature = run.plasma.init_electron_temperature if s == 0 else run.gas.temperature
    dx = run.dz
    load = np.asarray(geom.cell_load).astype(bool)
    ci, cj = np.nonzero(load)                      # 0-based cell (i-1, j-1)
    ci = np.repeat(ci, ppg).astype(np.float64)
    cj = np.repeat(cj, ppg).astype(np.float64)
    m = ci.size
    r = cj + rng.random(m)
    z = ci + rng.random(m)
    keep = rng.random(m) < r / (cj + 1.0)
    z, r, j = z[keep], r[keep], cj[keep] + 1.0
    n = z.size
    # rejection sampling of the speed, all candidates at once
    r2 = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        a = rng.random(todo.size)
        b = rng.random(todo.size)
        ok = a * a <= -C_EXP * b * np.log(b)
        r2[todo[ok]] = b[ok]
        todo = todo[~ok]
    beta = 1.0 / (xsec.KB * temperature)
    v = np.sqrt(-3.0 * np.log(r2) / beta / sp.mass)
    cos_t = 1.0 - 2.0 * rng.random(n)
    sin_t = np.sqrt(1.0 - cos_t * cos_t)
    phi = 2.0 * G.PI * rng.random(n)
    vres = 1.0 / (dx / run.dt)
    out = np.empty((8, n))
    out[0] = z
    out[1] = r
    out[2] = v * cos_t * vres
    out[3] = v * sin_t * np.cos(phi) * vres
    out[4] = v * sin_t * np.sin(phi) * vres
    out[5] = 0.0
    out[6] = 0.0
    out[7] = run.plasma.init_density / float(ppg) * 2.0 * G.PI * dx * dx * dx * j
    return out


def build_hotpath(run, root=None, seed=0, capacity=None, hotpath_cls=None):
    """A HotPath set up from the reference's inputs: the species of the gas
    set (charge, mass, array size npar_max unless ``capacity`` is given), the
    geometry rebuilt from its files, the collision tables, the particles at
    load (``initial_particles``, numpy's generator seeded with ``seed``) and
    the per-particle random streams.  The fields come from the caller's field
    solve (``set_fields`` / ``set_potential``).  Returns (hp, species, gases,
    geometry)."""
    from .hotpath import HotPath
    cls = HotPath if hotpath_cls is None else hotpath_cls
    species, gases = gases_and_species(run, root)
    tables = mcc_tables(run, species, gases, root)
    geom = geometry_from_files(run)
    cap = run.bundle_sizes()[1] if capacity is None else int(capacity)
    rng = np.random.default_rng(seed)
    parts = [initial_particles(run, species, s, geom, rng) for s in range(len(species))]
    hp = cls(run.nz, run.nr, run.dz, run.dt, len(species), len(gases))
    for s, sp in enumerate(species):
        if parts[s].shape[1] > cap:
            raise ValueError(f"species {sp.name}: {parts[s].shape[1]} particles at load exceed capacity {cap}")
        hp.set_species(s, sp.charge, sp.mass, cap)
    hp.set_geometry(geom)
    for (s, g), b in sorted(tables.items()):
        hp.set_mcc(s, g, b)
    for s in range(len(species)):
        hp.upload(s, parts[s])
    hp.set_rng(seed, 0)
    return hp, species, gases, geom
