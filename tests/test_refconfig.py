"""The reference's run configuration (ipm2d.refconfig), CPU checks:

* the reference's own config.json and the inputs it names (skipped where the
  reference tree is absent; nothing of it is copied here):
  - Nz, Nr come from the geometry file and dt from the diagnostic period
    (1 / 13.56 MHz / 150, not the 5e-10 written under control.dt), as
    InitControlFlowJson does (R/src/base/init.F90:105-167);
  - only the first circuits_number = 3 entries of the seven-entry
    circuits_list are taken, each with its type's keys;
  - the string "1" under plasma.load_type reads as the integer 1;
  - the geometry rebuilt from geom.h5 and geom.json agrees with what the
    file already holds: edge types follow the materials' types, the
    dielectrics' permittivities equal the file's epsilon_r, metal nodes are
    exactly the nodes the file marks as metal, and the node volumes are
    positive exactly where the file's volume fractions exceed 1e-6;
  - the base weight times the particles at load equals the initial density
    times the cylinder's volume;
  - the CF4 gas set gives the electron, F-, CF3+, CF3- in that order (the
    configuration's recombination list, written against the same indices,
    pairs opposite charges and gives its temperature-dependent reaction to
    the electron and a positive ion); the four collision tables build, the
    electron's normalised cumulative table peaks at exactly 1, its
    selection ratio is 1 - exp(-nu_max dt), and the reactive ratio, which
    depends on the reduced mass only, is equal for CF3+ and CF3-;
* json-fortran's reading rules on synthetic configurations: type
  conversions, 1-based paths, first of duplicate keys, keys read only where
  the reference reads them, missing keys, the diagnostic-period branches;
* the metal node marking against the reference's loop (i outer, j inner,
  later cells overwrite) on small random maps where metals touch;
* the GasPhysicsfile namelist: comments, d/D exponents, a missing species
  line;
* particles at load (AllInitializationParticleBundle's steps): only loadable
  cells, the R / j acceptance (count per row), a uniform density (total
  weight = n0 x loadable volume), and the speed sampler's closed form: its
  accepted R2 make the energy Maxwell-Boltzmann, mean 3/2 kT, isotropic;
* build_hotpath's calls on a recording stand-in, from the reference's own
  inputs.
"""
import copy
import json
import math
import os

import numpy as np
import pytest

from ipm2d import geometry as G
from ipm2d import h5lite, refconfig as RC

REF = "/root/reference/iPM2D-multi-EI/config.json"
needs_ref = pytest.mark.skipif(not os.path.exists(REF), reason="reference tree absent")

BASE = {
    "control": {"nrun": 10, "dt": 1e-10, "is_restart": False, "pic_type": 1},
    "diag": {"diag_type": 1, "diag_cycles": 1, "diag_preiod": 20, "diag_freq": 0.0, "diag_Nt": 5},
    "grid": {"z_length": 0.02, "deps_in_z": 1, "deps_in_r": 1, "deps_type": 0,
             "materials": "m.json", "geometry": "g.h5"},
    "plasma": {"init_density": 1e15, "init_electron_temperature": 3.0, "ppg": 8, "alpha": 0.05,
               "is_open_split_merge": True, "paritlc_upper_limit": 1.02, "paritlc_lower_limit": 0.98,
               "min_mp_to_split_k": 0.2, "merge_limit": 0.3, "load_type": 1,
               "particle_number_max_ratio": 4.0, "particle_limit_max_velocity": 10.0},
    "gas": {"ng": 1, "pressure": 50, "temperature": 300, "name": ["CF4"], "ratio": [1.0],
            "collision_section_type": 1, "is_open_recombination": False},
    "circuit": {"circuits_number": 1, "circuits_list": [{"type": 1, "metal_index": 2, "name": "ground"}]},
}


def _write(tmp_path, doc, text=None):
    p = tmp_path / "config.json"
    p.write_text(text if text is not None else json.dumps(doc))
    return p


def _load(tmp_path, doc=None, text=None):
    return RC.load(_write(tmp_path, doc, text), nz=9, nr=5)


# ------------------------------------------------------------ the reference's input
@needs_ref
def test_reference_config_derived_quantities():
    run = RC.load(REF)
    assert (run.nz, run.nr) == (201, 201)
    assert run.dt_read == 5e-10
    assert run.period == 150 and run.dt == 1.0 / 13.56e6 / 150.0
    assert run.dz == 0.1054 / 200.0 and run.r_length == run.dz * 200.0
    assert run.plasma.load_type == 1 and run.plasma.ppg == 50
    assert [(c.type, c.metal_index, c.index) for c in run.circuits] == [(5, 1, 1), (4, 2, 1), (1, 3, 1)]
    rlc = run.circuits[1]
    assert rlc.params == {"Rs": 0.5, "Cs": 5e-09, "Ls": 1e-08} and rlc.amplitude == [50.0]
    imn = run.imn()
    assert (imn.Rs, imn.Cm1, imn.Cm2, imn.Lm, imn.Rm) == (50.0, 1.5e-10, 2e-10, 4.3e-6, 0.5)
    assert imn.stray and (imn.C_stray, imn.L_stray, imn.R_stray) == (2.4e-11, 1e-8, 0.5)
    assert (imn.frequency, imn.amplitude, imn.phase_deg) == (13.56e6, 100.0, 0.0)
    r = run.recombination
    assert r.enabled and r.time_step_number == 50
    assert r.reactions == [(1, 2, 0), (2, 3, 0), (0, 2, 1)] and r.rates == [1e-13, 1e-13, 3.95e-15]
    lim = run.plasma.rezoning
    assert (lim.is_open_split_merge, lim.paritlc_upper_limit, lim.paritlc_lower_limit,
            lim.min_mp_to_split_k, lim.merge_limit) == (True, 1.01, 0.99, 0.1, 0.34)
    assert run.gas.names == ["CF4"] and run.gas.pressure == 100.0 and run.gas.temperature == 300.0
    # total weight at load = density x cylinder volume
    n0, _ = run.bundle_sizes()
    vol = math.pi * run.r_length ** 2 * run.grid.z_length
    assert run.base_weight() * n0 == pytest.approx(run.plasma.init_density * vol, rel=1e-14)
    assert run.bundle_sizes()[1] == 100 * n0


@needs_ref
def test_reference_geometry_rebuilt_from_its_inputs():
    run = RC.load(REF)
    mats = RC.read_materials(run.grid.materials)
    assert [m.type for m in mats] == [0, 1, 1, 1, 2, 2]
    g = RC.geometry_from_files(run, mats)
    f = h5lite.read_geometry(run.grid.geometry)
    assert (g.nmetal, g.ndiel) == (3, 2)
    assert set(np.unique(g.cell_material)) == {-2, -1, 0, 1, 2, 3}
    # metals 1..3 and dielectrics -1..-2 in list order
    for i, want in enumerate([0, 1, 2, 3, -1, -2]):
        assert np.all(g.cell_material[f["cell_material"] == i] == want)
    ctype = np.array([m.type for m in mats])[f["cell_material"]]
    he, ve = G._edge_types(ctype)
    np.testing.assert_array_equal(g.hedge_type, he)
    np.testing.assert_array_equal(g.vedge_type, ve)
    np.testing.assert_array_equal(g.cell_epsilon, f["epsilon_r"])
    np.testing.assert_array_equal(g.point_type > 0, f["point_type"] == 1)
    np.testing.assert_array_equal(g.point_type[g.point_type <= 0], f["point_type"][g.point_type <= 0])
    assert np.array_equal(g.inv_volume > 0, f["volume"] > 1e-6)
    assert np.all(g.mtype(g.cell_material)[f["load"] == 1] == G.MAT_VACUUM)


# ------------------------------------------------------------ json-fortran rules
def test_type_conversions(tmp_path):
    d = copy.deepcopy(BASE)
    d["plasma"]["load_type"] = " 1 2 "           # Iw read: blanks skipped
    d["plasma"]["ppg"] = 8.9                     # real -> integer toward zero
    d["control"]["nrun"] = True                  # logical -> 1
    d["control"]["is_restart"] = 2               # number > 0 -> true
    d["plasma"]["is_open_split_merge"] = "True"  # only "true" is true
    d["gas"]["pressure"] = "1.5D+01"             # string -> real
    d["plasma"]["alpha"] = None                  # null -> NaN
    run = _load(tmp_path, d)
    assert run.plasma.load_type == 12 and run.plasma.ppg == 8 and run.nrun == 1
    assert run.is_restart is True and run.plasma.rezoning.is_open_split_merge is False
    assert run.gas.pressure == 15.0 and math.isnan(run.plasma.alpha)
    d["plasma"]["ppg"] = -8.9
    assert _load(tmp_path, d).plasma.ppg == -8
    for bad in ("1.5", "x", ""):
        d["plasma"]["load_type"] = bad
        with pytest.raises(RC.ConfigError):
            _load(tmp_path, d)
    d["plasma"]["load_type"] = 2 ** 31
    with pytest.raises(RC.ConfigError):
        _load(tmp_path, d)


def test_first_duplicate_key_wins_and_keys_are_case_sensitive(tmp_path):
    text = json.dumps(BASE).replace('"ppg": 8', '"ppg": 8, "ppg": 99, "PPG": 7')
    assert _load(tmp_path, text=text).plasma.ppg == 8
    d = copy.deepcopy(BASE)
    d["plasma"]["PPG"] = d["plasma"].pop("ppg")
    with pytest.raises(RC.ConfigError, match="plasma.ppg"):
        _load(tmp_path, d)


def test_missing_keys_fail_and_conditional_keys_are_optional(tmp_path):
    d = copy.deepcopy(BASE)
    run = _load(tmp_path, d)                         # no recb, no specified deps: fine
    assert not run.recombination.enabled and run.grid.specified_z is None
    d["grid"]["deps_type"] = 1
    d["grid"]["deps_in_z"] = 2
    with pytest.raises(RC.ConfigError, match="specified_deps_in_z"):
        _load(tmp_path, d)
    d["grid"]["specified_deps_in_z"] = [1, 3]
    d["grid"]["specified_deps_in_r"] = [2]
    run = _load(tmp_path, d)
    assert run.grid.specified_z == [1.0, 3.0] and run.grid.specified_r == [2.0]
    d["gas"]["is_open_recombination"] = True
    with pytest.raises(RC.ConfigError, match="recb_time_step_num"):
        _load(tmp_path, d)
    d["gas"]["recb"] = {"recb_time_step_num": 5, "recb_reaction_num": 2,
                        "recb_list": [[1, 2, 0], [0, 1, 1]], "recb_rate": [1e-13]}
    with pytest.raises(RC.ConfigError, match=r"recb_rate\[2\]"):
        _load(tmp_path, d)
    d["gas"]["recb"]["recb_rate"].append(2e-14)
    r = _load(tmp_path, d).recombination
    assert r.reactions == [(1, 2, 0), (0, 1, 1)] and r.rates == [1e-13, 2e-14]
    for sec, key in (("gas", "temperature"), ("control", "dt"), ("diag", "diag_Nt")):
        e = copy.deepcopy(BASE)
        del e[sec][key]
        with pytest.raises(RC.ConfigError, match=f"{sec}.{key}"):
            _load(tmp_path, e)
    e = copy.deepcopy(BASE)
    e["gas"]["ng"] = 0
    with pytest.raises(RC.ConfigError):
        _load(tmp_path, e)


def test_circuits_read_as_the_reference_reads_them(tmp_path):
    d = copy.deepcopy(BASE)
    d["circuit"] = {"circuits_number": 5, "circuits_list": [
        {"type": 2, "metal_index": 1, "name": "v", "Vdc": 3, "n_freq": 2, "frequency": [1e6, 2e6],
         "amplitude": [10, 20], "phase": [0, 90]},
        {"type": 7, "metal_index": 4, "name": "unknown type"},
        {"type": 3, "metal_index": 2, "name": "pulse one segment", "seg": 1},
        {"type": 3, "metal_index": 2, "name": "pulse", "seg": 2, "time": [0, 1e-8], "amplitude": [0, 5]},
        {"type": 4, "metal_index": 3, "name": "rlc without frequencies", "n_freq": 0},
        {"type": 5, "metal_index": 1, "name": "beyond circuits_number"},
    ]}
    c = _load(tmp_path, d).circuits
    assert [(x.type, x.index, x.name) for x in c] == [
        (2, 1, "v"), (3, 1, "pulse one segment"), (3, 2, "pulse"), (4, 1, "rlc without frequencies")]
    assert c[0].vdc == 3.0 and c[0].frequency == [1e6, 2e6] and c[0].phase == [0.0, 90.0]
    assert c[1].time == [] and c[2].time == [0.0, 1e-8] and c[2].amplitude == [0.0, 5.0]
    assert c[3].params == {} and c[3].frequency == []
    with pytest.raises(ValueError, match="no IMN"):
        _load(tmp_path, d).imn()
    d["circuit"]["circuits_number"] = 6
    with pytest.raises(RC.ConfigError, match=r"circuits_list\[6\].n_freq"):
        _load(tmp_path, d)
    d["circuit"]["circuits_number"] = 7                    # every type is read before any entry
    with pytest.raises(RC.ConfigError, match=r"circuits_list\[7\].type"):
        _load(tmp_path, d)


def test_diagnostic_period_branches(tmp_path):
    d = copy.deepcopy(BASE)
    run = _load(tmp_path, d)                                # diag_freq 0: dt kept
    assert run.dt == 1e-10 and run.period == 20
    d["diag"]["diag_freq"] = 13.56e6                        # period given: dt replaced
    run = _load(tmp_path, d)
    assert run.period == 20 and run.dt == 1.0 / 13.56e6 / 20.0
    d["diag"]["diag_preiod"] = 0                            # period from dt
    run = _load(tmp_path, d)
    assert run.dt == 1e-10 and run.period == int(1.0 / 13.56e6 / 1e-10) == 737
    d["diag"]["diag_freq"] = 0.0
    with pytest.raises(RC.ConfigError, match="diag period"):
        _load(tmp_path, d)


def test_relative_paths_follow_the_config_folder(tmp_path):
    run = _load(tmp_path, BASE)
    assert run.grid.geometry == str(tmp_path / "g.h5") and run.grid.materials == str(tmp_path / "m.json")


# ------------------------------------------------------------ metal node marking
def _reference_marking(pt, cm, metal):
    out = np.array(pt, copy=True)
    for i in range(cm.shape[0]):
        for j in range(cm.shape[1]):
            if metal[i, j]:
                for a in (0, 1):
                    for b in (0, 1):
                        out[i + a, j + b] = cm[i, j]
    return out


def test_metal_marking_matches_the_loop_order(rng):
    for _ in range(50):
        cm = rng.integers(-1, 4, size=(7, 6)).astype(np.int32)
        metal = cm > 0
        pt = rng.integers(-2, 2, size=(8, 7)).astype(np.int32)
        np.testing.assert_array_equal(G.metal_point_types(pt, cm, metal), _reference_marking(pt, cm, metal))
    # the order matters where two metals share a node
    cm = np.array([[1, 2], [3, 1]], dtype=np.int32)
    got = G.metal_point_types(np.zeros((3, 3), np.int32), cm, cm > 0)
    assert got[1, 1] == 1 and got[0, 1] == 2 and got[1, 0] == 3


# ------------------------------------------------------------ gases, species, tables
@needs_ref
def test_reference_gas_set_species_and_tables():
    from ipm2d import xsec
    run = RC.load(REF)
    species, gases = RC.gases_and_species(run)
    assert [sp.name for sp in species] == ["Electron", "F-", "CF3+", "CF3-"]
    assert [round(sp.charge / xsec.ELECTRON_CHARGE) for sp in species] == [-1, -1, 1, -1]
    # the configuration's recombination list pairs opposite charges, and its
    # temperature-dependent reaction (type 1) is the electron with a positive ion
    for a, b, kind in run.recombination.reactions:
        assert species[a].charge * species[b].charge < 0
        assert kind == 0 or (a == 0 and species[b].charge > 0)
    g = gases[0]
    assert (g.name, g.natom, g.ns, g.index_start) == ("CF4", 5, 3, 0)
    assert g.density == (100.0 * 0.13332) / (1.3807e-23 * 300.0)
    assert g.mass == 88.0043128 * 1.660539040e-27
    assert species[0].mass == (9.1095e-31 / 1.660539040e-27) * 1.660539040e-27
    tables = RC.mcc_tables(run, species, gases)
    assert sorted(tables) == [(0, 0), (1, 0), (2, 0), (3, 0)]
    e = tables[(0, 0)]
    assert (e.model, e.nreaction, e.nsigma) == (1, 8, 10001)
    assert e.prob.max() == 1.0 and np.all(np.diff(e.prob, axis=1) >= 0)
    assert e.collision_ratio == 1.0 - math.exp(-e.sigma_max * run.dt)
    for s in (1, 2, 3):
        b = tables[(s, 0)]
        assert b.model == 3 and b.nsigma == 1001
        assert np.all((b.resultant >= 0) & (b.resultant < len(species)))
    # the reactive selection ratio depends on the reduced mass only: equal for CF3+ and CF3-
    assert tables[(2, 0)].collision_ratio == tables[(3, 0)].collision_ratio != tables[(1, 0)].collision_ratio


def test_gas_physics_namelist(tmp_path):
    p = tmp_path / "X.txt"
    p.write_text(""" &GasPhysicsfile
 GP%Name      =           Xe,   ! a comment
 GP%MCModel   =           2,
 GP%NAtom     =           1,
 GP%Ns        =           2,
 GP%Mass      =           131.293,
 GP%Radius    =           2.16d-10,
 GP%BetaMax   =           3.d0
 GP%SP(1)     =           Xe+,1,1,131.293,2.d-10,
 GP%SP(2)     =           Xe2+,1,2,131.293,1.9D-10,
 /
""")
    gp = RC.read_gas_physics(p)
    assert (gp["Name"], gp["MCModel"], gp["NAtom"], gp["Ns"]) == ("Xe", 2, 1, 2)
    assert gp["Mass"] == 131.293 and gp["Radius"] == 2.16e-10 and gp["BetaMax"] == 3.0
    assert gp["SP"] == {1: ("Xe+", 1, 1, 131.293, 2e-10), 2: ("Xe2+", 1, 2, 131.293, 1.9e-10)}
    p.write_text(p.read_text().replace(" GP%SP(2)     =           Xe2+,1,2,131.293,1.9D-10,\n", ""))
    with pytest.raises(RC.ConfigError, match="SP"):
        RC.read_gas_physics(p)


# ------------------------------------------------------------ particles at load
def test_initial_particles_follow_the_load_rules(tmp_path):
    from ipm2d import xsec
    d = copy.deepcopy(BASE)
    d["plasma"]["ppg"] = 40
    d["plasma"]["init_electron_temperature"] = 30000.0
    run = RC.load(_write(tmp_path, d), nz=33, nr=17)
    cm = np.zeros((32, 16), dtype=np.int32)
    cm[0, :] = 1
    cm[-1, :] = 2
    geom = G.build_geometry(33, 17, run.dz, cm, nmetal=2, ndiel=0)
    species = [RC.SpeciesDef("Electron", 1, -xsec.ELECTRON_CHARGE, xsec.ELECTRON_MASS, 0.0, 0),
               RC.SpeciesDef("CF3+", 4, xsec.ELECTRON_CHARGE, 69.0059096 * xsec.ATOMIC_MASS, 2e-10, 1)]
    gen = np.random.default_rng(5)
    for s, temp in ((0, 30000.0), (1, 300.0)):
        P = RC.initial_particles(run, species, s, geom, gen)
        n = P.shape[1]
        iz, ir = np.floor(P[0]).astype(int), np.floor(P[1]).astype(int)
        assert np.all(geom.cell_load[iz, ir] == 1)
        # kept with probability R / j: the expected count of row j is ppg (j - 1/2) / j per cell
        cells_per_row = geom.cell_load.sum(axis=0)
        j = np.arange(1, 17)
        p = (j - 0.5) / j
        mean = (cells_per_row * 40 * p).sum()
        sd = math.sqrt((cells_per_row * 40 * p * (1 - p)).sum())
        assert abs(n - mean) < 5 * sd
        # the weights make the density uniform: total weight = n0 x loadable volume
        vol = (cells_per_row * math.pi * run.dz ** 3 * (2 * j - 1)).sum()
        assert P[7].sum() == pytest.approx(run.plasma.init_density * vol, rel=5 * sd / mean)
        # the speed sampler gives a Maxwell-Boltzmann energy distribution: mean 3/2 kT,
        # isotropic components, kT/m each
        vf = run.dz / run.dt
        v2 = (P[2:5] * vf) ** 2
        kt_m = xsec.KB * temp / species[s].mass
        e_mean = 0.5 * species[s].mass * v2.sum(axis=0).mean() / (xsec.KB * temp)
        assert e_mean == pytest.approx(1.5, rel=5 * math.sqrt(2.0 / 3.0) / math.sqrt(n))
        for c in range(3):
            assert v2[c].mean() == pytest.approx(kt_m, rel=0.03)
            assert abs((P[2 + c] * vf).mean()) < 5 * math.sqrt(kt_m / n)
        assert np.all(P[5:7] == 0.0)


class _Recorder:
    """Stands in for HotPath and records the set-up calls."""

    def __init__(self, nz, nr, dx, dt, ns, ng):
        self.args = (nz, nr, dx, dt, ns, ng)
        self.species, self.mcc, self.uploads = {}, {}, {}
        self.geom = self.rng = None

    def set_species(self, s, q, m, cap):
        self.species[s] = (q, m, cap)

    def set_geometry(self, g):
        self.geom = g

    def set_mcc(self, s, g, b):
        self.mcc[(s, g)] = b

    def upload(self, s, P):
        self.uploads[s] = P.shape

    def set_rng(self, seed, step):
        self.rng = (seed, step)


@needs_ref
def test_build_hotpath_from_the_reference_inputs():
    run = RC.load(REF)
    hp, species, gases, geom = RC.build_hotpath(run, seed=9, capacity=4_000_000, hotpath_cls=_Recorder)
    assert hp.args == (201, 201, run.dz, run.dt, 4, 1)
    assert sorted(hp.species) == [0, 1, 2, 3] and all(v[2] == 4_000_000 for v in hp.species.values())
    assert [hp.species[s][:2] for s in range(4)] == [(sp.charge, sp.mass) for sp in species]
    assert sorted(hp.mcc) == [(0, 0), (1, 0), (2, 0), (3, 0)]
    assert hp.geom is geom and hp.rng == (9, 0)
    loadable = int(geom.cell_load.sum())
    for s in range(4):
        assert hp.uploads[s][0] == 8 and 0.4 * 50 * loadable < hp.uploads[s][1] < 50 * loadable


def test_refused_inputs_say_what_is_missing(tmp_path):
    d = copy.deepcopy(BASE)
    d["gas"]["collision_section_type"] = 3
    run = _load(tmp_path, d)
    with pytest.raises(NotImplementedError, match="collision_section_type 3"):
        RC.gases_and_species(run, root=str(tmp_path))
    d["gas"]["collision_section_type"] = 1
    run = _load(tmp_path, d)
    gas = tmp_path / "input" / "gas" / "lxcat" / "CF4"
    gas.mkdir(parents=True)
    (gas / "CF4.txt").write_text(""" &GasPhysicsfile
 GP%Name = CF4, GP%MCModel = 2, GP%NAtom = 5, GP%Ns = 1, GP%Mass = 88.0043128,
 GP%SP(1) = CF3+,4,1,69.0059096,2.d-10,
 /
""")
    species, gases = RC.gases_and_species(run, root=str(tmp_path))
    assert [sp.name for sp in species] == ["Electron", "CF3+"] and gases[0].ns == 1
    # no cross-section file for the electron: the reference has no fallback for it
    with pytest.raises(NotImplementedError, match="ElectronCF4.txt"):
        RC.mcc_tables(run, species, gases, root=str(tmp_path))
    # an ion without its file gets the hard-sphere table (MCCInit.F90:107-111):
    # the gas has one ion species, so elastic 101 and charge exchange 103
    (gas / "ElectronCF4.txt").write_text(_xsec_text())
    tabs = RC.mcc_tables(run, species, gases, root=str(tmp_path))
    b = tabs[(1, 0)]
    assert b.model == 2 and list(b.rtype) == [101, 103] and list(b.resultant) == [1, 1]
    sigma = math.pi * (2e-10 + gases[0].radius) ** 2 / 2.0
    mu = species[1].mass * gases[0].mass / (species[1].mass + gases[0].mass)
    assert b.sigma_max == pytest.approx(sigma * math.sqrt(2e3 * 1.6022e-19 / mu) * gases[0].density, rel=1e-12)
    # and so does a species of another gas (MCCInit.F90:113-116)
    species[1].gas = 2
    assert list(RC.mcc_tables(run, species, gases, root=str(tmp_path))[(1, 0)].rtype) == [101, 103]


@needs_ref
def test_hard_sphere_for_the_reference_be_gas(tmp_path):
    """input/gas/lxcat/Be has Be+ as its one ion and no Be+Be.txt: the ion
    gets the hard-sphere pair with the radii of Be.txt (112 pm each)."""
    d = copy.deepcopy(BASE)
    d["gas"]["name"] = ["Be"]
    d["gas"]["ratio"] = [1.0]
    run = _load(tmp_path, d)
    root = os.path.dirname(REF)
    species, gases = RC.gases_and_species(run, root=root)
    assert [sp.name for sp in species] == ["Electron", "Be+"] and gases[0].ns == 1
    tabs = RC.mcc_tables(run, species, gases, root=root)
    b = tabs[(1, 0)]
    assert list(b.rtype) == [101, 103]
    sigma = 2.0 * math.pi * (0.5 * 224e-12) ** 2
    mu = species[1].mass * gases[0].mass / (species[1].mass + gases[0].mass)
    assert b.sigma_max == pytest.approx(sigma * math.sqrt(2e3 * 1.6022e-19 / mu) * gases[0].density, rel=1e-12)
    assert tabs[(0, 0)].model == 1


def _xsec_text():
    from ipm2d import xsec
    return xsec.write_sigma_file(1, xsec.synthetic_electron_reactions(ion_product=1, att_product=1))


# ------------------------------------------------------------ Pegasus gas sets
@needs_ref
def test_pegasus_ar_set_from_the_reference(tmp_path):
    """input/gas/pegasus/Ar: the gas line 180 (39.95 amu, 3.68 A), one ion
    Ar(+) (Index 181); the electron's five reactions 180000/1/6/7/8 become
    elastic 11, ionization 13 into Ar(+), and three excitations 12 with the
    files' thresholds; Ar(+) gets charge exchange 103 (type 6) and elastic
    101 (type 1); cross sections are the files' values / 1e4 (cm^2 -> m^2)."""
    d = copy.deepcopy(BASE)
    d["gas"]["name"] = ["Ar"]
    d["gas"]["ratio"] = [1.0]
    d["gas"]["collision_section_type"] = 2
    run = _load(tmp_path, d)
    root = os.path.dirname(REF)
    species, gases = RC.gases_and_species(run, root=root)
    assert [sp.name for sp in species] == ["Electron", "Ar(+)"]
    g = gases[0]
    assert g.name == "Ar" and g.ns == 1 and g.sop_index == [181] and g.cs_model == 2
    assert g.mass == pytest.approx(39.95 * 1.660539040e-27, rel=1e-15) and g.radius == pytest.approx(3.68e-10)
    assert species[1].charge == pytest.approx(1.6022e-19) and species[1].radius == pytest.approx(3.33e-10)
    sf = RC.pegasus_sigma_file(root, g, "Electron", 0)
    assert sf.model == 1
    assert [(r.rtype, r.resultant) for r in sf.reactions] == [(11, 0), (13, 1), (12, 0), (12, 0), (12, 0)]
    assert [r.threshold for r in sf.reactions] == [0.0, 15.76, 11.828, 12.900, 14.090]
    ion = sf.reactions[1]
    assert ion.energy[13] == 40.0 and ion.sigma[13] == pytest.approx(2.39e-20, rel=1e-14)
    sfi = RC.pegasus_sigma_file(root, g, "Ar(+)", 1)
    assert sfi.model == 2 and [r.rtype for r in sfi.reactions] == [103, 101]
    tabs = RC.mcc_tables(run, species, gases, root=root)
    assert tabs[(0, 0)].model == 1 and tabs[(0, 0)].nreaction == 5
    assert tabs[(1, 0)].model == 2 and list(tabs[(1, 0)].rtype) == [103, 101]


def _pegasus_reaction(path, index, species, change, rtype, th=1.0):
    path.write_text(f"""# a comment
#
{index}
A+B->C
{th}
{len(species)}
{' '.join(str(x) for x in species)}
{' '.join(f'{x:+d}' for x in change)}
{rtype}
3
 1.0  1.0E-16
 10.0  2.0E-16
 100.0  3.0E-16
""")


@pytest.mark.parametrize("species,change,rtype,want", [
    ((7, 1), (0, 0), 1, (11, 0)),                 # elastic
    ((7, 1, 8), (-1, 0, 1), 100, (12, 0)),        # excitation
    ((7, 1, 8), (-1, 1, 1), 2, (13, 1)),          # ionization into the first species
    ((7, 1, 9), (-1, -1, 1), 3, (14, 2)),         # attachment (type 3, one product)
    ((7, 1, 8, 9), (-1, 0, 1, 1), 3, (15, 12)),   # dissociation into two
    ((7, 1, 9), (-1, -1, 1), 4, (14, 2)),         # type 4, the electron lost
    ((7, 1, 9), (-1, 0, 1), 4, (12, 2)),          # type 4, the electron kept
    ((7, 1, 8), (-1, 1, 1), 255, (13, 1)),        # type 255, an electron gained
    ((7, 1), (0, 0), 255, (12, 0)),               # type 255, no product
    ((7, 8), (0, 0), 1, (101, 1)),                # ion elastic
    ((7, 8), (0, 0), 6, (103, 1)),                # ion charge exchange
    ((7, 8, 9), (-1, -1, 1), 255, (112, 2)),      # ion type 255 with a product
])
def test_pegasus_reaction_codes(tmp_path, species, change, rtype, want):
    """ConvertReactionOneBase2ReactionOne (R/src/mc/ReactionOnePegasus.f90:241-404)
    on hand-made reaction files: particle 1 is the electron, 7 the gas, 8
    and 9 the gas's species 1 and 2."""
    f = tmp_path / "r"
    _pegasus_reaction(f, 99, species, change, rtype)
    rx = RC.read_pegasus_reaction(str(f))
    assert rx["index"] == 99 and rx["threshold"] == 1.0 and rx["type"] == rtype
    np.testing.assert_array_equal(rx["sigma_cm2"], [1e-16, 2e-16, 3e-16])
    assert RC.pegasus_reaction_code(rx, [8, 9], 1) == want


def test_pegasus_electron_flag_follows_the_second_particle(tmp_path):
    """The source's electron test keeps only the second particle's answer:
    a file listing the electron first and the gas second reads as an ion
    reaction."""
    f = tmp_path / "r"
    _pegasus_reaction(f, 5, (1, 7), (0, 0), 1)
    assert RC.pegasus_reaction_code(RC.read_pegasus_reaction(str(f)), [8], 1) == (101, 1)
