"""From a reference-format input folder to steps on the GPU.

The folder is written here in the reference's formats: config.json (the keys
R/src/base/init.F90 reads), input/geom.h5 (the geometry datasets and Nz, Nr
attributes, written by ipm2d.h5lite), input/geom.json (the materials list)
and the lxcat gas set input/gas/lxcat/CF4/ (GasPhysicsfile namelist and one
cross-section namelist per species).  The geometry is a 65 x 33 node
capacitive cell: top and bottom electrodes, a metal outer wall and a
dielectric ring.

``refconfig.build_hotpath`` sets up two HotPaths from it with the same seed.
Five steps launched kernel by kernel on one and replayed as a captured graph
on the other leave the same particles (multisets of 64-bit hashes, bit for
bit) and deposits equal to 1e-12 of their magnitude, and the metals collect
charge.
"""
import json

import numpy as np
import pytest

from ipm2d import geometry as G
from ipm2d import h5lite, refconfig as RC, xsec

pytestmark = pytest.mark.gpu

NZ, NR = 65, 33


def _hashes(P):
    b = np.ascontiguousarray(P).view(np.uint64)
    h = np.full(b.shape[1], 0x9E3779B97F4A7C15, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for f in range(8):
            h ^= b[f] * np.uint64(0xBF58476D1CE4E5B9 + 2 * f)
            h = (h << np.uint64(27)) | (h >> np.uint64(37))
            h *= np.uint64(0x94D049BB133111EB)
    return np.sort(h)


def _write_pegasus_ar(root):
    """A Pegasus-format Ar set (input/gas/pegasus/Ar/): the gas file, the
    electron's three reactions (elastic, ionization into Ar(+), an
    excitation), Ar(+)'s charge exchange and elastic scattering, and no file
    for Ar(2+), which therefore gets the hard-sphere table."""
    d = root / "input" / "gas" / "pegasus" / "Ar"
    d.mkdir(parents=True)
    (d / "Ar.txt").write_text("Ar\n1\n1\n3\n180  39.95  0.0  3.68  93.30  Ar\n"
                              "181  39.95  +1.0  3.33  -1  Ar(+)\n182  39.95  +2.0  3.00  -1  Ar(2+)\n")
    (d / "ElectronAr.txt").write_text("1  3\n180000  Elastic 0.0\n180001  Ionization 15.76\n"
                                      "180006  Excitation 11.828\n")
    (d / "Ar(+)Ar.txt").write_text("2  2\n180011  Exchange 0.0\n180012  Elastic 0.0\n")
    e = np.logspace(-2, 3, 60)

    def rx(index, name, th, species, change, rtype, sigma_cm2):
        rows = "\n".join(f"  {a:.6e}  {b:.6e}" for a, b in zip(e, sigma_cm2))
        (d / str(index)).write_text(f"# synthetic\n{index}\n{name}\n{th}\n{len(species)}\n"
                                    f"{' '.join(map(str, species))}\n{' '.join(f'{x:+d}' for x in change)}\n"
                                    f"{rtype}\n{len(e)}\n{rows}\n")
    rx(180000, "e(-)+Ar->Ar+e(-)", 0.0, (180, 1), (0, 0), 1, np.full(e.size, 1.0e-15))
    rx(180001, "e(-)+Ar->Ar(+)+e(-)+e(-)", 15.76, (180, 1, 181), (-1, 1, 1), 2,
       np.where(e > 15.76, 3e-16 * (1 - np.exp(-(e - 15.76) / 30.0)), 0.0))
    rx(180006, "e(-)+Ar->Ar(*)+e(-)", 11.828, (180, 1, 186), (-1, 0, 1), 100,
       np.where(e > 11.828, 1e-16, 0.0))
    rx(180011, "Ar(+)+Ar->Ar+Ar(+)", 0.0, (180, 181), (0, 0), 6, np.full(e.size, 6e-15))
    rx(180012, "Ar(+)+Ar->Ar(+)+Ar", 0.0, (180, 181), (0, 0), 1, np.full(e.size, 2e-14))


def _write_inputs(root, controls=False, pegasus=False):
    (root / "input" / "gas" / "lxcat" / "CF4").mkdir(parents=True)
    cm = np.zeros((NZ - 1, NR - 1), dtype=np.int32)      # indices into the materials list
    cm[:2, :] = 1
    cm[-2:, :] = 2
    cm[:, -1] = 3
    cm[24:40, -3:-1] = 4
    types = np.array([0, 1, 1, 1, 2])
    ctype = types[cm]
    he, ve = G._edge_types(ctype)
    metal = ctype == G.MAT_METAL
    geom = {"Nz": NZ, "Nr": NR, "cell_material": cm, "epsilon_r": np.where(cm == 4, 2.1, 1.0),
            "load": (ctype == 0).astype(np.int32),
            "point_type": G.metal_point_types(np.zeros((NZ, NR), np.int32), metal.astype(np.int32), metal),
            "volume": G._volume_factor(NZ, NR, ctype == 0), "hedge_type": he, "vedge_type": ve}
    h5lite.write_geometry(root / "input" / "geom.h5", geom)
    mats = [{"type": 0, "name": "vacuum"}, {"type": 1, "name": "top electrode"},
            {"type": 1, "name": "bottom electrode"}, {"type": 1, "name": "wall"},
            {"type": 2, "name": "ring", "permittivity": 2.1, "permeability": 1.0, "conductivity": 1.0}]
    (root / "input" / "geom.json").write_text(json.dumps({"material_amount": len(mats), "materials": mats}))
    gas = root / "input" / "gas" / "lxcat" / "CF4"
    (gas / "CF4.txt").write_text(""" &GasPhysicsfile
 GP%Name      =           CF4,
 GP%MCModel   =           2,
 GP%NAtom     =           5,
 GP%Ns        =           3,
 GP%Mass      =           88.0043128,
 GP%Radius    =           1.548d-10,
 GP%BetaMax   =           3.d0
 GP%SP(1)     =           F-,1,-1,18.9984030,1.35d-10,
 GP%SP(2)     =           CF3+,4,1,69.0059096,2.d-10,
 GP%SP(3)     =           CF3-,4,-1,69.0059096,2.d-10,
 /
""")
    R = xsec.Reaction
    (gas / "ElectronCF4.txt").write_text(
        xsec.write_sigma_file(1, xsec.synthetic_electron_reactions(ion_product=2, att_product=1)))
    (gas / "F-CF4.txt").write_text(xsec.write_sigma_file(3, [R(0, 101, 1, 0.0), R(0, 113, 0, 3.521),
                                                             R(0, 111, 1, 5.621)]))
    (gas / "CF3+CF4.txt").write_text(xsec.write_sigma_file(3, [R(0, 101, 2, 0.0), R(0, 111, 2, 5.621),
                                                               R(0, 100, 0, 5.843)]))
    (gas / "CF3-CF4.txt").write_text(xsec.write_sigma_file(3, [R(0, 101, 3, 0.0), R(0, 113, 0, 1.871),
                                                               R(0, 112, 1, 1.927)]))
    cfg = {
        "control": {"nrun": 10, "dt": 5e-11, "is_restart": False, "pic_type": 1},
        "diag": {"diag_type": 1, "diag_cycles": 1, "diag_preiod": 20, "diag_freq": 0.0, "diag_Nt": 5},
        "grid": {"z_length": 0.032, "deps_in_z": 1, "deps_in_r": 1, "deps_type": 0,
                 "materials": "./input/geom.json", "geometry": "./input/geom.h5"},
        "plasma": {"init_density": 1e16, "init_electron_temperature": 60000.0, "ppg": 12, "alpha": 0.01,
                   "is_open_split_merge": False, "paritlc_upper_limit": 1.01, "paritlc_lower_limit": 0.99,
                   "min_mp_to_split_k": 0.1, "merge_limit": 0.34, "load_type": "1",
                   "particle_number_max_ratio": 10.0, "particle_limit_max_velocity": 10.0},
        "gas": {"ng": 1, "pressure": 300, "temperature": 300, "name": ["CF4"], "ratio": [1.0],
                "collision_section_type": 1, "is_open_recombination": False},
        "circuit": {"circuits_number": 1, "circuits_list": [{"type": 1, "metal_index": 2, "name": "ground"}]},
    }
    if pegasus:
        _write_pegasus_ar(root)
        cfg["gas"].update({"name": ["Ar"], "collision_section_type": 2})
    if controls:                            # the shipped configuration's controls
        cfg["plasma"]["is_open_split_merge"] = True
        cfg["diag"]["diag_preiod"] = 4
        cfg["gas"]["is_open_recombination"] = True
        cfg["gas"]["recb"] = {"recb_time_step_num": 5, "recb_reaction_num": 3,
                              "recb_list": [[1, 2, 0], [2, 3, 0], [0, 2, 1]],
                              "recb_rate": [1e-6, 1e-6, 3.95e-15]}
    (root / "config.json").write_text(json.dumps(cfg))
    return root / "config.json"


def test_reference_inputs_to_gpu_steps(tmp_path):
    run = RC.load(_write_inputs(tmp_path))
    assert (run.nz, run.nr, run.dt) == (NZ, NR, 5e-11)
    a, species, gases, geom = RC.build_hotpath(run, seed=3, capacity=60000)
    b, _, _, _ = RC.build_hotpath(run, seed=3, capacity=60000)
    assert [sp.name for sp in species] == ["Electron", "F-", "CF3+", "CF3-"]
    assert (geom.nmetal, geom.ndiel) == (3, 1)
    n0 = [a.count(s) for s in range(4)]
    assert all(n > 10000 for n in n0)
    # 150 V on the top electrode falling linearly to the grounded bottom one
    z = np.arange(NZ)[:, None] * np.ones((1, NR))
    phi = 150.0 * (1.0 - z / (NZ - 1))
    for hp in (a, b):
        hp.set_potential(phi)
    for _ in range(5):
        a.one_step()
        b.one_step_graph()
    rho_a, _ = a.field_source_sum()
    rho_b, _ = b.field_source_sum()
    np.testing.assert_allclose(rho_a, rho_b, rtol=1e-12, atol=1e-12 * np.abs(rho_b).max())
    for s in range(4):
        np.testing.assert_array_equal(_hashes(a.download(s)), _hashes(b.download(s)), err_msg=species[s].name)
    metal, _ = a.take_tallies()
    assert np.any(metal != 0.0), "nothing reached the metals"
    a.close()
    b.close()


def test_pegasus_inputs_to_gpu_steps(tmp_path):
    """The same cell with a Pegasus-format Ar set (collision_section_type 2,
    R/src/mc/ReactionOnePegasus.f90): the tables come from the reaction files,
    Ar(2+) without a file gets the hard-sphere table, and five eager steps
    equal five graph steps bit for bit."""
    run = RC.load(_write_inputs(tmp_path, pegasus=True))
    species, gases = RC.gases_and_species(run)
    tabs = RC.mcc_tables(run, species, gases)
    assert [sp.name for sp in species] == ["Electron", "Ar(+)", "Ar(2+)"]
    assert list(tabs[(0, 0)].rtype) == [11, 13, 12] and tabs[(0, 0)].resultant[1] == 1
    assert list(tabs[(1, 0)].rtype) == [103, 101] and list(tabs[(2, 0)].rtype) == [101]
    a, _, _, _ = RC.build_hotpath(run, seed=5, capacity=60000)
    b, _, _, _ = RC.build_hotpath(run, seed=5, capacity=60000)
    n0 = [a.count(s) for s in range(3)]
    z = np.arange(NZ)[:, None] * np.ones((1, NR))
    phi = 300.0 * (1.0 - z / (NZ - 1))
    for hp in (a, b):
        hp.set_potential(phi)
        hp.set_event_counting(True)
    for _ in range(5):
        a.one_step()
        b.one_step_graph()
    b.synchronize()
    for s in range(3):
        np.testing.assert_array_equal(_hashes(a.download(s)), _hashes(b.download(s)), err_msg=species[s].name)
    counts = a.take_mcc_counts()
    assert counts[0, 0, 1] > 100                      # electron elastic events from file 180000
    assert all(a.count(s) > 0.9 * n0[s] for s in range(3))
    a.close()
    b.close()


def test_control_loop_from_reference_inputs(tmp_path):
    """The loop of the integration notes, with the shipped configuration's
    controls switched on: split/merge of the electrons every step and of
    every bundle every Period (4) steps, recombination every 5 steps with
    the shipped reaction list (the two constant rates raised to 1e-6 m^3/s
    so that twelve small steps see deletions).  Every control acts, and
    every particle stays inside the domain with a positive weight."""
    run = RC.load(_write_inputs(tmp_path, controls=True))
    hp, species, _, _ = RC.build_hotpath(run, seed=4, capacity=80000)
    rezone = run.rezoning(hp)
    recomb = run.recombination_control(hp)
    assert rezone.period == 4 and recomb.time_step_number == 5
    z = np.arange(NZ)[:, None] * np.ones((1, NR))
    hp.set_potential(150.0 * (1.0 - z / (NZ - 1)))
    touched, deleted = set(), np.zeros(4, dtype=np.int64)
    for timer in range(1, 13):
        hp.one_step_graph()
        hp.field_source_sum()
        touched |= set(rezone.control(timer))
        d = recomb.control(timer)
        if d is not None:
            deleted += np.asarray(d)
    assert touched == {0, 1, 2, 3}
    assert deleted[1] > 0 and deleted[2] > 0 and deleted[3] > 0
    for s in range(4):
        P = hp.download(s)
        assert P.shape[1] > 0, species[s].name
        assert np.all((P[0] > 0) & (P[0] < NZ - 1) & (P[1] >= 0) & (P[1] < NR - 1)), species[s].name
        assert np.all(P[7] > 0), species[s].name
    hp.close()
