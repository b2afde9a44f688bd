"""The oracle's pins catch plausible mistakes.

Each case plants one mistake of the kind a restatement could make (a
dropped or shifted index, a swapped operand, a wrong sign, a lost quirk of
the reference) in a copy of oracle/ipm_oracle.c, builds the copy on the CPU,
and runs the pin test that should catch it against that build: the pin must
fail.  The oracle itself is unchanged; only the copies carry the mistakes.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import test_oracle_absorb_mcc as AM
import test_oracle_mcc_channels as MC
import test_oracle_moments_circuit as MO
import test_oracle_push_deposit as PD
import test_oracle_recombination as RC
import test_oracle_set_e as SE
import test_oracle_solver as SO
import test_oracle_split_merge as SM
import test_xsec as XS
from oracle import ipm_oracle

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "ipm_oracle.c")


def _rng():
    return np.random.default_rng(2024)


# (name, text in ipm_oracle.c, planted mistake, pin that must fail)
MUTANTS = [
    ("push: half-cell stagger of the Ez gather dropped",
     "int N1H = (int)(z + 0.5);", "int N1H = (int)(z);",
     lambda O: PD.test_push_gather_stays_between_bracketing_faces(O, _rng())),
    ("push: Er gathered from the cell rows in place of its half-cell rows",
     "int N2H = (int)(r + 0.5);", "int N2H = (int)(r);",
     lambda O: PD.test_push_gather_stays_between_bracketing_faces(O, _rng())),
    ("push: sign of the rotation's Vt term",
     "vt = -Vx * SinA + Vy * CosA;", "vt = Vx * SinA + Vy * CosA;",
     lambda O: PD.test_push_zero_field_is_a_straight_line(O, _rng())),
    ("deposit: upper z weight in place of the lower one",
     "RH(NzL, NrU) = RH(NzL, NrU) + ZL * RU * WQ[p];", "RH(NzL, NrU) = RH(NzL, NrU) + ZU * RU * WQ[p];",
     lambda O: PD.test_deposit_conserves_weight_and_linear_moments(O, _rng())),
    ("SetE: zero-gradient ghost instead of linear extrapolation",
     "PP(0, j) = 2.0 * PP(1, j) - PP(2, j);", "PP(0, j) = PP(1, j);",
     lambda O: SE.test_linear_potential_gives_uniform_field(O)),
    ("ions 111-113: the Vy <- Vz quirk restored to Vy",
     "m->poi.Vt = B * m->poi.Vz + A * m->pot.Vt + B * m->gy;",
     "m->poi.Vt = B * m->poi.Vt + A * m->pot.Vt + B * m->gy;",
     lambda O: MC.test_ion_reactive_hand_worked_no_threshold(O, 111)),
    ("charge exchange keeps the ion's own velocity",
     "case 103: m->poi = m->pot; break;", "case 103: break;",
     lambda O: MC.test_charge_exchange_takes_target_velocity(O, 103)),
    ("Langevin selection: beta_max * R instead of beta_max * sqrt(R)",
     "m->beta = t->beta_max * sqrt(R);", "m->beta = t->beta_max * R;",
     lambda O: MC.test_langevin_selection_fraction(O, _rng(), 3.0)),
    ("dissociation: product digits swapped",
     "m->pon_idx[0] = resultant % 10;\n        m->pon_idx[1] = resultant / 10;",
     "m->pon_idx[0] = resultant / 10;\n        m->pon_idx[1] = resultant % 10;",
     lambda O: MC.test_dissociation_product_species(O, 20, {0: 1, 2: 1})),
    ("sigma interpolation: dble(i - 1) for the reference's dble(i)",
     "double S1 = ((double)i - LE) / (UE - LE);", "double S1 = ((double)(i - 1) - LE) / (UE - LE);",
     lambda O: MC.test_sigma_interpolation_uses_dble_i_weight(O, "oracle")),
    ("node volumes: axis shell with 1/4 in place of 1/3",
     "v = f * 1.0 / 3.0 * C_PI * dx * dx * dx;", "v = f * 1.0 / 4.0 * C_PI * dx * dx * dx;",
     lambda O: PD.test_node_volumes_tile_the_cylinder(O)),
    ("absorption: the axis reflection dropped",
     "R[i] = fabs(R[i]);", "R[i] = R[i];",
     lambda O: AM.test_absorption_rule_table(O)),
    ("Maxwellian sampler: the e factor of the acceptance test dropped",
     "FuncB = -C_EXP * R * log(R);", "FuncB = -R * log(R);",
     lambda O: AM.test_maxwellian_acceptance_rate(O)),
    ("Vahedi angle: (1 + E)^R taken as E^R",
     "return (E + 2.0 - 2.0 * pow(E + 1.0, R)) / E;", "return (E + 2.0 - 2.0 * pow(E, R)) / E;",
     lambda O: AM.test_vahedi_cos_chi_cdf(O, 10.0)),
    ("excitation: the threshold loss not applied to the velocity",
     "case 12: {\n        double VF = sqrt(1.0 - th / m->energy);", "case 12: {\n        double VF = 1.0;",
     lambda O: AM.test_excitation_removes_threshold(O, _rng())),
    ("moments: the 1/2 of the kinetic energy dropped",
     "const double EnergyFactor = 0.5 * mass * VFactor * VFactor;",
     "const double EnergyFactor = mass * VFactor * VFactor;",
     lambda O: MO.test_moments_uniform_velocity_and_speed(O, _rng())),
    ("hard sphere: elastic share 0.25 in place of 0.5",
     "prob[(size_t)(i - 1) * nreaction] = (nreaction == 2) ? 0.5 : 1.0;",
     "prob[(size_t)(i - 1) * nreaction] = (nreaction == 2) ? 0.25 : 1.0;",
     lambda O: XS.test_hard_sphere_closed_form(O, 1)),
    ("merge: kept when the draw is above WQ/ref",
     "if (u < WQ[i] / ref) {", "if (u > WQ[i] / ref) {",
     lambda O: SM.test_merge_keep_fraction_and_expected_weight(O)),
    ("merge: threshold made inclusive",
     "if (WQ[i] < merge_limit * ref) {", "if (WQ[i] <= merge_limit * ref) {",
     lambda O: SM.test_merge_threshold_is_strict(O)),
    ("recombination: the drawn deletion ignores the particle's weight",
     "if (u < recb[c] / P[7][i]) {", "if (u < recb[c]) {",
     lambda O: RC.test_visiting_order_outcomes(O)),
    ("Poisson system: CoeC with the left cells' permittivity",
     "AT(coeC, i, j) = (1.0 + 0.5 * (CHI(i, j) + CHI(i + 1, j))) * 0.5 * (EPSR(i, j - 1) + EPSR(i, j));",
     "AT(coeC, i, j) = (1.0 + 0.5 * (CHI(i, j) + CHI(i + 1, j))) * 0.5 * (EPSR(i - 1, j - 1) + EPSR(i, j));",
     lambda O: SO.test_flux_symmetry_for_any_medium()),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
@pytest.mark.parametrize("name,find,plant,pin", MUTANTS, ids=[m[0].split(":")[0] for m in MUTANTS])
def test_pin_catches_planted_mistake(tmp_path, name, find, plant, pin):
    src = open(SRC).read()
    assert src.count(find) == 1, f"mutation site not unique or missing: {find!r}"
    mutant = tmp_path / "mutant.c"
    mutant.write_text(src.replace(find, plant))
    lib = str(tmp_path / "libmutant.so")
    subprocess.run(["gcc", "-shared", "-fPIC", "-std=c99", "-O1", "-ffp-contract=off", "-fno-fast-math",
                    "-I", os.path.dirname(SRC), str(mutant), "-o", lib, "-lm"], check=True, capture_output=True)
    with ipm_oracle.library_override(lib):
        with pytest.raises(AssertionError):
            pin(ipm_oracle)


def test_pins_pass_on_the_oracle():
    """The same pins, run on the oracle itself, pass (so the failures above
    come from the planted mistakes)."""
    for name, _, _, pin in MUTANTS:
        pin(ipm_oracle)
