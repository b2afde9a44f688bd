"""Pins of the oracle's ion channels, reactive selection, dissociation
products, gas loop order and table interpolation, each against a
hand-worked value, a conservation law or a closed-form frequency that a
plausible misreading of the reference would fail.

* PostVelocityIonReactive (R/src/mc/MCCIon.F90:54-74), types 111/112/113:
  hand-worked post-collision velocities for mass ratio A = 1/4, including
  the source's Vy <- B*Vz + A*pot%Vt + B*Gy (a reading with Vt in place of
  Vz gives another number), the threshold's VF = sqrt(1 - Th/E) and the
  products (the target for 112, the updated ion for 113).
* Charge exchange 102/103 (MCCIon.F90:17-22): the ion leaves with the
  target's velocity, hand-worked on one particle and 3/2 kT_gas on average.
* Elastic scatter 101 / 11 (TypeMCC.f90:114-146): the speed relative to the
  centre of mass, MassRatio*g, is conserved, and the CM deflection angle is
  isotropic (101) or Vahedi's (11).
* Langevin selection (TypeMCC.f90:218-236): beta = beta_max*sqrt(R) <= 1 has
  probability 1/beta_max^2; energies at or below Threshold(2) give
  reaction 1 without the table lookup.
* Dissociation 15 (MCCElectron.F90:140-145): products in species
  mod(Resultant, 10) and Resultant/10.
* Gas loop (MCCPublic.F90:52-115): gases in order, each seeing the state the
  earlier ones left.
* SigmaNormalizedUpdate (MCCSigma.f90:131-190): raw sigma linear in the
  log-energy index interpolates to a + b*i with the source's dble(i) weight
  (dble(i-1) would give a + b*(i-1)); raw energies below 10**Emin get the
  index coordinate Emin itself; 0 at or below the threshold; Alfa/E beyond
  the last raw energy.
"""
import math

import numpy as np
import pytest

from helpers import AMU, ME, QE
from ipm2d import xsec

KB = xsec.KB
DX, DT = 1e-3, 5e-11


def _poi():
    # Z R Vz Vr Vt Az Ar WQ, velocities in m/s
    return np.array([2.5, 3.5, 3.0, 1.0, 2.0, 0.01, 0.02, 7.0])


POT = np.array([0.3, 0.1, 0.2])        # target Vz Vr Vt
SM, GM = 3.0 * AMU, 1.0 * AMU          # MassRatio A = GM/(SM+GM) = 1/4, B = 3/4


def _cm_energy(poi, pot):
    g = np.array([poi[3] - pot[1], poi[4] - pot[2], poi[2] - pot[0]])
    mu = SM * GM / (SM + GM)
    return 0.5 * mu * float(g @ g) / xsec.JTOEV


@pytest.mark.parametrize("rtype", [111, 112, 113])
def test_ion_reactive_hand_worked_no_threshold(oracle, rtype):
    out, en, prods = oracle.collide_given(3, 1, rtype, 2, 0.0, SM, GM, _poi(), POT)
    # g = poi - pot = (gx, gy, gz) = (0.9, 1.8, 2.7); VF = 1
    assert out[3] == pytest.approx(0.75 * 1.0 + 0.25 * 0.1 + 0.75 * 0.9, rel=1e-14)      # Vr = 1.45
    assert out[4] == pytest.approx(0.75 * 3.0 + 0.25 * 0.2 + 0.75 * 1.8, rel=1e-14)      # Vt = 3.65 (B*Vz, not B*Vt)
    assert out[2] == pytest.approx(0.75 * 3.0 + 0.25 * 0.3 + 0.75 * 2.7, rel=1e-14)      # Vz = 4.35
    assert out[4] != pytest.approx(0.75 * 2.0 + 0.25 * 0.2 + 0.75 * 1.8, rel=1e-6)
    np.testing.assert_array_equal(out[[0, 1, 5, 6, 7]], _poi()[[0, 1, 5, 6, 7]])
    assert en == pytest.approx(_cm_energy(_poi(), POT), rel=1e-14)
    if rtype == 111:
        assert prods == []
    elif rtype == 112:
        # the product is the target: parent's place and weight, target velocity,
        # Az/Ar scaled by Mass/GasMass (Updater, TypeMCC.f90:90-91)
        (d, p), = prods
        assert d == 2
        np.testing.assert_array_equal(p[[0, 1, 7]], _poi()[[0, 1, 7]])
        np.testing.assert_array_equal(p[2:5], POT)
        assert p[5] == pytest.approx(0.01 * 3.0, rel=1e-14) and p[6] == pytest.approx(0.02 * 3.0, rel=1e-14)
    else:
        (d, p), = prods
        assert d == 2
        np.testing.assert_array_equal(p, out)       # the updated ion itself


def test_ion_reactive_threshold_scales_velocities(oracle):
    E = _cm_energy(_poi(), POT)
    th = 0.75 * E                                # VF = sqrt(1 - 3/4) = 1/2
    out, en, _ = oracle.collide_given(3, 1, 111, 1, th, SM, GM, _poi(), POT)
    assert en == pytest.approx(0.25 * E, rel=1e-12)
    # poi -> (1.5, 0.5, 1.0) (Vz, Vr, Vt); g -> (0.45, 0.9, 1.35) (x, y, z)
    assert out[3] == pytest.approx(0.75 * 0.5 + 0.25 * 0.1 + 0.75 * 0.45, rel=1e-12)
    assert out[4] == pytest.approx(0.75 * 1.5 + 0.25 * 0.2 + 0.75 * 0.9, rel=1e-12)
    assert out[2] == pytest.approx(0.75 * 1.5 + 0.25 * 0.3 + 0.75 * 1.35, rel=1e-12)


@pytest.mark.parametrize("rtype", [102, 103])
def test_charge_exchange_takes_target_velocity(oracle, rtype):
    out, _, prods = oracle.collide_given(2, 1, rtype, 1, 0.0, SM, GM, _poi(), POT)
    assert prods == []
    np.testing.assert_array_equal(out[2:5], POT)
    np.testing.assert_array_equal(out[[0, 1, 7]], _poi()[[0, 1, 7]])
    assert out[5] == pytest.approx(0.01 * 3.0, rel=1e-14) and out[6] == pytest.approx(0.02 * 3.0, rel=1e-14)


def test_ion_100_changes_nothing(oracle):
    out, _, prods = oracle.collide_given(2, 1, 100, 0, 0.0, SM, GM, _poi(), POT)
    assert prods == []
    np.testing.assert_array_equal(out, _poi())


def _cm_frame(poi_v, pot_v, out_v, sm, gm):
    """(speed relative to the CM before, after, cos of the CM deflection)."""
    vcm = (sm * poi_v + gm * pot_v) / (sm + gm)
    a, b = poi_v - vcm, out_v - vcm
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    return na, nb, float(a @ b) / (na * nb)


@pytest.mark.parametrize("model,rtype,sm", [(2, 101, 3.0 * AMU), (1, 11, ME)])
def test_elastic_scatter_conserves_cm_speed(oracle, model, rtype, sm):
    rng = np.random.default_rng(4)
    dev = []
    for c0 in range(4000):
        poi = _poi()
        poi[2:5] = rng.standard_normal(3) * (2e6 if model == 1 else 500.0)
        pot = rng.standard_normal(3) * 300.0
        out, en, prods = oracle.collide_given(model, 0, rtype, 0, 0.0, sm, GM, poi, pot, seed=5, c0=c0)
        assert prods == []
        na, nb, c = _cm_frame(poi[2:5], pot, out[2:5], sm, GM)
        assert nb == pytest.approx(na, rel=1e-9)     # |v' - v_cm| = MassRatio * g = |v - v_cm|
        if rtype == 101:      # isotropic: E[cos chi] = 0
            dev.append(c)
        else:                 # Vahedi: E[cos chi] = 1 + 2/E - 2/ln(1 + E)
            dev.append(c - (1.0 + 2.0 / en - 2.0 / math.log1p(en)))
    d = np.array(dev)
    assert abs(d.mean()) < 6 * d.std() / math.sqrt(d.size)
    if rtype == 101:
        assert abs((d ** 2).mean() - 1 / 3) < 6 * math.sqrt(4 / 45 / d.size)


def test_charge_exchange_ions_leave_at_gas_temperature(oracle, rng):
    """Every ion charge-exchanges (CollisionRatio 1, Probility 1): the ions'
    velocities are the gas Maxwellian, mean 1/2 m_gas v^2 = 3/2 kT_gas."""
    Tg, mg = 600.0, 88.0 * AMU
    b = xsec.MCCBundle(model=2, nreaction=1, nsigma=4, emin=-1.0, einterval=1.0, emax=100.0, collision_ratio=1.0,
                       rtype=np.array([103], np.int32), resultant=np.array([1], np.int32), threshold=np.zeros(1),
                       prob=np.ones((4, 1)), gas_mass=mg, gas_temperature=Tg)
    st = oracle.MccSetup(DX, DT, [-QE, QE], [ME, 31 * AMU], {(1, 0): b})
    n = 20000
    P = np.zeros((8, n))
    P[0], P[1], P[7] = 3.5, 2.5, 1.0
    P[2:5] = rng.standard_normal((3, n)) * 2e-3
    out, cnt = oracle.mcc(st, 3, 0, [np.zeros((8, 0)), P], counts=True)
    assert cnt[1, 0, 1] == n
    v2 = ((out[1][2:5] * (DX / DT)) ** 2).sum(axis=0)
    e = 0.5 * mg * v2 / (KB * Tg)
    assert abs(e.mean() - 1.5) < 6 * e.std() / math.sqrt(n)


def _reactive_table(th2, beta_max=3.0):
    # reaction 1 carries no probability; reaction 2 all of it (cumulative 1)
    prob = np.zeros((4, 2))
    prob[:, 1] = 1.0
    return xsec.MCCBundle(model=3, nreaction=2, nsigma=4, emin=-1.0, einterval=1.0, emax=100.0, collision_ratio=1.0,
                          rtype=np.array([101, 111], np.int32), resultant=np.array([1, 1], np.int32),
                          threshold=np.array([0.0, th2]), prob=prob, gas_mass=88.0 * AMU, gas_temperature=300.0,
                          beta_max=beta_max)


@pytest.mark.parametrize("beta_max", [3.0, 1.5])
def test_langevin_selection_fraction(oracle, rng, beta_max):
    """beta_max*sqrt(R) <= 1 with probability 1/beta_max^2 (a reading with
    beta = beta_max*R would give 1/beta_max)."""
    st = oracle.MccSetup(DX, DT, [-QE, -QE], [ME, 19 * AMU], {(1, 0): _reactive_table(0.0, beta_max)})
    n = 60000
    P = np.zeros((8, n))
    P[0], P[1], P[7] = 3.5, 2.5, 1.0
    P[2:5] = rng.standard_normal((3, n)) * 2e-3
    _, cnt = oracle.mcc(st, 8, 0, [np.zeros((8, 0)), P], counts=True)
    n2 = cnt[1, 0, 2]
    p = 1.0 / beta_max ** 2
    assert cnt[1, 0, 1] + n2 == n
    assert abs(n2 - n * p) < 6 * math.sqrt(n * p * (1 - p))


def test_langevin_sub_threshold_gives_reaction_one(oracle, rng):
    """E <= Threshold(2) returns reaction 1 before the table (whose masked
    lookup would return 0, a null event)."""
    st = oracle.MccSetup(DX, DT, [-QE, -QE], [ME, 19 * AMU], {(1, 0): _reactive_table(1e6)})
    n = 5000
    P = np.zeros((8, n))
    P[0], P[1], P[7] = 3.5, 2.5, 1.0
    P[2:5] = rng.standard_normal((3, n)) * 2e-3
    _, cnt = oracle.mcc(st, 9, 0, [np.zeros((8, 0)), P], counts=True)
    assert cnt[1, 0, 1] == n and cnt[1, 0, 0] == 0


@pytest.mark.parametrize("res,expect", [(21, {1: 1, 2: 1}), (11, {1: 2}), (20, {0: 1, 2: 1})])
def test_dissociation_product_species(oracle, res, expect):
    poi = _poi()
    poi[2:5] = [3e6, 1e6, 2e6]
    out, en, prods = oracle.collide_given(1, 0, 15, res, 1.0, ME, 88 * AMU, poi, POT, seed=2)
    got = {}
    for d, p in prods:
        got[d] = got.get(d, 0) + 1
        np.testing.assert_array_equal(p[2:5], POT)       # both products are the target
    assert got == expect
    assert [d for d, _ in prods] == [res % 10, res // 10]
    assert en == pytest.approx(_cm_energy_e(poi, POT) - 1.0, rel=1e-12)


def _cm_energy_e(poi, pot):
    g = np.array([poi[3] - pot[1], poi[4] - pot[2], poi[2] - pot[0]])
    gm = 88 * AMU
    mu = ME * gm / (ME + gm)
    return 0.5 * mu * float(g @ g) / xsec.JTOEV


def test_ionization_secondary_keeps_primary_direction(oracle):
    """The secondary is the primary before the collision scaled by
    sqrt(EnergyNew/E0) (same direction); EnergyOld = E - Th - EnergyNew
    inverts CreateEnergy's 10 tan(R atan(E - Th)/20) (MCCChi.f90:49-59) to
    a uniform R; the ion is the target."""
    poi = _poi()
    poi[2:5] = [3e6, 1e6, 2e6]
    E, th = _cm_energy_e(poi, POT), 16.0
    R = []
    for c0 in range(3000):
        out, en, prods = oracle.collide_given(1, 0, 13, 1, th, ME, 88 * AMU, poi, POT, seed=6, c0=c0)
        (d0, sec), (d1, ion) = prods
        assert d0 == 0 and d1 == 1
        u, w = poi[2:5] / np.linalg.norm(poi[2:5]), sec[2:5] / np.linalg.norm(sec[2:5])
        assert float(u @ w) == pytest.approx(1.0, abs=1e-14)
        np.testing.assert_array_equal(ion[2:5], POT)
        np.testing.assert_array_equal(sec[[0, 1, 5, 6, 7]], poi[[0, 1, 5, 6, 7]])
        assert en == pytest.approx(E, rel=1e-14)        # MCPO%Energy is left as it was
        e_new = E * float(sec[2:5] @ sec[2:5]) / float(poi[2:5] @ poi[2:5])
        e_old = E - th - e_new
        R.append(20.0 * math.atan(e_old / 10.0) / math.atan(E - th))
    R = np.array(R)
    assert R.min() > 0.0 and R.max() < 1.0
    assert abs(R.mean() - 0.5) < 6 * math.sqrt(1 / 12 / R.size)
    assert abs(R.var() - 1 / 12) < 6 * math.sqrt(1 / 180 / R.size)


def test_gases_run_in_order_on_the_updated_state(oracle, rng):
    """Two gases, both charge exchange with every ion: the ion leaves with
    the last gas's velocity distribution (gas 2 at 30 K, gas 1 at 3000 K)."""
    def cx(T):
        return xsec.MCCBundle(model=2, nreaction=1, nsigma=4, emin=-1.0, einterval=1.0, emax=100.0,
                              collision_ratio=1.0, rtype=np.array([103], np.int32), resultant=np.array([1], np.int32),
                              threshold=np.zeros(1), prob=np.ones((4, 1)), gas_mass=40 * AMU, gas_temperature=T)
    st = oracle.MccSetup(DX, DT, [-QE, QE], [ME, 31 * AMU], {(1, 0): cx(3000.0), (1, 1): cx(30.0)}, ng=2)
    n = 20000
    P = np.zeros((8, n))
    P[0], P[1], P[7] = 3.5, 2.5, 1.0
    P[2:5] = rng.standard_normal((3, n)) * 2e-3
    out, cnt = oracle.mcc(st, 4, 0, [np.zeros((8, 0)), P], counts=True)
    assert cnt[1, 0, 1] == n and cnt[1, 1, 1] == n
    v2 = ((out[1][2:5] * (DX / DT)) ** 2).sum(axis=0)
    e = 0.5 * 40 * AMU * v2 / (KB * 30.0)
    assert abs(e.mean() - 1.5) < 6 * e.std() / math.sqrt(n)


# ------------------------------------------------------- table interpolation

def _linear_raw(emin, eint, x, a, b):
    e = 10.0 ** (emin + np.asarray(x) * eint)
    return e, a + b * np.asarray(x)


@pytest.mark.parametrize("builder", ["oracle", "numpy"])
def test_sigma_interpolation_uses_dble_i_weight(oracle, builder):
    emin, eint, _, ns = xsec.grid_for_model(1)
    x = np.array([100.3, 1000.7, 2500.2, 6000.9, 9000.4])
    a, b = 1e-20, 1e-23
    e, s = _linear_raw(emin, eint, x, a, b)
    th = 10.0 ** (emin + 50.5 * eint)              # between grid energies 50 and 51
    r = xsec.Reaction(0, 11, 0, th, e, s)
    sig = (oracle.sigma_normalize([r], emin, eint, ns) if builder == "oracle"
           else xsec.sigma_normalize([r], emin, eint, ns))[:, 0]
    i = np.arange(1, ns + 1, dtype=np.float64)      # Fortran index of each grid energy
    E = xsec.energy_grid(emin, eint, ns)
    inside = (E > th) & (E < e[-1])
    np.testing.assert_allclose(sig[inside], a + b * i[inside], rtol=1e-9)
    assert not np.allclose(sig[inside], a + b * (i[inside] - 1), rtol=1e-6, atol=0.0)
    assert np.all(sig[E <= th] == 0.0)
    tail = E >= e[-1]
    np.testing.assert_allclose(sig[tail], s[-1] * e[-1] / E[tail], rtol=1e-13)


@pytest.mark.parametrize("builder", ["oracle", "numpy"])
def test_sigma_raw_energy_below_grid_takes_emin_coordinate(oracle, builder):
    emin, eint, _, ns = xsec.grid_for_model(1)
    x2 = 2000.25
    e = np.array([1e-4, 10.0 ** (emin + x2 * eint), 10.0 ** (emin + 8000.5 * eint)])
    s = np.array([1e-20, 3e-20, 5e-20])
    r = xsec.Reaction(0, 11, 0, 0.0, e, s)
    sig = (oracle.sigma_normalize([r], emin, eint, ns) if builder == "oracle"
           else xsec.sigma_normalize([r], emin, eint, ns))[:, 0]
    # the first raw point sits at index coordinate Emin = -3, not at log10(1e-4)
    for i in (1, 500, 1500, 2000):
        want = s[0] + (i - emin) * (s[1] - s[0]) / (x2 - emin)
        assert sig[i - 1] == pytest.approx(want, rel=1e-9)


def _vt_bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def test_twin_products_get_distinct_streams(oracle, rng):
    """Dissociation into two fragments of one species (Resultant 22) appends
    two products that differ only in the low bit of Vt's pattern, so that
    their state-keyed streams differ afterwards; the same for 113 giving back
    the parent's species.  Fragments of different species are left as the
    reference makes them."""
    e = xsec.log_grid()
    def electron(res):
        r = xsec.Reaction(0, 15, res, 1.0, e, np.where(e > 1.0, 1e-19, 0.0))
        gas = dict(mass=88.0043128 * AMU, temperature=300.0, density=1e23, natom=5.0, beta_max=3.0)
        b = xsec.build_bundle(xsec.SigmaFile(1, [r]), ME, 1.0, gas, DT)
        b.collision_ratio = 1.0
        return b
    n = 3000
    P = np.zeros((8, n))
    P[0], P[1], P[7] = 3.5, 2.5, 1.0
    d = rng.standard_normal((3, n))
    P[2:5] = d / np.linalg.norm(d, axis=0) * (math.sqrt(2 * 20.0 * QE / ME) / (DX / DT))
    for res, twin in ((22, True), (21, False)):
        st = oracle.MccSetup(DX, DT, [-QE, QE, -QE], [ME, 69 * AMU, 19 * AMU], {(0, 0): electron(res)})
        out, cnt = oracle.mcc(st, 6, 0, [P, np.zeros((8, 0)), np.zeros((8, 0))], counts=True)
        k = int(cnt[0, 0, 1])
        assert k > 100
        if twin:
            A, B = out[2][:, 0::2], out[2][:, 1::2]
            assert A.shape[1] == B.shape[1] == k
            np.testing.assert_array_equal(np.delete(A, 4, axis=0), np.delete(B, 4, axis=0))
            assert np.all((_vt_bits(A[4]) ^ _vt_bits(B[4])) == 1)
        else:
            np.testing.assert_array_equal(out[1], out[2])      # the same target in both species
    # 113 with Resultant = the parent's species: the product is the parent but for Vt's low bit
    b = _reactive_table(0.0, beta_max=1.0)
    b.rtype = np.array([101, 113], np.int32)
    b.resultant = np.array([1, 1], np.int32)
    st = oracle.MccSetup(DX, DT, [-QE, -QE], [ME, 19 * AMU], {(1, 0): b})
    Q = P.copy()
    Q[2:5] *= 0.05
    out, cnt = oracle.mcc(st, 7, 0, [np.zeros((8, 0)), Q], counts=True)
    assert cnt[1, 0, 2] == n and out[1].shape[1] == 2 * n
    A, B = out[1][:, :n], out[1][:, n:]
    np.testing.assert_array_equal(np.delete(A, 4, axis=0), np.delete(B, 4, axis=0))
    assert np.all((_vt_bits(A[4]) ^ _vt_bits(B[4])) == 1)
