"""Cross-section table build (host side): the package's numpy builder against
the oracle's C restatement, plus closed forms the algorithm fixes.

* constant sigma, Model 1: Probility(1,i) = v(E_i)/v_max below the last raw
  energy, the 1/E tail beyond it, CollisionRatio = 1 - exp(-n sigma v_max dt)
* reactive fractions (E - Th)^S / sum with S = (3 N - 6)/2 - 1
* Langevin ratio sqrt(pi alpha e^2/(eps0 mu)) beta^2 n dt
* locate against bisect; namelist write/parse round trip
"""
import bisect
import math

import numpy as np
import pytest

from ipm2d import xsec

ME = xsec.ELECTRON_MASS
AMU = xsec.ATOMIC_MASS


def test_locate_matches_bisect(oracle, rng):
    xx = np.sort(rng.uniform(0, 100, 50))
    for x in rng.uniform(-5, 105, 300).tolist() + [xx[0], xx[-1], xx[10]]:
        k = xsec.locate(xx, xx.size, x)
        assert k == oracle.locate(xx, x)
        if xx[0] < x < xx[-1]:
            assert k == bisect.bisect_right(xx.tolist(), x)


def _electron_reactions():
    return xsec.synthetic_electron_reactions(excitation=(7.0, 2e-20), dissociation=(12.0, 1e-20, 41))


def test_sigma_normalize_matches_oracle(oracle):
    reacs = _electron_reactions()
    emin, eint, _, ns = xsec.grid_for_model(1)
    a = xsec.sigma_normalize(reacs, emin, eint, ns)
    b = oracle.sigma_normalize(reacs, emin, eint, ns)
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-40)


def test_probability_matches_oracle(oracle):
    reacs = _electron_reactions()
    emin, eint, _, ns = xsec.grid_for_model(1)
    sig = xsec.sigma_normalize(reacs, emin, eint, ns)
    n = xsec.gas_density(50.0, 300.0)
    pa, ma = xsec.probability_nonreactive(sig, emin, eint, ME, 88.0043128 * AMU, n)
    pb, mb = oracle.probability_nonreactive(sig, emin, eint, ME, 88.0043128 * AMU, n)
    np.testing.assert_allclose(pa, pb, rtol=1e-13)
    assert ma == pytest.approx(mb, rel=1e-14)


def test_reactive_matches_oracle(oracle):
    th = [0.0, 3.521, 5.621, 7.5]
    emin, eint, _, ns = xsec.grid_for_model(3)
    a = xsec.sigma_normalize_reactive(th, 1.0, 5.0, emin, eint, ns)
    b = oracle.sigma_normalize_reactive(th, 1.0, 5.0, emin, eint, ns)
    np.testing.assert_allclose(a, b, rtol=1e-13)
    n = xsec.gas_density(50.0, 300.0)
    pa, ra = xsec.probability_reactive(a, 19 * AMU, 88 * AMU, n, 3.0, xsec.polarizability(), 5e-11)
    pb, rb = oracle.probability_reactive(a, 19 * AMU, 88 * AMU, n, 3.0, xsec.polarizability(), 5e-11)
    np.testing.assert_allclose(pa, pb, rtol=1e-14)
    assert ra == pytest.approx(rb, rel=1e-14)


def test_constant_sigma_closed_form():
    e = xsec.log_grid()
    s0 = 1e-19
    sf = xsec.SigmaFile(1, [xsec.Reaction(0, 11, 0, 0.0, e, np.full(e.size, s0))])
    gas = dict(mass=88.0043128 * AMU, temperature=300.0, density=1e21, natom=5.0, beta_max=3.0)
    dt = 5e-11
    b = xsec.build_bundle(sf, ME, 1.0, gas, dt)
    emin, eint, _, ns = xsec.grid_for_model(1)
    E = xsec.energy_grid(emin, eint, ns)
    mu = ME * gas["mass"] / (ME + gas["mass"])
    v = np.sqrt(2 * E * xsec.JTOEV / mu)
    emax_raw = e[-1]
    sigma = np.where(E < emax_raw, s0, s0 * emax_raw / E)
    # interior of the raw range: the interpolation of a constant is the constant
    inside = (E > 0.012) & (E < 99.0)
    nu = sigma * v * gas["density"]
    np.testing.assert_allclose(b.prob[inside, 0], nu[inside] / nu.max(), rtol=1e-10)
    assert b.collision_ratio == pytest.approx(1 - math.exp(-nu.max() * dt), rel=1e-9)


def test_reactive_fraction_closed_form():
    th = np.array([0.0, 2.0, 5.0])
    emin, eint, _, ns = xsec.grid_for_model(3)
    sig = xsec.sigma_normalize_reactive(th, 4.0, 5.0, emin, eint, ns)
    S = (3 * 9 - 6) / 2 - 1
    E = xsec.energy_grid(emin, eint, ns)
    k = np.searchsorted(E, 20.0)
    w = np.maximum(E[k] - th, 0.0) ** S
    np.testing.assert_allclose(sig[k], w / w.sum(), rtol=1e-12)
    np.testing.assert_allclose(sig.sum(axis=1), 1.0, rtol=1e-13)


def test_langevin_ratio_closed_form():
    n = xsec.gas_density(50.0, 300.0)
    mu = 19 * AMU * 88 * AMU / (107 * AMU)
    alpha = 14 * xsec.A0 ** 3
    sig = np.ones((xsec.NSIGMA_ION, 2)) * 0.5
    _, r = xsec.probability_reactive(sig, 19 * AMU, 88 * AMU, n, 3.0, alpha, 5e-11)
    ref = math.sqrt(math.pi * alpha * xsec.ELECTRON_CHARGE ** 2 / (xsec.EPSILON0 * mu)) * 9.0 * n * 5e-11
    assert r == pytest.approx(ref, rel=1e-13)


def test_namelist_round_trip():
    reacs = xsec.synthetic_electron_reactions()
    text = xsec.write_sigma_file(1, reacs)
    sf = xsec.parse_sigma_file(text)
    assert sf.model == 1 and len(sf.reactions) == len(reacs)
    for a, b in zip(sf.reactions, reacs):
        assert (a.rtype, a.resultant, a.threshold) == (b.rtype, b.resultant, b.threshold)
        np.testing.assert_array_equal(a.energy, b.energy)
        np.testing.assert_array_equal(a.sigma, b.sigma)
    ions = [xsec.Reaction(2, 101, 2, 0.0), xsec.Reaction(2, 111, 2, 5.621), xsec.Reaction(2, 100, 0, 5.843)]
    sf3 = xsec.parse_sigma_file(xsec.write_sigma_file(3, ions))
    assert [(r.rtype, r.resultant, r.threshold) for r in sf3.reactions] == [(101, 2, 0.0), (111, 2, 5.621),
                                                                           (100, 0, 5.843)]


def test_parses_reference_style_namelist():
    """The format of input/gas/lxcat/CF4/*.txt: comments with '!', 'd' exponents,
    tabs, the array form SR%Reaction= for reactive sets."""
    text = """&MRList
Model     =     3,
NReaction  =    3,
/
&SRList
SR%Reaction=
1,101,1,0.d0        ! 1 Isotropic elastic
1,113,0,3.521d0     ! detachment
1,111,1,5.621d0
/"""
    sf = xsec.parse_sigma_file(text)
    assert sf.model == 3 and [r.rtype for r in sf.reactions] == [101, 113, 111]
    assert sf.reactions[1].threshold == 3.521
    text1 = """&MRList
Model = 1,
NReaction = 1,
/
!comment
&SRList
SR(1)%Reaction            =           0,11,0,6.2293000000d-6,
SR(1)%EnergySigma         =
 0.000000e+0\t7.144860e-19
 1.110000e-3\t7.144860e-19
 1.230000e-3\t6.462220e-19
/"""
    sf1 = xsec.parse_sigma_file(text1)
    r = sf1.reactions[0]
    assert r.threshold == 6.2293e-6 and r.energy.size == 3 and r.sigma[2] == 6.46222e-19


def _hs_gas(ns):
    return dict(mass=88.0043128 * AMU, temperature=300.0, density=xsec.gas_density(50.0, 300.0), radius=1.548e-10,
                ns=ns, index_start=0, beta_max=3.0)


@pytest.mark.parametrize("ns", [1, 3])
def test_hard_sphere_closed_form(oracle, ns):
    """ProbilityHardShpere: sigma = 2 pi ((r_g + r_s)/2)^2, Probility = base *
    sigma v(E) n / max with v = sqrt(2 E e / mu), so the top row is 1 at the
    grid's last energy and the shape is sqrt(E / E_max); two reactions (101,
    103) with base 0.5 and 1.0 for a one-ion gas, one (101) otherwise."""
    ms, rs, dt = 69.0059096 * AMU, 2e-10, 5e-11
    gas = _hs_gas(ns)
    b = xsec.hard_sphere_bundle(1, ms, rs, gas, dt)
    emin, eint, emax, nsig = xsec.grid_for_model(3)
    E = xsec.energy_grid(emin, eint, nsig)
    mu = ms * gas["mass"] / (ms + gas["mass"])
    sigma = math.pi * (rs + gas["radius"]) ** 2 / 2.0
    vmax = math.sqrt(2.0 * 10.0 ** emax * xsec.JTOEV / mu)
    assert b.sigma_max == pytest.approx(sigma * vmax * gas["density"], rel=1e-12)
    assert b.collision_ratio == pytest.approx(1.0 - math.exp(-sigma * vmax * gas["density"] * dt), rel=1e-12)
    shape = np.sqrt(E / E[-1])
    if ns == 1:
        assert list(b.rtype) == [101, 103] and list(b.resultant) == [1, 1]
        np.testing.assert_allclose(b.prob[:, 1], shape, rtol=1e-12)
        np.testing.assert_allclose(b.prob[:, 0], 0.5 * shape, rtol=1e-12)
    else:
        assert list(b.rtype) == [101] and list(b.resultant) == [1]
        np.testing.assert_allclose(b.prob[:, 0], shape, rtol=1e-12)
    assert b.model == 2 and b.nsigma == 1001
    p, mx = oracle.probability_hard_sphere(b.nreaction, emin, eint, nsig, ms, gas["mass"], gas["density"], rs,
                                           gas["radius"])
    np.testing.assert_allclose(p, b.prob, rtol=1e-13)
    assert mx == pytest.approx(b.sigma_max, rel=1e-14)


def test_hard_sphere_events_split_evenly(oracle, rng):
    """With the 0.5 / 1.0 table a selected ion's elastic and charge-exchange
    events are equally likely at every energy (SelectProbility,
    TypeMCC.f90:173-216), and the selected fraction is the CollisionRatio."""
    ms = 9.012182 * AMU
    gas = dict(mass=ms, temperature=300.0, density=xsec.gas_density(5000.0, 300.0), radius=112e-12, ns=1,
               index_start=0, beta_max=0.0)
    b = xsec.hard_sphere_bundle(1, ms, 112e-12, gas, 5e-9)
    st = oracle.MccSetup(1e-3, 5e-9, [-xsec.ELECTRON_CHARGE, xsec.ELECTRON_CHARGE], [ME, ms], {(1, 0): b})
    n = 200000
    P = np.zeros((8, n))
    P[0], P[1], P[7] = 3.5, 2.5, 1.0
    P[2:5] = rng.standard_normal((3, n)) * (5000.0 / (1e-3 / 5e-9))
    _, cnt = oracle.mcc(st, 3, 0, [np.zeros((8, 0)), P], counts=True)
    sel = cnt[1, 0].sum()
    p = b.collision_ratio
    assert abs(sel - n * p) < 6 * math.sqrt(n * p * (1 - p))
    n1, n2 = cnt[1, 0, 1], cnt[1, 0, 2]
    assert n1 + n2 > 2000
    assert abs(n1 - n2) < 6 * math.sqrt(n1 + n2)
