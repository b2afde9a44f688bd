"""Volume recombination, R/src/mc/Recombination.F90:164-273
(executingRecombination2D, deletingParticles), CPU checks of the oracle on a
grid with dx = 1, dt = 1 (cell volume pi * (2 ir - 1), VFactor 1):

* densities and inverse mean energies against closed forms;
* a single cell with a hand-computed budget of 2.5 macro-particles: the two
  last-stored particles go deterministically, and the failed-draw retry
  makes a third deletion happen with probability 1 - (1/2)^2 = 3/4 (a
  restatement without the retry gives 1/2);
* unequal weights [1, 1, 3], budget 2.5: exactly two outcomes, {slot 2}
  with probability 5/6 and {slots 0, 1} with 1/6 -- the visiting order;
* the temperature form: budget n_e n_i dt V rate sqrt(1/T_e) / T_i (with
  1/T_e = 4 and 1/T_i = 1/4, swapped roles or a dropped sqrt change it);
* cells are independent: any interleaving that keeps each cell's own order
  gives the same survivors;
* a later reaction's pass skips particles an earlier one deleted;
* the host controller's schedule (every recb_time_step_num steps, over that
  many time steps) and its checks, on a recording stand-in.
"""
import numpy as np
import pytest

JTOEV = 1.6022e-19
NZ, NR, DX, DT = 5, 4, 1.0, 1.0


def _bundle(cells, weights, speeds=None, rng=None):
    """Particles in the given 1-based cells (iz, ir), strictly inside them."""
    rng = rng or np.random.default_rng(1)
    n = len(weights)
    P = np.zeros((8, n))
    for i, (iz, ir) in enumerate(cells):
        P[0, i] = iz - rng.uniform(0.05, 0.95)
        P[1, i] = ir - rng.uniform(0.05, 0.95)
    d = rng.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    v = np.ones(n) if speeds is None else np.asarray(speeds, dtype=np.float64)
    P[2:5] = d * v
    P[7] = weights
    return P


def test_density_and_mean_energy_closed_form(oracle):
    m = 2.0 * JTOEV                 # 0.5 m v^2 = v^2 eV with VFactor 1
    P = _bundle([(1, 1)] * 3 + [(2, 3)] * 2, [1.0, 2.0, 3.0, 4.0, 5.0], speeds=[2.0] * 3 + [0.5] * 2)
    d, t = oracle.recomb_density(NZ, NR, DX, DT, m, P)
    assert d[0, 0] == pytest.approx(6.0 / np.pi, rel=1e-15)
    assert d[1, 2] == pytest.approx(9.0 / (5 * np.pi), rel=1e-15)
    assert t[0, 0] == pytest.approx(1.0 / 4.0, rel=1e-14)
    assert t[1, 2] == pytest.approx(1.0 / 0.25, rel=1e-14)
    assert np.count_nonzero(d) == 2 and np.count_nonzero(t) == 2


def _two_species_cell(nA, nB, wA=None):
    A = _bundle([(1, 1)] * nA, np.ones(nA) if wA is None else wA, rng=np.random.default_rng(2))
    B = _bundle([(1, 1)] * nB, np.ones(nB), rng=np.random.default_rng(3))
    return A, B


def test_constant_rate_budget_and_retry(oracle):
    A, B = _two_species_cell(4, 2)
    # b = (4/pi)(2/pi) rate * 1 * pi = 8 rate / pi = 2.5
    rate = 2.5 * np.pi / 8.0
    outcomes = {}
    for seed in range(3000):
        (Ao, Bo), dl = oracle.recombine(NZ, NR, DX, DT, [A, B], [1.0, 1.0], [(0, 1, 0)], [rate], 1.0, seed, 0)
        assert dl[1] == 2 and Bo.shape[1] == 0
        assert dl[0] in (2, 3)
        left = {int(np.nonzero(A[0] == z)[0][0]) for z in Ao[0]}
        assert not left & {2, 3} and left <= {0, 1}
        outcomes[int(dl[0])] = outcomes.get(int(dl[0]), 0) + 1
    p3 = outcomes.get(3, 0) / 3000
    assert abs(p3 - 0.75) <= 5 * np.sqrt(0.75 * 0.25 / 3000), p3


def test_visiting_order_outcomes(oracle):
    A, B = _two_species_cell(3, 2, wA=[1.0, 1.0, 3.0])
    # dens_A = 5/pi, dens_B = 2/pi: b = 10 rate / pi = 2.5
    rate = 2.5 * np.pi / 10.0
    first = 0
    for seed in range(3000):
        (Ao, _), dl = oracle.recombine(NZ, NR, DX, DT, [A, B], [1.0, 1.0], [(0, 1, 0)], [rate], 1.0, seed, 7)
        left = sorted(Ao[7].tolist())
        assert left in ([1.0, 1.0], [3.0])
        first += left == [1.0, 1.0]
    assert abs(first / 3000 - 5 / 6) <= 5 * np.sqrt(5 / 36 / 3000)


def test_temperature_form(oracle):
    me, mi = 2.0 * JTOEV, 4.0 * JTOEV
    # electrons: 0.5 me v^2 = v^2 eV, v = 0.5 -> 1/T = 4; ions: 0.5 mi v^2 = 2 v^2 eV, v = sqrt(2) -> 1/T = 1/4
    E = _bundle([(1, 1)] * 4, np.ones(4), speeds=[0.5] * 4, rng=np.random.default_rng(4))
    I = _bundle([(1, 1)] * 2, np.ones(2), speeds=[np.sqrt(2.0)] * 2, rng=np.random.default_rng(5))
    # b = (4/pi)(2/pi) * 1 * pi * rate * sqrt(4) * (1/4) = 4 rate / pi = 2.5
    rate = 2.5 * np.pi / 4.0
    counts = []
    for seed in range(400):
        _, dl = oracle.recombine(NZ, NR, DX, DT, [E, I], [me, mi], [(0, 1, 1)], [rate], 1.0, seed, 0)
        assert dl[1] == 2 and dl[0] in (2, 3)
        counts.append(int(dl[0]))
    assert 0.6 < np.mean(np.array(counts) == 3) < 0.9


def test_cells_are_independent_of_interleaving(oracle):
    rng = np.random.default_rng(9)
    nz, nr = 17, 9
    cells = [(int(rng.integers(1, nz)), int(rng.integers(1, nr))) for _ in range(3000)]
    A = _bundle(cells, rng.uniform(0.5, 1.5, 3000), rng=rng)
    B = _bundle(cells[:1500], rng.uniform(0.5, 1.5, 1500), rng=rng)
    args = (nz, nr, DX, DT)
    (A1, B1), d1 = oracle.recombine(*args, [A, B], [1.0, 1.0], [(0, 1, 0)], [3.0], 1.0, 5, 2)
    assert d1[0] > 100 and d1[1] > 100
    cell = np.array([iz - 1 + (ir - 1) * (nz - 1) for iz, ir in cells])
    key = rng.permutation(cell.max() + 1)[cell]
    perm = np.argsort(key, kind="stable")               # regroup cells, keep each cell's order
    (A2, B2), d2 = oracle.recombine(*args, [A[:, perm], B], [1.0, 1.0], [(0, 1, 0)], [3.0], 1.0, 5, 2)
    assert tuple(d1) == tuple(d2)
    assert sorted(map(tuple, A1.T)) == sorted(map(tuple, A2.T))
    assert np.array_equal(B1, B2)


def test_later_reaction_skips_deleted(oracle):
    A = _bundle([(1, 1)] * 4, np.ones(4), rng=np.random.default_rng(6))
    B = _bundle([(1, 1)] * 50, np.ones(50), rng=np.random.default_rng(7))
    C = _bundle([(1, 1)] * 50, np.ones(50), rng=np.random.default_rng(8))
    # budgets 200 (4 * 50 / pi * rate) -> every A goes in the first reaction
    rate = 200.0 * np.pi / 200.0
    (Ao, Bo, Co), dl = oracle.recombine(NZ, NR, DX, DT, [A, B, C], [1.0] * 3, [(0, 1, 0), (0, 2, 0)],
                                        [rate, rate], 1.0, 1, 0)
    assert dl[0] == 4 and Ao.shape[1] == 0
    assert dl[1] == 50 and dl[2] == 50


def test_invalid_reaction_rejected(oracle):
    A, B = _two_species_cell(2, 2)
    with pytest.raises(ValueError):
        oracle.recombine(NZ, NR, DX, DT, [A, B], [1.0, 1.0], [(0, 2, 0)], [1.0], 1.0, 0, 0)
    with pytest.raises(ValueError):
        oracle.recombine(NZ, NR, DX, DT, [A, B], [1.0, 1.0], [(0, 1, 2)], [1.0], 1.0, 0, 0)


class _Recorder:
    ns = 3

    def __init__(self):
        self.calls = []

    def recombine(self, reactions, rates, time_step):
        self.calls.append((reactions, rates, time_step))
        return np.zeros(3, dtype=np.int64)


def test_controller_schedule():
    from ipm2d.recombination import Recombination
    hp = _Recorder()
    rc = Recombination(hp, [(1, 2, 0), (2, 0, 0), (0, 2, 1)], [1e-13, 1e-13, 3.95e-15], 50, 5e-11)
    assert rc.control(49) is None and rc.control(51) is None
    rc.control(50)
    rc.control(100)
    assert len(hp.calls) == 2
    assert hp.calls[0][2] == 50 * 5e-11
    assert hp.calls[0][0] == [(1, 2, 0), (2, 0, 0), (0, 2, 1)]
    with pytest.raises(ValueError):
        Recombination(hp, [(0, 3, 0)], [1.0], 50, 5e-11)
    with pytest.raises(ValueError):
        Recombination(hp, [(0, 1, 2)], [1.0], 50, 5e-11)
    Recombination(hp, [(0, 1, 0)], [1.0], 1, 1.0, enabled=False).control(0)
    assert len(hp.calls) == 2
