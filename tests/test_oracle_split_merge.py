"""Weight rezoning, R/src/base/Rezoning.F90 (split_particle_weight,
merge_particle_random, particle_split_merge_control,
weighting_update_base_particle_number), CPU checks of the oracle and of the
host controller:

* known answers with exactly representable weights: CalRW's form (a swapped
  alpha changes the copy count), the copy count int(WQ/ref) - 1 at and just
  below the 2*ref boundary, remainders, copies equal to the parent but for
  WQ and the XOR tag in Vt's low bits;
* merge: strict threshold WQ < merge_limit*ref, the kept fraction equals
  WQ/ref within 5 binomial sigma, total weight conserved in expectation;
* the swap-with-last deletion visits every particle once: survivors are
  exactly the particles whose own draw keeps them, in any input order;
* the controller's schedule and base-weight rule on a recording stand-in.
"""
import struct

import numpy as np
import pytest

from ipm2d.rezoning import Rezoning, RezoningLimits

# ref = 8 * ((1 - 0.25) * 2 / 4 + 0.25) = 5 exactly
W, ALPHA, GRID_R, RPOS, REF = 8.0, 0.25, 4.0, 2.0, 5.0


def _parts(wq, r=RPOS, rng=None):
    wq = np.atleast_1d(np.asarray(wq, dtype=np.float64))
    n = wq.size
    rng = rng or np.random.default_rng(7)
    P = np.zeros((8, n))
    P[0] = rng.uniform(1, 60, n)
    P[1] = r
    P[2:7] = rng.normal(0, 0.1, (5, n))
    P[7] = wq
    return P


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def test_split_known_answer(oracle):
    P = _parts([17.5])
    Q, st = oracle.split_merge(P, W, ALPHA, GRID_R, 0.34, 1, 0, 0)
    # ratio 3.5 -> Ncr = 3 -> two copies of weight 5, remainder 7.5
    assert Q.shape[1] == 3 and tuple(st) == (2, 0, 0)
    np.testing.assert_array_equal(Q[7], [7.5, 5.0, 5.0])
    for j in (1, 2):
        for f in (0, 1, 2, 3, 5, 6):
            assert Q[f, j] == P[f, 0]
        assert _bits(Q[4, j]) == _bits(P[4, 0]) ^ j
    assert Q[:, 0].tolist()[:7] == P[:, 0].tolist()[:7]


def test_split_boundaries(oracle):
    below = np.nextafter(2 * REF, 0.0)
    P = _parts([2 * REF, below, REF, 0.99 * REF, 10 * REF])
    Q, st = oracle.split_merge(P, W, ALPHA, GRID_R, 0.34, 1, 0, 0, mode=1)
    # 2 ref -> one copy (int(2.0) = 2 > 1); just below -> none; 10 ref -> nine
    assert st[0] == 1 + 9
    np.testing.assert_array_equal(Q[7, :5], [REF, below, REF, 0.99 * REF, REF])
    assert np.all(Q[7, 5:] == REF)
    assert Q[7].sum() == pytest.approx(P[7].sum(), rel=1e-15)


def test_radial_reference_weight(oracle):
    # at R = 0 the reference weight is W * alpha, at R = grid_r it is W
    P = _parts([3.0 * W * ALPHA, 3.0 * W], r=np.array([0.0, GRID_R]))
    Q, st = oracle.split_merge(P, W, ALPHA, GRID_R, 0.34, 1, 0, 0, mode=1)
    assert st[0] == 4
    np.testing.assert_allclose(np.sort(Q[7]), np.sort([W * ALPHA] * 3 + [W] * 3), rtol=1e-15)


def test_merge_threshold_is_strict(oracle):
    lim = 0.34
    at = lim * REF
    P = _parts([at, np.nextafter(at, 0.0), REF])
    Q, st = oracle.split_merge(P, W, ALPHA, GRID_R, lim, 3, 0, 0, mode=2)
    assert st[1] + st[2] == 1
    assert at in Q[7] and REF in Q[7]


def test_merge_keep_fraction_and_expected_weight(oracle):
    rng = np.random.default_rng(11)
    n = 200_000
    frac = rng.uniform(0.0, 0.34, n)
    P = _parts(frac * REF, rng=rng)
    Q, st = oracle.split_merge(P, W, ALPHA, GRID_R, 0.34, 5, 9, 1, mode=2)
    kept, deleted = int(st[1]), int(st[2])
    assert kept + deleted == n and Q.shape[1] == kept
    assert np.all(Q[7] == REF)
    mean = frac.sum()
    sigma = np.sqrt((frac * (1 - frac)).sum())
    assert abs(kept - mean) <= 5 * sigma
    # expected weight after = sum of WQ before
    assert abs(Q[7].sum() - P[7].sum()) <= 5 * sigma * REF


def test_merge_survivors_are_the_draws_own(oracle):
    rng = np.random.default_rng(3)
    n = 400
    frac = np.concatenate([rng.uniform(0, 0.34, n), rng.uniform(0.5, 1.5, 50)])
    P = _parts(frac * REF, rng=rng)
    seed, step, s = 123, 4, 2
    keep = []
    for i in range(P.shape[1]):
        if P[7, i] < 0.34 * REF:
            k = oracle.state_key(*P[:5, i])
            u = oracle.uniform(seed, step, 0x4D520000 | s, k[0], k[1], k[2], 0)
            if u < P[7, i] / REF:
                keep.append(i)
        else:
            keep.append(i)
    want = {tuple(P[:7, i]) for i in keep}
    for perm in (np.arange(P.shape[1]), rng.permutation(P.shape[1])):
        Q, st = oracle.split_merge(P[:, perm], W, ALPHA, GRID_R, 0.34, seed, step, s, mode=2)
        assert {tuple(Q[:7, i]) for i in range(Q.shape[1])} == want
        assert Q.shape[1] == len(keep)


def test_split_copies_draw_independently(oracle):
    P = _parts([6 * REF])
    Q, _ = oracle.split_merge(P, W, ALPHA, GRID_R, 0.34, 1, 0, 0, mode=1)
    keys = {tuple(oracle.state_key(*Q[:5, i])) for i in range(Q.shape[1])}
    assert len(keys) == Q.shape[1] == 6
    u = [oracle.uniform(1, 0, 0, *oracle.state_key(*Q[:5, i]), 0) for i in range(6)]
    assert len(set(u)) == 6


def test_split_then_merge_are_disjoint(oracle):
    rng = np.random.default_rng(5)
    frac = np.concatenate([rng.uniform(0.01, 0.33, 300), rng.uniform(2.0, 5.0, 300), rng.uniform(0.4, 1.9, 300)])
    P = _parts(frac * REF, rng=rng)
    A, sa = oracle.split_merge(P, W, ALPHA, GRID_R, 0.34, 8, 1, 0, mode=3)
    B, sb = oracle.split_merge(P, W, ALPHA, GRID_R, 0.34, 8, 1, 0, mode=1)
    B, sb2 = oracle.split_merge(B, W, ALPHA, GRID_R, 0.34, 8, 1, 0, mode=2)
    assert sa[0] == sb[0] and sa[1] == sb2[1] and sa[2] == sb2[2]
    assert sorted(map(tuple, A.T)) == sorted(map(tuple, B.T))


class _Recorder:
    ns = 3

    def __init__(self, counts):
        self.counts = counts
        self.calls = []

    def count(self, s):
        return self.counts[s]

    def split_merge(self, s, weight, alpha, grid_r, merge_limit):
        self.calls.append((s, weight, alpha, grid_r, merge_limit))
        return (0, 0, 0)

    def set_step_rezone(self, s, weight, alpha, grid_r, merge_limit):
        self.calls.append(("step", s, weight, alpha, grid_r, merge_limit))


def test_controller_schedule_and_base_weight():
    nz, nr, ppg = 11, 9, 10             # 80 cells * 10 = 800 reference particles
    hp = _Recorder([900, 700, 50])      # above 808, below 792, below the 80 limit
    rz = Rezoning(hp, nz, nr, [1.0, 2.0, 3.0], [ppg] * 3, alpha=0.01, period=4)
    rz.control(1)
    assert [c[0] for c in hp.calls] == [0]
    assert rz.weight == [1.0, 2.0, 3.0]
    rz.control(4)
    assert [c[0] for c in hp.calls] == [0, 0, 1, 2]
    assert rz.weight == [1.0 * 1.01, 2.0 * 0.99, 3.0]
    # electrons split/merged with the weight they had before the update
    assert hp.calls[1][1] == 1.0 and hp.calls[2][1] == 2.0 * 0.99
    assert all(c[3] == nr - 1 and c[4] == 0.34 for c in hp.calls)
    hp.calls.clear()
    Rezoning(hp, nz, nr, [1.0] * 3, [ppg] * 3, limits=RezoningLimits(is_open_split_merge=False)).control(0)
    assert hp.calls == []


def test_controller_with_electrons_in_the_step():
    """fused_electrons: the electrons' split/merge is handed to the step once
    (and again with each new Weight every period, taking effect from the
    next step), never run as a separate call; the ions keep theirs."""
    nz, nr, ppg = 11, 9, 10
    hp = _Recorder([900, 700, 50])
    rz = Rezoning(hp, nz, nr, [1.0, 2.0, 3.0], [ppg] * 3, alpha=0.01, period=4, fused_electrons=True)
    assert hp.calls == [("step", 0, 1.0, 0.01, nr - 1, 0.34)]
    hp.calls.clear()
    rz.control(1)
    assert hp.calls == []
    rz.control(4)
    assert hp.calls[0] == ("step", 0, 1.0 * 1.01, 0.01, nr - 1, 0.34)
    assert [c[0] for c in hp.calls[1:]] == [1, 2]
