"""Pins the variogram-fit oracle (oracle/vgfit_oracle.py) with closed forms
on tiny inputs and parameter recovery from noise-free variograms."""
import math

import numpy as np
import pytest

from oracle import vgfit_oracle as V


def test_pairs_order_and_distances():
    x = np.array([0.0, 3.0, 0.0])
    y = np.array([0.0, 0.0, 4.0])
    z = np.array([1.0, 2.0, 5.0])
    pr = V.pairs(x, y, z)
    # pairs (i > j): (1,0) d=3, (2,0) d=4, (2,1) d=5, already ascending
    np.testing.assert_array_equal(pr, [[2, 1, 3], [5, 1, 4], [5, 2, 5]])


def test_clip_keeps_pairs_below_nearest_distance():
    pr = np.array([[0, 0, d] for d in (1.0, 2.0, 3.0, 4.0, 10.0)])
    # max 10, mdr 0.35 -> 3.5: nearest pair distance is 3.0 (first of the
    # two at 0.5), not beyond 3.5, so clip_ul = 3.5 and d < 3.5 stay
    out = V.clip_pairs(pr, 0.35)
    np.testing.assert_array_equal(out[:, 2], [1.0, 2.0, 3.0])
    # mdr 0.38 -> 3.8: nearest is 4.0 > 3.8 -> clip_ul = 4.0, d < 4 stay
    np.testing.assert_array_equal(V.clip_pairs(pr, 0.38)[:, 2], [1, 2, 3])
    # mdr 1 -> nearest is the last pair: nothing is dropped
    assert V.clip_pairs(pr, 1.0).shape[0] == 5


def test_robust_estimator_closed_form():
    # |dz| = 1 and 4: ((1 + 2) / 2)^4 / (0.457 + 0.494 / 2) / 2
    g = V.evg_robust(np.array([0.0, 0.0]), np.array([1.0, 4.0]))
    assert math.isclose(g, 1.5 ** 4 / (0.457 + 0.247) / 2, rel_tol=1e-15)
    assert V.evg_classic(np.array([0.0, 0.0]), np.array([1.0, 3.0])) == 5.0


def test_bins_of_nk_pairs():
    pr = np.array([[0.0, float(k), float(k + 1)] for k in range(7)])
    b = V.var_bins(pr, 3)
    assert b.shape == (3, 3)
    np.testing.assert_allclose(b[:, 0], [2.0, 5.0, 7.0])
    np.testing.assert_array_equal(b[:, 2], [3, 3, 1])


def test_candidates_follow_vgs_rules():
    c = V.candidate_sets(['Sph', 'Exp'], [1, 2], 'Nug')
    assert c == [('Sph',), ('Exp',), ('Nug',), ('Nug', 'Sph'), ('Nug', 'Exp')]
    # r clamped to the two models before the nugget joins
    assert V.candidate_sets(['Sph', 'Exp'], [5], 'Nug') == [('Nug', 'Sph'),
                                                         ('Nug', 'Exp')]


@pytest.mark.parametrize('name,r,s', [('Exp', 2.0e4, 3.0), ('Sph', 3.5e4, 1.2),
                                      ('Gau', 1.5e4, 0.7)])
def test_de_recovers_noise_free_parameters(name, r, s):
    h = np.linspace(500.0, 6e4, 60)
    g = V.FIT_FTNS[name](h, (r, s))
    from scipy.optimize import differential_evolution
    # bounds that contain the answer; the reference's (h0, 1e9) range bound
    # lets the energy-spread test stop DE early on saturated variograms
    res = differential_evolution(V.objective, [(1e3, 1e5), (1e-5, 2 * s)],
                                 args=((name,), h, g, None), maxiter=10000,
                                 popsize=100, polish=False, seed=3)
    assert res.fun < 1e-6 * np.sum(g ** 2)
    assert abs(res.x[0] / r - 1) < 2e-2 and abs(res.x[1] / s - 1) < 2e-2


# ---- model choice and output strings (vgs.py:771-846), hand-built values ----

def _fake(table):
    calls = []

    def solve(o):
        calls.append(o)
        return table[o]
    return solve, calls


def test_aic_choice_threshold_and_n_best():
    """tot_pts 100: Sph fun 2.0 -> AIC 100 ln 2 + 4 = 73.31 (first: always
    taken); Exp fun 1.99 -> 72.81, not below 0.99 x 73.31 = 72.58, so
    rejected; Gau fun 1.9 -> 68.19, taken."""
    cands = [('Sph',), ('Exp',), ('Gau',)]
    table = {('Sph',): (2.0, [1000.0, 1.0]), ('Exp',): (1.99, [900.0, 1.1]),
             ('Gau',): (1.9, [800.0, 1.2])}
    solve, calls = _fake(table)
    best, skip, _ = V.select_models(cands, solve, 100, 0.01, n_best=1)
    assert calls == cands and skip == []
    assert [b[0] for b in best] == [['Gau']]
    assert math.isclose(best[0][4], 100 * math.log(1.9) + 4, rel_tol=1e-15)
    # n_best = 2 keeps the last two accepted, oldest first (Exp rejected)
    best2, _, _ = V.select_models(cands, _fake(table)[0], 100, 0.01, 2)
    assert [b[0] for b in best2] == [['Sph'], ['Gau']]
    assert V.vg_strings(best2) == ['1.00000 Sph(1000.00000)',
                                   '1.20000 Gau(800.00000)']


def test_aic_rule_with_negative_aic_as_written():
    """With fun < 1 the AIC is negative and min_AIC (1 - 0.01) lies above
    min_AIC: the reference's rule then takes a candidate that is slightly
    worse.  Sph fun 0.5 -> AIC -65.31; Exp fun 0.51 -> -63.34, below
    0.99 x -65.31 = -64.66?  No: -63.34 > -64.66, rejected; Gau fun
    0.502 -> -64.91 < -64.66, taken although its fun is larger."""
    cands = [('Sph',), ('Exp',), ('Gau',)]
    table = {('Sph',): (0.5, [1.0, 1.0]), ('Exp',): (0.51, [1.0, 1.0]),
             ('Gau',): (0.502, [1.0, 1.0])}
    best, _, _ = V.select_models(cands, _fake(table)[0], 100, 0.01)
    assert [b[0] for b in best] == [['Gau']]


def test_out_of_order_ranges_are_sorted_and_skipped():
    """A nested candidate whose ranges descend is sorted (names and sills
    with it), accepted in sorted order, and its sorted permutation is not
    fitted when it comes up later."""
    cands = [('Nug', 'Sph'), ('Nug', 'Exp'), ('Sph', 'Nug')]
    table = {('Nug', 'Sph'): (1.0, [500.0, 0.1, 200.0, 2.0]),
             ('Nug', 'Exp'): (0.999, [500.0, 0.1, 900.0, 2.0]),
             ('Sph', 'Nug'): (0.1, [1.0, 1.0, 2.0, 2.0])}
    solve, calls = _fake(table)
    best, skip, _ = V.select_models(cands, solve, 50, 0.01, n_best=2)
    assert skip == [('Sph', 'Nug')]
    assert ('Sph', 'Nug') not in calls
    assert best[0][0] == ['Sph', 'Nug']
    np.testing.assert_array_equal(best[0][1], [200.0, 500.0])
    np.testing.assert_array_equal(best[0][2], [2.0, 0.1])
    # AIC 50 ln 1 + 8 = 8; Exp nested: 50 ln 0.999 + 8 = 7.95 > 0.99 x 8
    assert len(best) == 1
    assert V.vg_strings(best) == [
        '2.00000 Sph(200.00000) + 0.10000 Nug(500.00000)']
