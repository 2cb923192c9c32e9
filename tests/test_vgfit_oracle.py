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
