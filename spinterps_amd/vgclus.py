"""The kriging-weight heuristic of the variogram clustering
(vgclus/tvgs_clus.py:92-137) on the GPU class API.

``sorted_kriging_weights`` gives the ordinary-kriging weights of a random
station configuration for its first point (the class API's formulation,
pyx:251-362), sorted ascending, with the optional cut of small weights and
rescaling that the clustering applies before comparing configurations.
"""
from __future__ import annotations

import numpy as np

from .cyth import OrdinaryKriging


def _threshold(wts, min_kg_abs_wt):
    """tvgs_clus.py:109-130: weights whose share of sum |w| is at most
    min_kg_abs_wt become 0 (unless that is all of them); the rest are
    rescaled to sum to one."""
    if min_kg_abs_wt:
        abs_wts = np.abs(wts)
        rel_wts = abs_wts / abs_wts.sum()
        zero_idxs = rel_wts <= min_kg_abs_wt
        if zero_idxs.sum() == zero_idxs.size:
            zero_idxs = np.zeros_like(zero_idxs, dtype=bool)
        non_zero_wts = wts[~zero_idxs]
        sclr = 1.0 / non_zero_wts.sum()
        assert zero_idxs.sum() < zero_idxs.size
        wts[zero_idxs] = 0.0
        wts[~zero_idxs] *= sclr
    assert (wts[1:] - wts[:-1]).min() >= 0.0, (wts[1:] - wts[:-1]).min()
    assert np.isclose(wts.sum(), 1.0)
    return wts


def sorted_kriging_weights(tvg_str, rnd_pts, min_kg_abs_wt=None):
    """rnd_pts [n+1 x 2]: the reference point first, then the n stations.
    Returns the n weights sorted ascending (tvgs_clus.py:92-137)."""
    rnd_pts = np.ascontiguousarray(rnd_pts, dtype=np.float64)
    assert rnd_pts.ndim == 2 and rnd_pts.shape[1] >= 2
    ok = OrdinaryKriging(
        xi=np.ascontiguousarray(rnd_pts[1:, 0]),
        yi=np.ascontiguousarray(rnd_pts[1:, 1]),
        zi=np.zeros(rnd_pts.shape[0] - 1),
        xk=np.array([rnd_pts[0, 0]]),
        yk=np.array([rnd_pts[0, 1]]),
        model=tvg_str)
    ok.krige()
    wts = ok.lambdas.ravel().copy()
    wts.sort()
    return _threshold(wts, min_kg_abs_wt)
