"""Drop-in for the reference's ``cyth`` module (cyth/__init__.py:9-21).

Same names, argument order, dtypes and in-place semantics as the cpdefs of
cyth/interpmthds.pyx; each call runs on the GPU through libspk.so.  Arrays
must be C-contiguous like the reference's typed memoryviews ``DT_D[::1]``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, context, dptr, i64ptr, lib, u64ptr

__all__ = ['fill_dists_one_pt', 'fill_wts_and_sum', 'get_mults_sum',
           'sel_equidist_refs', 'fill_dists_2d_mat', 'fill_vg_var_arr',
           'fill_theo_vg_vals', 'copy_2d_arr_at_idxs', 'get_nd_dists',
           'OrdinaryKriging', 'SimpleKriging', 'ExternalDriftKriging_MD']


def _ctx():
    return context().h


def _f64(a, name):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64
            and a.flags['C_CONTIGUOUS']):
        raise ValueError(f'{name}: expected a C-contiguous float64 array')
    return a


def fill_dists_2d_mat(x1s, y1s, x2s, y2s, dists):
    """pyx:123-143."""
    for a, nm in ((x1s, 'x1s'), (y1s, 'y1s'), (x2s, 'x2s'), (y2s, 'y2s'),
                  (dists, 'dists')):
        _f64(a, nm)
    assert x1s.shape == y1s.shape and x2s.shape == y2s.shape
    assert dists.shape == (x1s.size, x2s.size)
    check(lib().spk_fill_dists_2d_mat(_ctx(), dptr(x1s), dptr(y1s), x1s.size,
                                      dptr(x2s), dptr(y2s), x2s.size,
                                      dptr(dists)))


def fill_vg_var_arr(dists, in_vars, covar_flag, diag_mat_flag,
                    vg_models_str, min_vg_val):
    """pyx:146-226."""
    _f64(dists, 'dists')
    _f64(in_vars, 'in_vars')
    assert dists.ndim == 2 and in_vars.shape == dists.shape
    check(lib().spk_fill_vg_var_arr(
        _ctx(), dptr(dists), dists.shape[0], dists.shape[1], dptr(in_vars),
        int(covar_flag), int(diag_mat_flag), str(vg_models_str).encode(),
        float(min_vg_val)))


def copy_2d_arr_at_idxs(arr, row_idxs, col_idxs, subset_arr):
    """pyx:229-248 (int64 indices; only the leading block is written)."""
    _f64(arr, 'arr')
    _f64(subset_arr, 'subset_arr')
    row_idxs = np.ascontiguousarray(row_idxs, dtype=np.int64)
    col_idxs = np.ascontiguousarray(col_idxs, dtype=np.int64)
    check(lib().spk_copy_2d_arr_at_idxs(
        _ctx(), dptr(arr), arr.shape[0], arr.shape[1], i64ptr(row_idxs),
        row_idxs.size, i64ptr(col_idxs), col_idxs.size, dptr(subset_arr),
        subset_arr.shape[0], subset_arr.shape[1]))


def fill_dists_one_pt(x, y, xs, ys, dists):
    """pyx:768-781."""
    _f64(xs, 'xs')
    _f64(ys, 'ys')
    _f64(dists, 'dists')
    check(lib().spk_fill_dists_one_pt(_ctx(), float(x), float(y), dptr(xs),
                                      dptr(ys), xs.size, dptr(dists)))


def fill_wts_and_sum(dists, wts, idw_exp):
    """pyx:784-795: wts = 1/d^p in place; returns their sum."""
    _f64(dists, 'dists')
    _f64(wts, 'wts')
    out = ctypes.c_double()
    check(lib().spk_fill_wts_and_sum(_ctx(), dptr(dists), dptr(wts),
                                     dists.size, float(idw_exp),
                                     ctypes.byref(out)))
    return out.value


def get_mults_sum(wts, data):
    """pyx:798-808."""
    _f64(wts, 'wts')
    _f64(data, 'data')
    out = ctypes.c_double()
    check(lib().spk_get_mults_sum(_ctx(), dptr(wts), dptr(data), wts.size,
                                  ctypes.byref(out)))
    return out.value


def fill_theo_vg_vals(vg_str, h_arr, r, s, vg_arr):
    """pyx:98-120: vg_arr += model(h, r, s)."""
    _f64(h_arr, 'h_arr')
    _f64(vg_arr, 'vg_arr')
    assert h_arr.shape[0] == vg_arr.shape[0]
    check(lib().spk_fill_theo_vg_vals(_ctx(), str(vg_str).encode(),
                                      dptr(h_arr), h_arr.shape[0], float(r),
                                      float(s), dptr(vg_arr)))


def get_nd_dists(pts):
    """pyx:893-923: condensed pairwise distances (i > j)."""
    pts = _f64(pts, 'pts')
    n = pts.shape[0]
    out = np.full(n * (n - 1) // 2, np.nan)
    check(lib().spk_get_nd_dists(_ctx(), dptr(pts), n, pts.shape[1],
                                 dptr(out)))
    return out


def sel_equidist_refs(dst_x, dst_y, ref_xs, ref_ys, n_pies, min_dist_thresh,
                      not_neb_flag, dists, tem_ref_sel_dists,
                      ref_sel_pie_idxs, ref_pie_idxs, ref_pie_cts):
    """pyx:811-890 (DT_UL arrays are uint64 as on LP64 platforms)."""
    for a, nm in ((ref_xs, 'ref_xs'), (ref_ys, 'ref_ys'), (dists, 'dists'),
                  (tem_ref_sel_dists, 'tem_ref_sel_dists')):
        _f64(a, nm)
    check(lib().spk_sel_equidist_refs(
        _ctx(), float(dst_x), float(dst_y), dptr(ref_xs), dptr(ref_ys),
        ref_xs.size, int(n_pies), float(min_dist_thresh), int(not_neb_flag),
        dptr(dists), dptr(tem_ref_sel_dists), i64ptr(ref_sel_pie_idxs),
        u64ptr(ref_pie_idxs), u64ptr(ref_pie_cts)))



# ---- class API (pyx:251-719) ------------------------------------------------
class _KrigeClass:
    """Common part of the kriging classes: inputs as the reference stores
    them, outputs filled by krige() on the GPU."""

    _kind = 0

    def _run(self, si=None, sk=None):
        xi = np.ascontiguousarray(self.xi, np.float64)
        yi = np.ascontiguousarray(self.yi, np.float64)
        zi = np.ascontiguousarray(self.zi, np.float64)
        xk = np.ascontiguousarray(self.xk, np.float64)
        yk = np.ascontiguousarray(self.yk, np.float64)
        n, C = xi.size, xk.size
        m = 0 if si is None else si.shape[0]
        P = n + (0 if self._kind == 1 else 1 + m)
        zk = np.zeros(C)
        ev = np.zeros(C)
        lam = np.zeros((C, n))
        mus = np.zeros((1 + m, C)) if self._kind == 2 else np.zeros(C)
        rhss = np.zeros((C, P))
        kin = np.zeros((P, P))
        kinv = np.zeros((P, P))
        dists = np.zeros((n, n))
        if si is not None:
            si = np.ascontiguousarray(si, np.float64)
            sk = np.ascontiguousarray(sk, np.float64)
        check(lib().spk_krige_class(
            _ctx(), self._kind, self.model, dptr(xi), dptr(yi), dptr(zi), n,
            dptr(si) if si is not None else None, m, dptr(xk), dptr(yk),
            dptr(sk) if sk is not None else None, C, dptr(zk),
            dptr(ev) if self._kind != 2 else None, dptr(lam),
            dptr(mus) if self._kind != 1 else None, dptr(rhss), dptr(kin),
            dptr(kinv), dptr(dists)))
        return zk, ev, lam, mus, rhss, kin, kinv, dists


class OrdinaryKriging(_KrigeClass):
    """pyx:251-362: zero-diagonal semivariogram matrix, right-hand sides
    that skip stations at distance 0, pinv-equivalent inverse, variance
    max(0, lambda.gamma + mu)."""

    _kind = 0

    def __init__(self, xi, yi, zi, xk, yk, model='1.0 Sph(2)'):
        self.xi, self.yi, self.zi, self.xk, self.yk = xi, yi, zi, xk, yk
        self.model = bytes(model, 'utf-8')
        self.in_count = self.xi.shape[0]
        self.out_count = self.xk.shape[0]
        n, C = self.in_count, self.out_count
        self.zk = np.zeros(C)
        self.in_dists = np.zeros((n, n))
        self.mus = np.zeros(C)
        self.est_vars = np.zeros(C)
        self.lambdas = np.zeros((C, n))
        self.in_vars = np.zeros((n + 1, n + 1))
        self.in_vars_inv = self.in_vars.copy()
        self.rhss = np.zeros((C, n + 1))

    def krige(self):
        (self.zk, self.est_vars, self.lambdas, self.mus, self.rhss,
         self.in_vars, self.in_vars_inv, self.in_dists) = self._run()


class SimpleKriging(_KrigeClass):
    """pyx:365-471, including its order of initialisation: the covariance
    matrix is filled before the sills are summed, so it is -Gamma with a
    zero diagonal; est_covars = max(0, covar - lambda.rhs)."""

    _kind = 1

    def __init__(self, xi, yi, zi, xk, yk, model='1.0 Sph(2)'):
        self.xi, self.yi, self.zi, self.xk, self.yk = xi, yi, zi, xk, yk
        self.model = bytes(model, 'utf-8')
        self.in_count = self.xi.shape[0]
        self.out_count = self.xk.shape[0]
        n, C = self.in_count, self.out_count
        self.zk = np.zeros(C)
        self.in_dists = np.zeros((n, n))
        self.est_covars = np.zeros(C)
        self.lambdas = np.zeros((C, n))
        self.in_covars = np.zeros((n, n))
        self.in_covars_inv = self.in_covars.copy()
        self.out_covars = np.zeros((C, n))
        self.rhss = np.zeros((C, n))

    def krige(self):
        (self.zk, self.est_covars, self.lambdas, _, self.rhss,
         self.in_covars, self.in_covars_inv, self.in_dists) = self._run()
        self.out_covars = self.rhss


class ExternalDriftKriging_MD(_KrigeClass):
    """pyx:588-719: OK system with a [1 | drifts] border; mus_arr holds the
    m + 1 Lagrange multipliers per output point."""

    _kind = 2

    def __init__(self, xi, yi, zi, si, xk, yk, sk, model='1.0 Sph(2)'):
        assert len(xi.shape) == 1
        assert len(si.shape) == 2
        assert xi.shape[0] == yi.shape[0] == zi.shape[0] == si.shape[1], \
            'Observation points and drift shapes are unequal!'
        assert xk.shape[0] == yk.shape[0] == sk.shape[1], \
            'Resulting points and drifts shapes are unequal!'
        assert si.shape[0] == sk.shape[0], \
            'Observation and reulting drifts have unequal shapes!'
        self.xi, self.yi, self.zi, self.si = xi, yi, zi, si
        self.xk, self.yk, self.sk = xk, yk, sk
        self.model = bytes(model, 'utf-8')
        self.n_drifts = sk.shape[0]
        self.in_count = self.xi.shape[0]
        self.out_count = self.xk.shape[0]
        n, C, m = self.in_count, self.out_count, self.n_drifts
        self.zk = np.zeros(C)
        self.in_dists = np.zeros((n, n))
        self.mus_arr = np.zeros((m + 1, C))
        self.lambdas = np.zeros((C, n))
        self.in_vars = np.zeros((n + m + 1, n + m + 1))
        self.in_vars_inv = self.in_vars.copy()
        self.rhss = np.zeros((C, n + m + 1))

    def krige(self):
        (self.zk, _, self.lambdas, self.mus_arr, self.rhss, self.in_vars,
         self.in_vars_inv, self.in_dists) = self._run(self.si, self.sk)
