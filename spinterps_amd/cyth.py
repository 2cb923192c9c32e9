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
           'fill_theo_vg_vals', 'copy_2d_arr_at_idxs', 'get_nd_dists']


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

