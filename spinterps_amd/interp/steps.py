"""GPU step driver: interpolate many time steps in one call.

The reference interpolates one (time group x cell group x method) at a time
(interp/steps.py:478-835, per-step Python loops at :293-313 and :315-434).
Here a call takes all steps at once, each with its own NaN mask and
variogram string, and returns [T x C] estimates (and OK variances) on the
GPU.  Torch tensors only carry device memory and the stream.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _lib
from .._lib import (METHODS, NEB_SEL, PRECISIONS, SpkF32Out, SpkNeb, SpkOpts,
                    SpkProblem, check, context, lib, vg_array)


def _dev(device):
    if not torch.cuda.is_available():
        raise RuntimeError('a ROCm GPU is required (torch.cuda unavailable)')
    return torch.device('cuda', device)


def _t64(a, dev):
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64),
                           device=dev)


class DeviceProblem:
    """Station and cell coordinates (and drifts) resident in HBM.

    Coordinates are metres as float64.  ``origin`` (default: the integer
    floor of the smallest station/cell coordinate) is subtracted on the
    device before any arithmetic, which keeps FP32 work accurate for
    UTM-sized coordinates and leaves FP64 differences bit-identical.
    """

    def __init__(self, stn_x, stn_y, cell_x, cell_y, stn_drift=None,
                 cell_drift=None, device=0, origin=None):
        self.device = _dev(device)
        d = self.device
        self.stn_x = _t64(stn_x, d)
        self.stn_y = _t64(stn_y, d)
        self.cell_x = _t64(cell_x, d)
        self.cell_y = _t64(cell_y, d)
        assert self.stn_x.shape == self.stn_y.shape and self.stn_x.ndim == 1
        assert self.cell_x.shape == self.cell_y.shape and self.cell_x.ndim == 1
        if origin is None:
            origin = (math.floor(float(min(self.stn_x.min(),
                                           self.cell_x.min()))),
                      math.floor(float(min(self.stn_y.min(),
                                           self.cell_y.min()))))
        self.origin = (float(origin[0]), float(origin[1]))
        self.n_drifts = 0
        self.stn_drift = self.cell_drift = None
        if stn_drift is not None:
            sd = _t64(stn_drift, d)
            cd = _t64(cell_drift, d)
            if sd.ndim == 1:
                sd = sd[:, None].contiguous()
            if cd.ndim == 1:
                cd = cd[None, :].contiguous()
            assert sd.shape[0] == self.n_stns and cd.shape[1] == self.n_cells
            assert sd.shape[1] == cd.shape[0]
            self.stn_drift, self.cell_drift = sd, cd
            self.n_drifts = sd.shape[1]

    @property
    def n_stns(self):
        return self.stn_x.numel()

    @property
    def n_cells(self):
        return self.cell_x.numel()

    def c_struct(self):
        p = SpkProblem()
        p.n_stns = self.n_stns
        p.n_cells = self.n_cells
        p.n_drifts = self.n_drifts
        p.stn_x = self.stn_x.data_ptr()
        p.stn_y = self.stn_y.data_ptr()
        p.cell_x = self.cell_x.data_ptr()
        p.cell_y = self.cell_y.data_ptr()
        p.stn_drift = self.stn_drift.data_ptr() if self.n_drifts else None
        p.cell_drift = self.cell_drift.data_ptr() if self.n_drifts else None
        p.origin_x, p.origin_y = self.origin
        return p

    @classmethod
    def from_workload(cls, w, device=0):
        return cls(w.stn_x, w.stn_y, w.cell_x, w.cell_y, w.stn_drift,
                   w.cell_drift, device=device)


def vg_device(vgs, device):
    """Parse variogram strings into a device array of spk_vg records."""
    arr = vg_array(list(vgs))
    raw = torch.from_numpy(arr.view(np.uint8).reshape(-1).copy())
    return raw.to(device)


def _ctx_for(device):
    ctx = context(device.index or 0)
    ctx.set_stream(torch.cuda.current_stream(device).cuda_stream)
    return ctx


def interp_steps(prob: DeviceProblem, vals, vgs=None, method='OK',
                 precision='fp64', want_var=True, idw_exp=2.0,
                 min_vg_val=0.0, min_var_thr=-np.inf, est=None, var=None,
                 vg_dev=None, kappa=None, fp64var_kappa=0.0,
                 int8var_kappa=0.0, out32=None, min_var_cut=None,
                 max_var_cut=None):
    """Interpolate T steps.  vals [T x N] (NaN = missing station).

    method: 'OK', 'SK' (the reference's SK: lambda^T z, nearest station
    where sum(lambda) is not ~1), 'SKM' (simple kriging with the step mean
    as known mean, m + lambda^T (z - m)), 'EDK', 'IDW' or 'NNB'.

    precision: 'fp64' (each step on its own, FP64 end to end), 'mixed'
    (batched; the well-conditioned steps' variance contraction on
    split-BF16 operands), 'mixed_f32' (the same with FP32 operands) or
    'fp64b' (batched, FP64 contraction).

    Returns (est [T x C] float64, var [T x C] float64 or None,
    info [T] int32) as device tensors.  info holds SPK_ST_* flags per step:
    1 no station (outputs left NaN), 2 mean fill, 4 failed system
    (nearest-station fill), 8 single station, 16 ill-conditioned
    (condition estimate > 1e12), 32 variance promoted to FP64 (mixed).
    ``kappa``: optional float64 device tensor [T] that receives each
    system's condition estimate.  ``fp64var_kappa`` (mixed precision):
    condition estimate above which a step's variance contraction runs in
    FP64, for every variogram model (0 = the library's policy: 1e5, and
    1e3 for steps with a Gaussian component).  ``int8var_kappa``: up to
    this condition estimate such steps use exact int8 plane products
    (flag 64), above it FP64 (0 = no limit, negative = never).
    ``out32``: optional (est32, var32) float32 CPU tensors [T x C] (var32
    may be None; page-locked for overlap) that receive the fields as the
    reference keeps them (estimates clipped to [min_var_cut,
    max_var_cut], interp/steps.py:466-476), copied batch by batch while
    later batches compute; complete once the device is synchronised.
    """
    dev = prob.device
    v = _t64(vals, dev)
    assert v.ndim == 2 and v.shape[1] == prob.n_stns
    T = v.shape[0]
    # 'SKM': simple kriging with the step mean as known mean,
    # m + lambda^T (z - m), reported apart from the reference's SK
    sk_mean = method == 'SKM'
    if sk_mean:
        method = 'SK'
    m = METHODS[method]
    krig = method in ('OK', 'SK', 'EDK')
    if krig:
        if vg_dev is None:
            assert vgs is not None and len(vgs) == T
            if any(str(s) == 'nan' for s in vgs):
                raise ValueError('NaN VG!')
            vg_dev = vg_device(vgs, dev)
    if method == 'EDK' and prob.n_drifts < 1:
        raise AssertionError('EDK needs drifts')
    C = prob.n_cells
    if est is None:
        est = torch.full((T, C), float('nan'), dtype=torch.float64,
                         device=dev)
    wv = bool(want_var) and method == 'OK'
    if wv and var is None:
        var = torch.full((T, C), float('nan'), dtype=torch.float64,
                         device=dev)
    info = torch.zeros(T, dtype=torch.int32, device=dev)
    opts = SpkOpts()
    opts.method = m
    opts.precision = PRECISIONS[precision]
    opts.want_var = 1 if wv else 0
    opts.idw_exp = float(idw_exp)
    opts.min_vg_val = float(min_vg_val)
    opts.min_var_thr = float(min_var_thr)
    opts.fp64var_kappa = float(fp64var_kappa)
    opts.int8var_kappa = float(int8var_kappa)
    opts.sk_mean = 1 if sk_mean else 0
    ctx = _ctx_for(dev)
    pst = prob.c_struct()
    args = (ctx.h, ctypes.byref(pst), ctypes.byref(opts), T, v.data_ptr(),
            vg_dev.data_ptr() if krig else None, est.data_ptr(),
            var.data_ptr() if wv else None, info.data_ptr(),
            kappa.data_ptr() if kappa is not None else None)
    if out32 is None:
        check(lib().spk_interp_steps_dev(*args))
    else:
        e32, v32 = out32
        for t in (e32, v32):
            if t is not None:
                assert (t.device.type == 'cpu' and t.dtype == torch.float32
                        and t.is_contiguous() and tuple(t.shape) == (T, C))
        o = SpkF32Out()
        o.est32 = e32.data_ptr()
        o.var32 = v32.data_ptr() if (v32 is not None and wv) else None
        o.has_min = 0 if min_var_cut is None else 1
        o.min_cut = 0.0 if min_var_cut is None else float(min_var_cut)
        o.has_max = 0 if max_var_cut is None else 1
        o.max_cut = 0.0 if max_var_cut is None else float(max_var_cut)
        check(lib().spk_interp_steps_f32_dev(*args, ctypes.byref(o)))
    return est, (var if wv else None), info


def interp_neb_steps(prob: DeviceProblem, vals, vgs=None, method='OK',
                     neb_sel_mthd='nrst', n_nebs=50, n_pies=None,
                     want_var=True, idw_exp=2.0, min_vg_val=0.0,
                     min_var_thr=-np.inf, est=None, var=None, vg_dev=None,
                     kappa=None):
    """interp_steps with the reference's neighbourhoods: every cell uses
    only the n_nebs nearest active stations ('nrst') or n_nebs taken
    sector by sector over n_pies sectors ('pie'); stations that no cell
    selects with every station present count as missing
    (interp/grps.py:141-288, interp/steps.py:592-608, 732-831).  FP64.

    Returns (est, var or None, info) like interp_steps; info holds the OR
    over each step's neighbourhood groups of the SPK_ST_* flags.
    """
    if neb_sel_mthd not in NEB_SEL:
        raise NotImplementedError(neb_sel_mthd)
    dev = prob.device
    v = _t64(vals, dev)
    assert v.ndim == 2 and v.shape[1] == prob.n_stns
    T = v.shape[0]
    sk_mean = method == 'SKM'
    if sk_mean:
        method = 'SK'
    m = METHODS[method]
    krig = method in ('OK', 'SK', 'EDK')
    if krig and vg_dev is None:
        assert vgs is not None and len(vgs) == T
        if any(str(s) == 'nan' for s in vgs):
            raise ValueError('NaN VG!')
        vg_dev = vg_device(vgs, dev)
    if method == 'EDK' and prob.n_drifts < 1:
        raise AssertionError('EDK needs drifts')
    C = prob.n_cells
    if est is None:
        est = torch.full((T, C), float('nan'), dtype=torch.float64,
                         device=dev)
    wv = bool(want_var) and method == 'OK'
    if wv and var is None:
        var = torch.full((T, C), float('nan'), dtype=torch.float64,
                         device=dev)
    info = torch.zeros(T, dtype=torch.int32, device=dev)
    opts = SpkOpts()
    opts.method = m
    opts.precision = PRECISIONS['fp64']
    opts.want_var = 1 if wv else 0
    opts.idw_exp = float(idw_exp)
    opts.min_vg_val = float(min_vg_val)
    opts.min_var_thr = float(min_var_thr)
    opts.sk_mean = 1 if sk_mean else 0
    neb = SpkNeb()
    neb.sel = NEB_SEL[neb_sel_mthd]
    neb.n_nebs = int(n_nebs)
    neb.n_pies = int(n_pies) if n_pies is not None else 0
    ctx = _ctx_for(dev)
    pst = prob.c_struct()
    check(lib().spk_interp_neb_steps_dev(
        ctx.h, ctypes.byref(pst), ctypes.byref(opts), ctypes.byref(neb), T,
        v.data_ptr(), vg_dev.data_ptr() if krig else None, est.data_ptr(),
        var.data_ptr() if wv else None, info.data_ptr(),
        kappa.data_ptr() if kappa is not None else None))
    return est, (var if wv else None), info


def clip_cast(f64, min_var_cut=None, max_var_cut=None, out=None):
    """interp/steps.py:466-476 clip + float32 cast, on the device (into
    ``out``, a float32 device tensor of f64's shape, when given)."""
    if out is None:
        out = torch.empty(f64.shape, dtype=torch.float32, device=f64.device)
    assert out.dtype == torch.float32 and out.shape == f64.shape
    assert out.is_contiguous() and f64.is_contiguous()
    ctx = _ctx_for(f64.device)
    check(lib().spk_clip_cast_dev(
        ctx.h, f64.data_ptr(), out.data_ptr(), f64.numel(),
        0 if min_var_cut is None else 1,
        0.0 if min_var_cut is None else float(min_var_cut),
        0 if max_var_cut is None else 1,
        0.0 if max_var_cut is None else float(max_var_cut)))
    return out


def prof_enable(device=0, on=True):
    ctx = context(device)
    check(lib().spk_prof_enable(ctx.h, 1 if on else 0))
    check(lib().spk_prof_reset(ctx.h))


def prof_get(name, device=0):
    ctx = context(device)
    ms = ctypes.c_double()
    cnt = ctypes.c_int64()
    check(lib().spk_prof_get(ctx.h, name.encode(), ctypes.byref(ms),
                             ctypes.byref(cnt)))
    return ms.value, cnt.value


def prof_work(name, device=0):
    """Executed work counted while profiling is on ('gemm_tri_f32_macs',
    'gemm_tri_bf16_macs', 'gemm_tri_i8_ops', 'neb_chunks')."""
    ctx = context(device)
    v = ctypes.c_double()
    check(lib().spk_prof_work(ctx.h, name.encode(), ctypes.byref(v)))
    return v.value


def interpolate(prob: DeviceProblem, vals, vgs, methods, *,
                min_var_thr=-np.inf, min_var_cut=None, max_var_cut=None,
                min_vg_val=0.0, est_vars_flag=True, precision='fp64',
                neb_sel_mthd='all', n_nebs=None, n_pies=None):
    """Fields of all methods for all steps, as the reference's chunk
    routine produces them (interp/steps.py:478-835).  neb_sel_mthd 'all'
    (the reference's default, interp/data.py:62), 'nrst' or 'pie' with
    n_nebs (and n_pies) as the reference takes them; the neighbourhood
    modes compute in FP64 whatever ``precision`` says.

    methods: list of (type, label, idw_exp) like the reference's
    interp_args.  Returns (fields, fields64, prblm): fields[label] is the
    float32 [T x C] tensor after clipping (steps.py:805, 815-831),
    fields64 the float64 values before the cast, prblm the step indices
    without stations or with a failed system (steps.py:666-688, 363-371).
    """
    fields, f64 = {}, {}
    prblm = set()
    v = _t64(vals, prob.device)
    vg_dev = None
    if any(m[0] in ('OK', 'SK', 'SKM', 'EDK') for m in methods):
        if any(str(s) == 'nan' for s in vgs):
            raise AssertionError('NaN VGs not allowed! Use Nugget or any '
                                 'other appropriate one!')
        vg_dev = vg_device(vgs, prob.device)
    for typ, lab, idw_exp in methods:
        if lab == 'EST_VARS_OK':
            continue
        kw = dict(want_var=est_vars_flag and typ == 'OK',
                  idw_exp=idw_exp if typ == 'IDW' else 2.0,
                  min_vg_val=min_vg_val, min_var_thr=min_var_thr,
                  vg_dev=vg_dev)
        if neb_sel_mthd == 'all':
            e, var, info = interp_steps(prob, v, vgs, typ,
                                        precision=precision, **kw)
        else:
            e, var, info = interp_neb_steps(
                prob, v, vgs, typ, neb_sel_mthd=neb_sel_mthd, n_nebs=n_nebs,
                n_pies=n_pies, **kw)
        f64[lab] = e
        fields[lab] = clip_cast(e, min_var_cut, max_var_cut)
        if var is not None:
            f64['EST_VARS_OK'] = var
            fields['EST_VARS_OK'] = clip_cast(var)
        inf = info.cpu().numpy()
        prblm.update(np.where((inf & (_lib.ST_NOSTNS | _lib.ST_FAILED))
                              != 0)[0].tolist())
    return fields, f64, sorted(prblm)
