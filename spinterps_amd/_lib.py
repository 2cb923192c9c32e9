"""ctypes binding of libspk.so (the HIP kernels behind include/spk.h).

The library must be built (``python -m spinterps_amd.build``); there is no
CPU fallback: every call below runs on the GPU or raises.
"""
from __future__ import annotations

import ctypes
import threading
from ctypes import (POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64,
                    c_uint64, c_void_p)
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().with_name('libspk.so')

SPK_OK, SPK_EBADMODEL, SPK_ESHAPE, SPK_EHIP, SPK_ENOTSPD, SPK_EARG, \
    SPK_ENOMEM, SPK_EINDEX = range(8)

VG_NAMES = ['Rng', 'Nug', 'Sph', 'Exp', 'Lin', 'Gau', 'Pow', 'Hol']
VG_MAX = 8

M_OK, M_SK, M_EDK, M_IDW, M_NNB = range(5)
METHODS = {'OK': M_OK, 'SK': M_SK, 'EDK': M_EDK, 'IDW': M_IDW, 'NNB': M_NNB}
P_FP64, P_MIXED, P_FP64B, P_MIXED_F32 = range(4)
TUNE_OFF, TUNE_AUTO = 0, 1
PRECISIONS = {'fp64': P_FP64, 'mixed': P_MIXED, 'fp64b': P_FP64B,
              'mixed_f32': P_MIXED_F32}

ST_NOSTNS, ST_MEANFILL, ST_FAILED, ST_ONESTN = 1, 2, 4, 8
ST_ILLCOND, ST_FP64VAR, ST_INT8VAR, ST_PINV = 16, 32, 64, 128


class SpkVg(ctypes.Structure):
    _fields_ = [('n', c_int32), ('type', c_int32 * VG_MAX),
                ('sill', c_double * VG_MAX), ('range', c_double * VG_MAX),
                ('sill_sum', c_double), ('range_max_raw', c_double)]


class SpkProblem(ctypes.Structure):
    _fields_ = [('n_stns', c_int64), ('n_cells', c_int64),
                ('n_drifts', c_int32), ('pad_', c_int32),
                ('stn_x', c_void_p), ('stn_y', c_void_p),
                ('cell_x', c_void_p), ('cell_y', c_void_p),
                ('stn_drift', c_void_p), ('cell_drift', c_void_p),
                ('origin_x', c_double), ('origin_y', c_double)]


class SpkOpts(ctypes.Structure):
    _fields_ = [('method', c_int32), ('precision', c_int32),
                ('want_var', c_int32), ('sk_mean', c_int32),
                ('idw_exp', c_double), ('min_vg_val', c_double),
                ('min_var_thr', c_double), ('fp64var_kappa', c_double),
                ('int8var_kappa', c_double)]


class SpkF32Out(ctypes.Structure):
    _fields_ = [('est32', c_void_p), ('var32', c_void_p),
                ('has_min', c_int32), ('has_max', c_int32),
                ('min_cut', c_double), ('max_cut', c_double)]


class SpkNeb(ctypes.Structure):
    _fields_ = [('sel', c_int32), ('n_nebs', c_int32), ('n_pies', c_int32),
                ('pad_', c_int32)]


NEB_SEL = {'nrst': 1, 'pie': 2}


VG_DTYPE = np.dtype([('n', '<i4'), ('type', '<i4', (VG_MAX,)),
                     ('sill', '<f8', (VG_MAX,)), ('range', '<f8', (VG_MAX,)),
                     ('sill_sum', '<f8'), ('range_max_raw', '<f8')],
                    align=True)
assert VG_DTYPE.itemsize == ctypes.sizeof(SpkVg)

_PD = POINTER(c_double)
_PI64 = POINTER(c_int64)
_PU64 = POINTER(c_uint64)

# name -> (restype, argtypes); every symbol of include/spk.h
SIGNATURES = {
    'spk_last_error': (c_char_p, []),
    'spk_version': (c_int, []),
    'spk_ctx_create': (c_int, [c_int, POINTER(c_void_p)]),
    'spk_ctx_destroy': (c_int, [c_void_p]),
    'spk_ctx_set_stream': (c_int, [c_void_p, c_void_p]),
    'spk_ctx_sync': (c_int, [c_void_p]),
    'spk_ctx_trim': (c_int, [c_void_p]),
    'spk_ctx_set_graphs': (c_int, [c_void_p, c_int]),
    'spk_ctx_set_tuning': (c_int, [c_void_p, c_int]),
    'spk_ctx_force_variant': (c_int, [c_void_p, c_char_p, c_int]),
    'spk_ctx_tuning_report': (c_int, [c_void_p, c_char_p, c_int64, _PI64]),
    'spk_vg_parse': (c_int, [c_char_p, POINTER(SpkVg)]),
    'spk_vg_full_nuggetness': (c_int, [c_char_p, c_double, POINTER(c_int)]),
    'spk_fill_dists_2d_mat': (c_int, [c_void_p, _PD, _PD, c_int64, _PD, _PD,
                                      c_int64, _PD]),
    'spk_fill_vg_var_arr': (c_int, [c_void_p, _PD, c_int64, c_int64, _PD,
                                    c_int, c_int, c_char_p, c_double]),
    'spk_copy_2d_arr_at_idxs': (c_int, [c_void_p, _PD, c_int64, c_int64,
                                        _PI64, c_int64, _PI64, c_int64, _PD,
                                        c_int64, c_int64]),
    'spk_fill_dists_one_pt': (c_int, [c_void_p, c_double, c_double, _PD, _PD,
                                      c_int64, _PD]),
    'spk_fill_wts_and_sum': (c_int, [c_void_p, _PD, _PD, c_int64, c_double,
                                     _PD]),
    'spk_get_mults_sum': (c_int, [c_void_p, _PD, _PD, c_int64, _PD]),
    'spk_fill_theo_vg_vals': (c_int, [c_void_p, c_char_p, _PD, c_int64,
                                      c_double, c_double, _PD]),
    'spk_get_nd_dists': (c_int, [c_void_p, _PD, c_int64, c_int64, _PD]),
    'spk_sel_equidist_refs': (c_int, [c_void_p, c_double, c_double, _PD, _PD,
                                      c_int64, c_uint64, c_double, c_int64,
                                      _PD, _PD, _PI64, _PU64, _PU64]),
    'spk_neb_idxs_nrst': (c_int, [c_void_p, _PD, _PD, c_int64, _PD, _PD,
                                  c_int64, c_int64, _PI64]),
    'spk_neb_idxs_pie': (c_int, [c_void_p, _PD, _PD, c_int64, _PD, _PD,
                                 c_int64, c_int64, c_int64, _PI64]),
    'spk_interp_steps_dev': (c_int, [c_void_p, POINTER(SpkProblem),
                                     POINTER(SpkOpts), c_int64, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    'spk_interp_steps_f32_dev': (c_int, [c_void_p, POINTER(SpkProblem),
                                         POINTER(SpkOpts), c_int64, c_void_p,
                                         c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p,
                                         POINTER(SpkF32Out)]),
    'spk_interp_neb_steps_dev': (c_int, [c_void_p, POINTER(SpkProblem),
                                         POINTER(SpkOpts), POINTER(SpkNeb),
                                         c_int64, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p,
                                         c_void_p]),
    'spk_idw_steps_noalloc_dev': (c_int, [c_void_p, POINTER(SpkProblem),
                                          POINTER(SpkOpts), c_int64,
                                          c_void_p, c_void_p, c_void_p,
                                          c_void_p]),
    'spk_clip_cast_dev': (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int,
                                  c_double, c_int, c_double]),
    'spk_last_timings': (c_int, [c_void_p, _PD, c_int, POINTER(c_int)]),
    'spk_prof_enable': (c_int, [c_void_p, c_int]),
    'spk_prof_reset': (c_int, [c_void_p]),
    'spk_prof_get': (c_int, [c_void_p, c_char_p, _PD, _PI64]),
    'spk_prof_work': (c_int, [c_void_p, c_char_p, _PD]),
    'spk_vgfit_pairs_dev': (c_int, [c_void_p, c_void_p, c_void_p, c_int64,
                                    c_void_p, c_void_p, c_void_p]),
    'spk_vgfit_evg_dev': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_int64, c_int64, c_void_p, c_int64,
                                  c_double, c_int, c_int, c_int, c_void_p,
                                  c_void_p, c_void_p, c_void_p]),
    'spk_vgfit_de_dev': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_int, c_void_p, c_int64,
                                 POINTER(c_int32), c_int, c_int, c_int,
                                 c_double, c_double, c_uint64, c_void_p,
                                 c_void_p, c_void_p]),
    'spk_vgfit_de_seeded_dev': (c_int, [c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_int, c_void_p, c_int64,
                                        POINTER(c_int32), c_int, c_int,
                                        c_int, c_double, c_double, c_uint64,
                                        c_void_p, c_double, c_void_p,
                                        c_void_p, c_void_p]),
    'spk_simulate_fields_dev': (c_int, [c_void_p, c_void_p, c_int64,
                                        c_void_p, c_void_p, c_void_p,
                                        c_uint64, c_void_p, c_void_p]),
    'spk_vg_features_dev': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_int64, c_int64, c_void_p, c_int64,
                                    c_double, c_int, c_void_p, c_void_p]),
    'spk_krige_class': (c_int, [c_void_p, c_int, c_char_p, c_void_p,
                                c_void_p, c_void_p, c_int64, c_void_p, c_int,
                                c_void_p, c_void_p, c_void_p, c_int64,
                                c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p]),
    'spk_mlp_n_params': (c_int64, []),
    'spk_mlp_init_dev': (c_int, [c_void_p, c_uint64, c_void_p]),
    'spk_mlp_forward_dev': (c_int, [c_void_p, c_void_p, c_void_p, c_int,
                                    c_void_p]),
    'spk_mlp_loss_grad_dev': (c_int, [c_void_p, c_void_p, c_void_p, c_int,
                                      c_void_p, c_void_p, c_float, c_void_p,
                                      c_void_p]),
    'spk_mlp_adam_dev': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_int64, c_float, c_float, c_float,
                                 c_float]),
}

_lib = None
_lock = threading.Lock()


def lib():
    """Load libspk.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f'{LIB_PATH} is missing: build it with '
                    '`python -m spinterps_amd.build` (HIP is required)')
            h = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                f = getattr(h, name)
                f.restype = res
                f.argtypes = args
            _lib = h
    return _lib


class SpkError(RuntimeError):
    pass


def last_error():
    return lib().spk_last_error().decode(errors='replace')


def check(rc):
    """Map a status code to the exception the reference would raise."""
    if rc == SPK_OK:
        return
    msg = last_error()
    if rc in (SPK_ESHAPE, SPK_EARG):
        raise AssertionError(msg)
    if rc == SPK_EBADMODEL:
        raise ValueError(msg)
    if rc == SPK_EINDEX:
        raise IndexError(msg)
    raise SpkError(f'spk error {rc}: {msg}')


class Context:
    """A device context (stream + cached workspaces) of libspk."""

    def __init__(self, device=0):
        self.device = device
        h = c_void_p()
        check(lib().spk_ctx_create(device, ctypes.byref(h)))
        self.h = h

    def set_stream(self, stream_ptr):
        check(lib().spk_ctx_set_stream(self.h, c_void_p(stream_ptr)))

    def sync(self):
        check(lib().spk_ctx_sync(self.h))

    def trim(self):
        check(lib().spk_ctx_trim(self.h))

    def set_graphs(self, on):
        """Run the batched factorisation's launch sequences as captured
        hipGraphs (default) or launch by launch."""
        check(lib().spk_ctx_set_graphs(self.h, 1 if on else 0))

    def set_tuning(self, auto):
        """Tile autotuning on (time each kernel variant at the first launch
        of a size class and keep the fastest) or off (library defaults)."""
        check(lib().spk_ctx_set_tuning(self.h, TUNE_AUTO if auto else TUNE_OFF))

    def force_variant(self, family, v):
        """Launch variant v of a kernel family ('gemm_plain_f64' 0-2,
        'gemm_tri_bf16' 0-1); v = -1 clears."""
        check(lib().spk_ctx_force_variant(self.h, family.encode(), int(v)))

    def tuning_report(self):
        """[(family, size class, variant, [ms per variant])] picked so far."""
        n = c_int64(0)
        check(lib().spk_ctx_tuning_report(self.h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        check(lib().spk_ctx_tuning_report(self.h, buf, n.value + 1, None))
        out = []
        for line in buf.value.decode().splitlines():
            f = line.split()
            out.append((f[0], int(f[1]), int(f[2]), [float(x) for x in f[3:]]))
        return out

    def close(self):
        if self.h:
            lib().spk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctxs = {}


def context(device=0):
    """Process-wide default context per device."""
    with _lock:
        if device not in _ctxs:
            _ctxs[device] = Context(device)
        return _ctxs[device]


def parse_vg(vg_str):
    out = SpkVg()
    check(lib().spk_vg_parse(str(vg_str).encode(), ctypes.byref(out)))
    return out


def vg_array(vg_strs):
    """Parse strings into a structured array matching spk_vg (cached)."""
    cache = {}
    arr = np.zeros(len(vg_strs), dtype=VG_DTYPE)
    raw = arr.view(np.uint8).reshape(len(vg_strs), VG_DTYPE.itemsize)
    for i, s in enumerate(vg_strs):
        if s not in cache:
            cache[s] = bytes(parse_vg(s))
        raw[i] = np.frombuffer(cache[s], dtype=np.uint8)
    return arr


def dptr(a):
    """ctypes double* to a C-contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(_PD)


def i64ptr(a):
    assert a.dtype == np.int64 and a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(_PI64)


def u64ptr(a):
    assert a.dtype == np.uint64 and a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(_PU64)
