"""The NetCDF writer of the interpolated fields (interp/prepare.py:290-432,
interp/steps.py:879-969): CDF-2 files read back by scipy.io's independent
reader, CDF-5 files against a header assembled by hand from the format
specification, the reference's in-place rounding."""
import struct

import numpy as np
import pytest

from spinterps_amd.interp.ncfile import NcWriter, read_nc, reference_attrs


def _case():
    rng = np.random.default_rng(3)
    nx, ny, nt = 7, 5, 6
    x = 350000.0 + 1000.0 * np.arange(nx) + 500.0
    y = 5400000.0 - 1000.0 * np.arange(ny) - 500.0
    t = np.arange(nt) * 24 + 1000
    f1 = rng.gamma(0.8, 5.0, (nt, ny, nx)).astype(np.float32)
    f2 = rng.normal(size=(nt, ny * nx)).astype(np.float32)
    return x, y, t, f1, f2


def test_cdf2_read_by_scipy(tmp_path):
    from scipy.io import netcdf_file
    x, y, t, f1, f2 = _case()
    args = [('OK', 'ok', 'OK'), ('IDW', 'idw', 'IDW_000', 2.0)]
    va, ga = reference_attrs(args, 'precipitation', 'mm',
                             {'n_nebs': 50, 'neb_sel_mthd': 'nrst'},
                             source='out.nc')
    w = NcWriter(tmp_path / 'a.nc', x, y, t, ['OK', 'IDW_000'],
                 var_attrs=va, global_attrs=ga,
                 time_attrs={'units': 'hours since 1990-01-01',
                             'calendar': 'gregorian'}, version=2)
    w.write('OK', 0, f1[:4])
    w.write('OK', 4, f1[4:])
    w.write('IDW_000', 0, f2)
    with netcdf_file(tmp_path / 'a.nc', 'r', mmap=False) as nc:
        assert nc.dimensions == {'dimx': 7, 'dimy': 5, 'dimt': 6}
        np.testing.assert_array_equal(nc.variables['X'][:], x)
        np.testing.assert_array_equal(nc.variables['Y'][:], y)
        np.testing.assert_array_equal(nc.variables['time'][:], t)
        assert nc.variables['time'].units == b'hours since 1990-01-01'
        ok = nc.variables['OK']
        assert ok.dimensions == ('dimt', 'dimy', 'dimx')
        assert ok.data.dtype == np.dtype('>f4')
        np.testing.assert_array_equal(ok[:], f1)
        np.testing.assert_array_equal(nc.variables['IDW_000'][:],
                                      f2.reshape(6, 5, 7))
        assert ok.standard_name == b'precipitation (OK)'
        assert nc.variables['IDW_000'].standard_name == \
            b'precipitation (IDW_exp_2.0)'
        assert nc.sett_n_nebs == b'50' and nc.Source == b'out.nc'
    d, g, v = read_nc(tmp_path / 'a.nc')
    np.testing.assert_array_equal(v['OK'][2], f1)


def test_rounding_is_the_references(tmp_path):
    x, y, t, f1, _ = _case()
    w = NcWriter(tmp_path / 'r.nc', x, y, t, ['OK'])
    w.write('OK', 0, f1.copy(), nmrl_prcn=2)
    _, _, v = read_nc(tmp_path / 'r.nc')
    np.testing.assert_array_equal(v['OK'][2], np.round(f1, 2))


def test_cdf5_header_by_hand(tmp_path):
    """A 2 x 1 grid, one step, one label, one global string attribute:
    every byte of the header assembled from the CDF-5 layout (32-bit tags
    and types, 64-bit counts, lengths, sizes and offsets, names padded to
    4 bytes)."""
    w = NcWriter(tmp_path / 'h.nc', [1.0, 2.0], [5.0], [7], ['OK'],
                 global_attrs={'Source': 'ab'}, version=5)
    w.write('OK', 0, np.array([[[1.5, -2.0]]], dtype=np.float32))
    q = lambda v: struct.pack('>q', v)
    i = lambda v: struct.pack('>i', v)

    def name(s):
        b = s.encode()
        return q(len(b)) + b + b'\0' * ((4 - len(b) % 4) % 4)

    hdr = b'CDF\x05' + q(0)
    hdr += i(0x0A) + q(3) + name('dimx') + q(2) + name('dimy') + q(1) + \
        name('dimt') + q(1)
    hdr += i(0x0C) + q(1) + name('Source') + i(2) + q(2) + b'ab\0\0'
    absent = i(0) + q(0)
    vars_ = [('X', [0], 6, 16), ('Y', [1], 6, 8), ('time', [2], 10, 8),
             ('OK', [2, 1, 0], 5, 8)]
    # the header's length fixes the first offset; build it twice
    def var_list(begins):
        out = i(0x0B) + q(len(vars_))
        for (nm, dids, t, size), b in zip(vars_, begins):
            out += name(nm) + q(len(dids)) + b''.join(q(d) for d in dids)
            out += absent + i(t) + q(size) + q(b)
        return out
    n0 = len(hdr + var_list([0] * 4))
    begins = [n0, n0 + 16, n0 + 24, n0 + 32]
    exp = hdr + var_list(begins)
    raw = (tmp_path / 'h.nc').read_bytes()
    assert raw[:len(exp)] == exp
    assert len(raw) == n0 + 40
    assert raw[n0:n0 + 16] == struct.pack('>2d', 1.0, 2.0)
    assert raw[n0 + 24:n0 + 32] == q(7)
    assert raw[n0 + 32:n0 + 40] == struct.pack('>2f', 1.5, -2.0)


def test_cdf2_refuses_a_large_variable_before_the_last(tmp_path):
    with pytest.raises(ValueError, match='version=5'):
        NcWriter(tmp_path / 'big.nc', np.arange(40000.0),
                 np.arange(40000.0), [0], ['A', 'B'], version=2)
