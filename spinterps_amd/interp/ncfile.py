"""NetCDF output of the interpolated fields, without netCDF4 or HDF5.

The reference writes one netCDF4 file per run (interp/prepare.py:290-432):
dimensions ``dimx``, ``dimy``, ``dimt``; double coordinate variables for
the cell centres along x and y; an int64 time variable (with ``units`` and
``calendar`` for a date index); one float32 variable per interpolation
label on (dimt, dimy, dimx) with ``units`` and ``standard_name``; the run's
settings as ``sett_*`` global attributes and ``Source``.  Chunks of steps
are written as they finish, each field rounded in place to ``nmrl_prcn``
decimals first (interp/steps.py:879-969).

Neither netCDF4 nor h5py is available here, and the HDF5 container behind
netCDF-4 is far outside this path, so this module writes the same names,
types, attributes and layout in the netCDF classic format, which every
netCDF library reads: CDF-5 (64-bit data: int64 time, variables beyond
4 GiB) by default, or CDF-2 (64-bit offsets) where the data allow it.  The
netCDF-4-only zlib compression and (1, ny, nx) chunking are therefore not
applied; fields are stored uncompressed, each step a contiguous (ny, nx)
slab.  ``read_nc`` reads both formats back (the tests also read CDF-2 with
scipy.io's independent reader).
"""
from __future__ import annotations

import os
import struct
from pathlib import Path

import numpy as np

# classic-format tags and types (the netCDF file format specification)
_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 0x0A, 0x0B, 0x0C
_TYPES = {  # numpy dtype -> (nc_type, size)
    np.dtype('i1'): (1, 1), np.dtype('S1'): (2, 1), np.dtype('>i2'): (3, 2),
    np.dtype('>i4'): (4, 4), np.dtype('>f4'): (5, 4), np.dtype('>f8'): (6, 8),
    np.dtype('u1'): (7, 1), np.dtype('>u2'): (8, 2), np.dtype('>u4'): (9, 4),
    np.dtype('>i8'): (10, 8), np.dtype('>u8'): (11, 8)}
_NC_TO_DTYPE = {v[0]: k for k, v in _TYPES.items()}


def _be(dt):
    dt = np.dtype(dt)
    if dt.kind == 'S':
        return np.dtype('S1')
    return dt.newbyteorder('>') if dt.itemsize > 1 else dt


def _pad4(n):
    return (4 - n % 4) % 4


def round_like_reference(fld, nmrl_prcn):
    """np.round(fld, nmrl_prcn, fld) as interp/steps.py:917-918 does it,
    in the field's own dtype (float32 for the output fields)."""
    if nmrl_prcn is not None and np.issubdtype(fld.dtype, np.floating):
        np.round(fld, nmrl_prcn, fld)
    return fld


class NcWriter:
    """One output file of ``n_steps`` x ``ny`` x ``nx`` fields.

    labels: the interpolation labels (variable names), in the reference's
    order; var_attrs: {label: {attr: value}} (``units``, ``standard_name``);
    time_vals and time_attrs: the time variable (int64) and its attributes;
    global_attrs: {name: value} (the reference writes ``sett_*`` strings
    and ``Source``).  version 5 (CDF-5) or 2 (CDF-2: the time variable is
    stored as int32, or as double when it does not fit).
    """

    def __init__(self, path, x_crds, y_crds, time_vals, labels, *,
                 var_attrs=None, time_attrs=None, global_attrs=None,
                 x_lab='X', y_lab='Y', t_lab='time', dtype=np.float32,
                 version=5):
        if version not in (2, 5):
            raise ValueError('version must be 2 (CDF-2) or 5 (CDF-5)')
        self.path = Path(path)
        self.version = version
        self.nx, self.ny = int(len(x_crds)), int(len(y_crds))
        self.nt = int(len(time_vals))
        self.labels = list(labels)
        self.dtype = np.dtype(dtype)
        tv = np.asarray(time_vals)
        if version == 5:
            tv = tv.astype('>i8')
        elif np.all(np.abs(tv) < 2 ** 31) and np.issubdtype(tv.dtype,
                                                             np.integer):
            tv = tv.astype('>i4')
        else:
            tv = tv.astype('>f8')
        # (name, dims, dtype, attrs, data or None)
        self._vars = [
            (x_lab, ('dimx',), np.dtype('>f8'), {},
             np.asarray(x_crds, dtype='>f8')),
            (y_lab, ('dimy',), np.dtype('>f8'), {},
             np.asarray(y_crds, dtype='>f8')),
            (t_lab, ('dimt',), tv.dtype, dict(time_attrs or {}), tv)]
        for lab in self.labels:
            self._vars.append((lab, ('dimt', 'dimy', 'dimx'), _be(dtype),
                               dict((var_attrs or {}).get(lab, {})), None))
        self._dims = [('dimx', self.nx), ('dimy', self.ny), ('dimt', self.nt)]
        self._gatts = dict(global_attrs or {})
        self._begin = {}
        self._write_header_and_coords()

    # ---- header -------------------------------------------------------------
    def _nn(self, v):  # NON_NEG
        return struct.pack('>q' if self.version == 5 else '>i', v)

    def _name(self, s):
        b = s.encode('utf-8')
        return self._nn(len(b)) + b + b'\0' * _pad4(len(b))

    def _attr_list(self, atts):
        if not atts:
            return struct.pack('>i', 0) + self._nn(0)
        out = struct.pack('>i', _NC_ATTRIBUTE) + self._nn(len(atts))
        for k, v in atts.items():
            out += self._name(k)
            if isinstance(v, (str, bytes)):
                b = v.encode('utf-8') if isinstance(v, str) else v
                out += struct.pack('>i', 2) + self._nn(len(b)) + b
                out += b'\0' * _pad4(len(b))
            else:
                a = np.atleast_1d(np.asarray(v))
                if a.dtype.kind == 'f':
                    a = a.astype('>f8')
                elif a.dtype.kind in 'iu':
                    a = a.astype('>i4' if self.version == 2 else '>i8')
                t, size = _TYPES[a.dtype]
                b = a.tobytes()
                out += struct.pack('>i', t) + self._nn(a.size) + b
                out += b'\0' * _pad4(len(b))
        return out

    def _header(self, begins):
        dimid = {d: i for i, (d, _) in enumerate(self._dims)}
        magic = b'CDF' + bytes([self.version])
        out = magic + self._nn(0)  # numrecs: no record dimension
        out += struct.pack('>i', _NC_DIMENSION) + self._nn(len(self._dims))
        for d, n in self._dims:
            out += self._name(d) + self._nn(n)
        out += self._attr_list(self._gatts)
        out += struct.pack('>i', _NC_VARIABLE) + self._nn(len(self._vars))
        for (name, dims, dt, atts, _), b in zip(self._vars, begins):
            out += self._name(name) + self._nn(len(dims))
            for d in dims:
                out += self._nn(dimid[d])
            out += self._attr_list(atts)
            t, size = _TYPES[dt]
            vsize = self._vsize(dims, size)
            if self.version == 2:
                vs = vsize if vsize < 2 ** 32 - 4 else 2 ** 32 - 1
                out += struct.pack('>i', t) + struct.pack('>I', vs)
                out += struct.pack('>q', b)
            else:
                out += struct.pack('>i', t) + struct.pack('>q', vsize)
                out += struct.pack('>q', b)
        return out

    def _vsize(self, dims, size):
        n = size
        for d in dims:
            n *= dict(self._dims)[d]
        return n + _pad4(n)

    def _write_header_and_coords(self):
        begins = [0] * len(self._vars)
        hdr = self._header(begins)
        off = len(hdr)
        for i, (name, dims, dt, atts, _) in enumerate(self._vars):
            begins[i] = off
            off += self._vsize(dims, _TYPES[dt][1])
        for i, (name, dims, dt, atts, _) in enumerate(self._vars):
            if self.version == 2 and begins[i] >= 2 ** 63:
                raise ValueError('file too large')
            if (self.version == 2 and i < len(self._vars) - 1 and
                    self._vsize(dims, _TYPES[dt][1]) >= 2 ** 32 - 4):
                raise ValueError(
                    f'variable {name} exceeds 4 GiB: CDF-2 allows that only '
                    'for the last variable; use version=5')
        hdr2 = self._header(begins)
        assert len(hdr2) == len(hdr)
        self._begin = {v[0]: b for v, b in zip(self._vars, begins)}
        self._end = off
        with open(self.path, 'wb') as f:
            f.write(hdr2)
            for (name, dims, dt, atts, data), b in zip(self._vars, begins):
                if data is not None:
                    f.seek(b)
                    f.write(np.ascontiguousarray(data, dtype=dt).tobytes())
            f.truncate(off)

    # ---- data ---------------------------------------------------------------
    def write(self, label, t0, fields, nmrl_prcn=None):
        """Fields of steps t0 .. t0 + len(fields) - 1 of ``label``
        ([k, ny, nx] or [k, ny * nx]), rounded in place to ``nmrl_prcn``
        decimals first as the reference does."""
        a = np.asarray(fields)
        if a.dtype != self.dtype:
            a = a.astype(self.dtype)
        a = a.reshape(-1, self.ny, self.nx)
        if not (0 <= t0 and t0 + a.shape[0] <= self.nt):
            raise IndexError('steps outside the file')
        round_like_reference(a, nmrl_prcn)
        slab = self.ny * self.nx * self.dtype.itemsize
        with open(self.path, 'r+b') as f:
            f.seek(self._begin[label] + t0 * slab)
            f.write(np.ascontiguousarray(a, dtype=_be(self.dtype)).tobytes())

    def close(self):
        pass


def read_nc(path):
    """Read a classic-format file (CDF-1, CDF-2 or CDF-5) written by
    NcWriter: returns (dims {name: len}, global attrs, {var: (dims, attrs,
    data)})."""
    buf = Path(path).read_bytes()
    if buf[:3] != b'CDF' or buf[3] not in (1, 2, 5):
        raise ValueError('not a classic netCDF file')
    ver = buf[3]
    pos = 4
    w = 8 if ver == 5 else 4

    def nn():
        nonlocal pos
        v = struct.unpack_from('>q' if w == 8 else '>i', buf, pos)[0]
        pos += w
        return v

    def i32():
        nonlocal pos
        v = struct.unpack_from('>i', buf, pos)[0]
        pos += 4
        return v

    def name():
        nonlocal pos
        n = nn()
        s = buf[pos:pos + n].decode('utf-8')
        pos += n + _pad4(n)
        return s

    def atts():
        nonlocal pos
        tag, n = i32(), nn()
        out = {}
        for _ in range(n if tag == _NC_ATTRIBUTE else 0):
            k = name()
            t, cnt = i32(), nn()
            dt = _NC_TO_DTYPE[t]
            nb = cnt * dt.itemsize
            raw = buf[pos:pos + nb]
            pos += nb + _pad4(nb)
            out[k] = raw.decode('utf-8') if t == 2 else np.frombuffer(raw, dt)
        return out

    nn()  # numrecs
    tag, nd = i32(), nn()
    dims = []
    for _ in range(nd if tag == _NC_DIMENSION else 0):
        dims.append((name(), nn()))
    gatts = atts()
    tag, nv = i32(), nn()
    out = {}
    for _ in range(nv if tag == _NC_VARIABLE else 0):
        vn = name()
        ndim = nn()
        dids = [nn() for _ in range(ndim)]
        vatts = atts()
        t = i32()
        if ver == 5:
            nn()
        else:
            struct.unpack_from('>I', buf, pos)
            pos += 4
        begin = struct.unpack_from('>q' if ver != 1 else '>i', buf, pos)[0]
        pos += 8 if ver != 1 else 4
        shape = tuple(dims[d][1] for d in dids)
        dt = _NC_TO_DTYPE[t]
        cnt = int(np.prod(shape)) if shape else 1
        data = np.frombuffer(buf, dt, cnt, begin).reshape(shape)
        out[vn] = (tuple(dims[d][0] for d in dids), vatts, data)
    return dict(dims), gatts, out


def reference_attrs(interp_args, nc_vlab, nc_vunits, settings=None,
                    source=None):
    """Variable and global attributes as interp/prepare.py:355-430 sets
    them: interp_args are the reference's (type, ..., label[, idw_exp])
    tuples (label third, IDW exponent fourth); settings {name: value}
    become ``sett_<name>`` strings."""
    var_attrs = {}
    for arg in interp_args:
        lab = arg[2]
        if arg[0] == 'IDW':
            sn = nc_vlab + f' ({lab[:3]}_exp_{arg[3]})'
        else:
            sn = nc_vlab + f' ({lab})'
        var_attrs[lab] = {'units': nc_vunits, 'standard_name': sn}
    gatts = {f'sett_{k}': str(v) for k, v in (settings or {}).items()}
    if source is not None:
        gatts['Source'] = str(source)
    return var_attrs, gatts
