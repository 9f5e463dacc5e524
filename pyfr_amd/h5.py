"""A small HDF5 binding (ctypes over the C library, 1.10+) for the native
mesh and solution files: what the reference does through h5py
(``pyfr/readers/native.py``, ``pyfr/readers/base.py:67-105``,
``pyfr/writers/native.py``), for the types those files hold -- integers,
floats, booleans (h5py's enum of int8 FALSE / TRUE), fixed-length byte
strings, fixed-size arrays and compounds of these -- as NumPy arrays.

The library is ``$PYFR_AMD_HDF5_LIBRARY``, else libhdf5 from the Python
prefix or the system's library path.
"""
import ctypes as ct
import ctypes.util
import os
import sys

import numpy as np

_hid, _herr, _hsize = ct.c_int64, ct.c_int, ct.c_ulonglong
H5P_DEFAULT = 0
H5S_ALL = 0
_H5F_ACC_RDONLY, _H5F_ACC_TRUNC = 0, 2
_H5T_INTEGER, _H5T_FLOAT, _H5T_STRING, _H5T_COMPOUND = 0, 1, 3, 6
_H5T_ENUM, _H5T_ARRAY = 8, 10
_H5T_SGN_NONE = 0
_H5T_STR_NULLPAD = 1
_H5I_GROUP, _H5I_DATASET = 2, 5


class H5Error(RuntimeError):
    pass


def _find_lib():
    env = os.environ.get('PYFR_AMD_HDF5_LIBRARY')
    cands = [env] if env else []
    for pre in {sys.prefix, sys.base_prefix, '/opt/conda'}:
        d = os.path.join(pre, 'lib')
        if os.path.isdir(d):
            cands += sorted((os.path.join(d, f) for f in os.listdir(d)
                             if f.startswith('libhdf5.so')), key=len)
    found = ctypes.util.find_library('hdf5')
    if found:
        cands.append(found)
    for c in cands:
        try:
            return ct.CDLL(c)
        except OSError:
            continue
    raise H5Error('libhdf5 not found (set PYFR_AMD_HDF5_LIBRARY)')


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = _find_lib()
        sig = {
            'H5open': ([], _herr), 'H5Fcreate': ([ct.c_char_p, ct.c_uint,
                                                  _hid, _hid], _hid),
            'H5Fopen': ([ct.c_char_p, ct.c_uint, _hid], _hid),
            'H5Fclose': ([_hid], _herr),
            'H5Pcreate': ([_hid], _hid), 'H5Pclose': ([_hid], _herr),
            'H5Pset_create_intermediate_group': ([_hid, ct.c_uint], _herr),
            'H5Screate': ([ct.c_int], _hid),
            'H5Screate_simple': ([ct.c_int, ct.POINTER(_hsize),
                                  ct.c_void_p], _hid),
            'H5Sclose': ([_hid], _herr),
            'H5Sget_simple_extent_ndims': ([_hid], ct.c_int),
            'H5Sget_simple_extent_dims': ([_hid, ct.POINTER(_hsize),
                                           ct.c_void_p], ct.c_int),
            'H5Dcreate2': ([_hid, ct.c_char_p, _hid, _hid, _hid, _hid, _hid],
                           _hid),
            'H5Dopen2': ([_hid, ct.c_char_p, _hid], _hid),
            'H5Dwrite': ([_hid, _hid, _hid, _hid, _hid, ct.c_void_p], _herr),
            'H5Dread': ([_hid, _hid, _hid, _hid, _hid, ct.c_void_p], _herr),
            'H5Dget_type': ([_hid], _hid), 'H5Dget_space': ([_hid], _hid),
            'H5Dclose': ([_hid], _herr),
            'H5Acreate2': ([_hid, ct.c_char_p, _hid, _hid, _hid, _hid], _hid),
            'H5Aopen': ([_hid, ct.c_char_p, _hid], _hid),
            'H5Aexists': ([_hid, ct.c_char_p], ct.c_int),
            'H5Awrite': ([_hid, _hid, ct.c_void_p], _herr),
            'H5Aread': ([_hid, _hid, ct.c_void_p], _herr),
            'H5Aget_type': ([_hid], _hid), 'H5Aget_space': ([_hid], _hid),
            'H5Aclose': ([_hid], _herr),
            'H5Tcreate': ([ct.c_int, ct.c_size_t], _hid),
            'H5Tcopy': ([_hid], _hid), 'H5Tclose': ([_hid], _herr),
            'H5Tset_size': ([_hid, ct.c_size_t], _herr),
            'H5Tset_strpad': ([_hid, ct.c_int], _herr),
            'H5Tinsert': ([_hid, ct.c_char_p, ct.c_size_t, _hid], _herr),
            'H5Tarray_create2': ([_hid, ct.c_uint, ct.POINTER(_hsize)], _hid),
            'H5Tenum_create': ([_hid], _hid),
            'H5Tenum_insert': ([_hid, ct.c_char_p, ct.c_void_p], _herr),
            'H5Tget_class': ([_hid], ct.c_int),
            'H5Tget_size': ([_hid], ct.c_size_t),
            'H5Tget_sign': ([_hid], ct.c_int),
            'H5Tget_nmembers': ([_hid], ct.c_int),
            'H5Tget_member_name': ([_hid, ct.c_uint], ct.c_void_p),
            'H5Tget_member_offset': ([_hid, ct.c_uint], ct.c_size_t),
            'H5Tget_member_type': ([_hid, ct.c_uint], _hid),
            'H5Tget_member_value': ([_hid, ct.c_uint, ct.c_void_p], _herr),
            'H5Tget_array_ndims': ([_hid], ct.c_int),
            'H5Tget_array_dims2': ([_hid, ct.POINTER(_hsize)], ct.c_int),
            'H5Tget_super': ([_hid], _hid),
            'H5Tis_variable_str': ([_hid], ct.c_int),
            'H5free_memory': ([ct.c_void_p], _herr),
            'H5Lexists': ([_hid, ct.c_char_p, _hid], ct.c_int),
            'H5Gopen2': ([_hid, ct.c_char_p, _hid], _hid),
            'H5Gclose': ([_hid], _herr),
            'H5Oopen': ([_hid, ct.c_char_p, _hid], _hid),
            'H5Oclose': ([_hid], _herr),
            'H5Iget_type': ([_hid], ct.c_int),
            'H5Lget_name_by_idx': ([_hid, ct.c_char_p, ct.c_int, ct.c_int,
                                    _hsize, ct.c_char_p, ct.c_size_t, _hid],
                                   ct.c_ssize_t),
            'H5Gget_info': ([_hid, ct.c_void_p], _herr),
            'H5Eset_auto2': ([_hid, ct.c_void_p, ct.c_void_p], _herr),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes, fn.restype = args, res
        if L.H5open() < 0:
            raise H5Error('H5open failed')
        # failures are reported through the return codes, not printed
        L.H5Eset_auto2(0, None, None)
        _lib = L
    return _lib


def _chk(v, what):
    if v < 0:
        raise H5Error(f'HDF5 call failed: {what}')
    return v


def _gl(name):
    return _hid.in_dll(lib(), name + '_g').value


_INTS = {(1, True): 'INT8', (2, True): 'INT16', (4, True): 'INT32',
         (8, True): 'INT64', (1, False): 'UINT8', (2, False): 'UINT16',
         (4, False): 'UINT32', (8, False): 'UINT64'}


# -- NumPy dtype -> HDF5 type --------------------------------------------------

def _h5type(dt):
    """A new HDF5 type (to be closed) for a NumPy dtype."""
    L = lib()
    dt = np.dtype(dt)
    if dt.subdtype is not None:
        base, shape = dt.subdtype
        bt = _h5type(base)
        dims = (_hsize*len(shape))(*shape)
        t = _chk(L.H5Tarray_create2(bt, len(shape), dims), 'array type')
        L.H5Tclose(bt)
        return t
    if dt.names:
        t = _chk(L.H5Tcreate(_H5T_COMPOUND, dt.itemsize), 'compound')
        for n in dt.names:
            sub, off = dt.fields[n][:2]
            mt = _h5type(sub)
            _chk(L.H5Tinsert(t, n.encode(), off, mt), 'insert')
            L.H5Tclose(mt)
        return t
    if dt.kind == 'b':
        t = _chk(L.H5Tenum_create(_gl('H5T_NATIVE_INT8')), 'enum')
        for name, v in ((b'FALSE', 0), (b'TRUE', 1)):
            val = ct.c_int8(v)
            _chk(L.H5Tenum_insert(t, name, ct.byref(val)), 'enum insert')
        return t
    if dt.kind == 'S':
        t = _chk(L.H5Tcopy(_gl('H5T_C_S1')), 'string')
        _chk(L.H5Tset_size(t, max(dt.itemsize, 1)), 'string size')
        _chk(L.H5Tset_strpad(t, _H5T_STR_NULLPAD), 'string pad')
        return t
    if dt.kind in 'iu':
        return L.H5Tcopy(_gl('H5T_NATIVE_' + _INTS[dt.itemsize,
                                                   dt.kind == 'i']))
    if dt.kind == 'f':
        return L.H5Tcopy(_gl({4: 'H5T_NATIVE_FLOAT',
                              8: 'H5T_NATIVE_DOUBLE'}[dt.itemsize]))
    raise H5Error(f'Unsupported dtype {dt}')


# -- HDF5 type -> NumPy dtype --------------------------------------------------

def _npdtype(t):
    """The NumPy dtype of an HDF5 type (its memory layout for reads: a
    compound keeps its members' names and order, packed)."""
    L = lib()
    cls = L.H5Tget_class(t)
    size = L.H5Tget_size(t)
    if cls == _H5T_INTEGER:
        signed = L.H5Tget_sign(t) != _H5T_SGN_NONE
        return np.dtype(f'{"i" if signed else "u"}{size}')
    if cls == _H5T_FLOAT:
        return np.dtype(f'f{size}')
    if cls == _H5T_STRING:
        if L.H5Tis_variable_str(t) > 0:
            raise H5Error('Variable-length strings are not supported')
        return np.dtype(f'S{size}')
    if cls == _H5T_ENUM:
        n = L.H5Tget_nmembers(t)
        names = set()
        for i in range(n):
            p = L.H5Tget_member_name(t, i)
            names.add(ct.string_at(p).decode())
            L.H5free_memory(p)
        if names == {'FALSE', 'TRUE'} and size == 1:
            return np.dtype(bool)
        bt = L.H5Tget_super(t)
        try:
            return _npdtype(bt)
        finally:
            L.H5Tclose(bt)
    if cls == _H5T_ARRAY:
        nd = L.H5Tget_array_ndims(t)
        dims = (_hsize*nd)()
        L.H5Tget_array_dims2(t, dims)
        bt = L.H5Tget_super(t)
        try:
            return np.dtype((_npdtype(bt), tuple(int(d) for d in dims)))
        finally:
            L.H5Tclose(bt)
    if cls == _H5T_COMPOUND:
        fields = []
        for i in range(L.H5Tget_nmembers(t)):
            p = L.H5Tget_member_name(t, i)
            name = ct.string_at(p).decode()
            L.H5free_memory(p)
            mt = L.H5Tget_member_type(t, i)
            try:
                fields.append((name, _npdtype(mt)))
            finally:
                L.H5Tclose(mt)
        return np.dtype(fields)
    raise H5Error(f'Unsupported HDF5 type class {cls}')


def _space(shape):
    L = lib()
    if shape == ():
        return _chk(L.H5Screate(0), 'scalar space')
    dims = (_hsize*len(shape))(*shape)
    return _chk(L.H5Screate_simple(len(shape), dims, None), 'space')


def _shape(space):
    L = lib()
    nd = L.H5Sget_simple_extent_ndims(space)
    if nd <= 0:
        return ()
    dims = (_hsize*nd)()
    L.H5Sget_simple_extent_dims(space, dims, None)
    return tuple(int(d) for d in dims)


class _GInfo(ct.Structure):
    _fields_ = [('storage_type', ct.c_int), ('nlinks', _hsize),
                ('max_corder', ct.c_int64), ('mounted', ct.c_bool)]


def _carray(value):
    arr = np.asarray(value)
    if arr.dtype.kind == 'U':
        arr = arr.astype('S')
    return arr if arr.flags.c_contiguous else arr.copy()


class File:
    """An HDF5 file: ``f[path] = array`` writes a dataset (intermediate
    groups created), ``f[path]`` reads one, ``f.set_attr`` / ``f.attr``
    handle attributes, ``f.keys(group)`` lists a group."""

    def __init__(self, fname, mode='r'):
        L = lib()
        if mode == 'w':
            self.fid = _chk(L.H5Fcreate(os.fsencode(fname), _H5F_ACC_TRUNC,
                                        H5P_DEFAULT, H5P_DEFAULT),
                            f'create {fname}')
        elif mode == 'r':
            self.fid = _chk(L.H5Fopen(os.fsencode(fname), _H5F_ACC_RDONLY,
                                      H5P_DEFAULT), f'open {fname}')
        else:
            raise ValueError('mode must be r or w')

    def close(self):
        if self.fid is not None:
            lib().H5Fclose(self.fid)
            self.fid = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __contains__(self, path):
        L = lib()
        parts = path.strip('/').split('/')
        for i in range(len(parts)):
            sub = '/'.join(parts[:i + 1]).encode()
            if L.H5Lexists(self.fid, sub, H5P_DEFAULT) <= 0:
                return False
        return True

    def __setitem__(self, path, value):
        L = lib()
        arr = _carray(value)
        t, sp = _h5type(arr.dtype), _space(arr.shape)
        lcpl = L.H5Pcreate(_gl('H5P_CLS_LINK_CREATE_ID'))
        L.H5Pset_create_intermediate_group(lcpl, 1)
        try:
            d = _chk(L.H5Dcreate2(self.fid, path.encode(), t, sp, lcpl,
                                  H5P_DEFAULT, H5P_DEFAULT), f'create {path}')
            try:
                _chk(L.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                arr.ctypes.data), f'write {path}')
            finally:
                L.H5Dclose(d)
        finally:
            L.H5Pclose(lcpl)
            L.H5Sclose(sp)
            L.H5Tclose(t)

    def __getitem__(self, path):
        L = lib()
        d = _chk(L.H5Dopen2(self.fid, path.encode(), H5P_DEFAULT),
                 f'open {path}')
        try:
            ft, sp = L.H5Dget_type(d), L.H5Dget_space(d)
            try:
                dt, shape = _npdtype(ft), _shape(sp)
            finally:
                L.H5Tclose(ft)
                L.H5Sclose(sp)
            out = np.empty(shape, dtype=dt)
            mt = _h5type(dt)
            try:
                _chk(L.H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                               out.ctypes.data), f'read {path}')
            finally:
                L.H5Tclose(mt)
            return out
        finally:
            L.H5Dclose(d)

    def set_attr(self, path, name, value):
        L = lib()
        arr = _carray(value)
        o = _chk(L.H5Oopen(self.fid, path.encode(), H5P_DEFAULT),
                 f'open {path}')
        t, sp = _h5type(arr.dtype), _space(arr.shape)
        try:
            a = _chk(L.H5Acreate2(o, name.encode(), t, sp, H5P_DEFAULT,
                                  H5P_DEFAULT), f'attribute {name}')
            try:
                _chk(L.H5Awrite(a, t, arr.ctypes.data), f'write {name}')
            finally:
                L.H5Aclose(a)
        finally:
            L.H5Sclose(sp)
            L.H5Tclose(t)
            L.H5Oclose(o)

    def attr(self, path, name):
        L = lib()
        o = _chk(L.H5Oopen(self.fid, path.encode(), H5P_DEFAULT),
                 f'open {path}')
        try:
            a = _chk(L.H5Aopen(o, name.encode(), H5P_DEFAULT),
                     f'attribute {name}')
            try:
                ft, sp = L.H5Aget_type(a), L.H5Aget_space(a)
                try:
                    dt, shape = _npdtype(ft), _shape(sp)
                finally:
                    L.H5Tclose(ft)
                    L.H5Sclose(sp)
                out = np.empty(shape, dtype=dt)
                mt = _h5type(dt)
                try:
                    _chk(L.H5Aread(a, mt, out.ctypes.data), f'read {name}')
                finally:
                    L.H5Tclose(mt)
                return out
            finally:
                L.H5Aclose(a)
        finally:
            L.H5Oclose(o)

    def keys(self, group='/'):
        """Names in a group, in name order."""
        L = lib()
        g = _chk(L.H5Gopen2(self.fid, group.encode(), H5P_DEFAULT),
                 f'open group {group}')
        try:
            info = _GInfo()
            _chk(L.H5Gget_info(g, ct.byref(info)), f'info {group}')
            out = []
            buf = ct.create_string_buffer(1024)
            for i in range(info.nlinks):
                _chk(L.H5Lget_name_by_idx(g, b'.', 0, 0, i, buf, 1024,
                                          H5P_DEFAULT), f'name in {group}')
                out.append(buf.value.decode())
            return out
        finally:
            L.H5Gclose(g)

    def is_group(self, path):
        L = lib()
        o = _chk(L.H5Oopen(self.fid, path.encode(), H5P_DEFAULT),
                 f'open {path}')
        try:
            return L.H5Iget_type(o) == _H5I_GROUP
        finally:
            L.H5Oclose(o)
