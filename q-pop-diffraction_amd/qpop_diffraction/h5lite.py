"""A small h5py-compatible subset over the HDF5 C library, through ctypes.

The package reads FerroDyn field files and writes ``diffraction.h5`` with the
layout h5py produces (src/python/io_utils.py:95-141 of the reference).  When
h5py is importable it is used as is; otherwise this module binds the HDF5 1.10+
shared library found on the system (``QPOP_HDF5_LIB``, then the usual
locations) and offers the pieces the package needs:

* ``File(path, mode)`` with ``in``, ``[name]`` (a ``Dataset``), ``[name] = array``,
  ``create_dataset(name, data=..., compression="gzip", shuffle=True)``,
  ``attrs`` (scalar numbers and strings; strings are written as variable-length
  UTF-8, as h5py writes them) and ``filename``;
* ``Dataset`` with ``shape``, ``dtype``, ``ndim``, ``len`` and reads of
  unit-step slices (hyperslabs), with conversion to the dtype asked for by
  ``read(dtype=...)`` done inside the library.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

try:  # pragma: no cover - depends on the environment
    import h5py as _h5py
except ImportError:  # pragma: no cover
    _h5py = None

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
hsize_t = ctypes.c_uint64
htri_t = ctypes.c_int

_H5F_ACC_RDONLY = 0
_H5F_ACC_RDWR = 1
_H5F_ACC_TRUNC = 2
_H5P_DEFAULT = 0
_H5S_ALL = 0
_H5S_SCALAR = 0
_H5S_SELECT_SET = 0
_H5T_INTEGER = 0
_H5T_FLOAT = 1
_H5T_STRING = 3
_H5T_SGN_NONE = 0
_H5T_CSET_UTF8 = 1
_H5T_VARIABLE = ctypes.c_size_t(-1).value

_CANDIDATES = (
    "/opt/conda/lib/libhdf5.so",
    "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so",
    "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
    "/usr/lib/x86_64-linux-gnu/libhdf5.so",
    "/usr/local/lib/libhdf5.so",
    "libhdf5.so",
)

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    paths = [os.environ["QPOP_HDF5_LIB"]] if os.environ.get("QPOP_HDF5_LIB") else []
    paths += list(_CANDIDATES)
    errors = []
    for p in paths:
        if p.startswith("/") and not Path(p).exists():
            continue
        try:
            lib = ctypes.CDLL(p)
            break
        except OSError as exc:
            errors.append(f"{p}: {exc}")
    else:
        raise ImportError("Neither h5py nor an HDF5 shared library is available "
                          f"(set QPOP_HDF5_LIB); tried {paths}: {errors}")

    sig = {
        "H5open": (herr_t, []),
        "H5Eset_auto2": (herr_t, [hid_t, ctypes.c_void_p, ctypes.c_void_p]),
        "H5Fopen": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
        "H5Fcreate": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]),
        "H5Fclose": (herr_t, [hid_t]),
        "H5Lexists": (htri_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Ldelete": (herr_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Dopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Dcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dclose": (herr_t, [hid_t]),
        "H5Dget_space": (hid_t, [hid_t]),
        "H5Dget_type": (hid_t, [hid_t]),
        "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Screate": (hid_t, [ctypes.c_int]),
        "H5Screate_simple": (hid_t, [ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Sclose": (herr_t, [hid_t]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Sselect_hyperslab": (herr_t, [hid_t, ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t),
                                         ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Tclose": (herr_t, [hid_t]),
        "H5Tcopy": (hid_t, [hid_t]),
        "H5Tget_class": (ctypes.c_int, [hid_t]),
        "H5Tget_size": (ctypes.c_size_t, [hid_t]),
        "H5Tget_sign": (ctypes.c_int, [hid_t]),
        "H5Tset_size": (herr_t, [hid_t, ctypes.c_size_t]),
        "H5Tset_cset": (herr_t, [hid_t, ctypes.c_int]),
        "H5Tis_variable_str": (htri_t, [hid_t]),
        "H5Pcreate": (hid_t, [hid_t]),
        "H5Pclose": (herr_t, [hid_t]),
        "H5Pset_chunk": (herr_t, [hid_t, ctypes.c_int, ctypes.POINTER(hsize_t)]),
        "H5Pset_shuffle": (herr_t, [hid_t]),
        "H5Pset_deflate": (herr_t, [hid_t, ctypes.c_uint]),
        "H5Aexists": (htri_t, [hid_t, ctypes.c_char_p]),
        "H5Aopen": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Acreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t]),
        "H5Adelete": (herr_t, [hid_t, ctypes.c_char_p]),
        "H5Aclose": (herr_t, [hid_t]),
        "H5Aread": (herr_t, [hid_t, hid_t, ctypes.c_void_p]),
        "H5Awrite": (herr_t, [hid_t, hid_t, ctypes.c_void_p]),
        "H5Aget_type": (hid_t, [hid_t]),
        "H5Aget_space": (hid_t, [hid_t]),
        "H5free_memory": (herr_t, [ctypes.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.H5open() < 0:
        raise ImportError("H5open failed")
    lib.H5Eset_auto2(0, None, None)  # errors become Python exceptions, not stderr dumps
    _lib = lib
    return lib


def _global_id(name):
    return hid_t.in_dll(_load(), name).value


def _native(dtype):
    dtype = np.dtype(dtype)
    table = {
        np.dtype(np.float64): "H5T_NATIVE_DOUBLE_g", np.dtype(np.float32): "H5T_NATIVE_FLOAT_g",
        np.dtype(np.int8): "H5T_NATIVE_SCHAR_g", np.dtype(np.uint8): "H5T_NATIVE_UCHAR_g",
        np.dtype(np.int16): "H5T_NATIVE_SHORT_g", np.dtype(np.uint16): "H5T_NATIVE_USHORT_g",
        np.dtype(np.int32): "H5T_NATIVE_INT_g", np.dtype(np.uint32): "H5T_NATIVE_UINT_g",
        np.dtype(np.int64): "H5T_NATIVE_LLONG_g", np.dtype(np.uint64): "H5T_NATIVE_ULLONG_g",
    }
    if dtype not in table:
        raise TypeError(f"unsupported HDF5 element type {dtype}")
    return _global_id(table[dtype])


def _check(ret, what):
    if ret < 0:
        raise OSError(f"HDF5 call failed: {what}")
    return ret


def _dtype_of(type_id):
    lib = _load()
    cls = lib.H5Tget_class(type_id)
    size = lib.H5Tget_size(type_id)
    if cls == _H5T_FLOAT:
        return np.dtype({4: np.float32, 8: np.float64, 2: np.float16}[size])
    if cls == _H5T_INTEGER:
        signed = lib.H5Tget_sign(type_id) != _H5T_SGN_NONE
        return np.dtype(f"{'i' if signed else 'u'}{size}")
    if cls == _H5T_STRING:
        return np.dtype(object)
    raise TypeError(f"unsupported HDF5 type class {cls}")


def _dims(space):
    lib = _load()
    rank = lib.H5Sget_simple_extent_ndims(space)
    if rank <= 0:
        return ()
    dims = (hsize_t * rank)()
    lib.H5Sget_simple_extent_dims(space, dims, None)
    return tuple(int(d) for d in dims)


class Dataset:
    """One dataset of an open file; reads hyperslabs of unit step."""

    def __init__(self, file, name):
        lib = _load()
        self.file = file
        self.name = name
        self._id = _check(lib.H5Dopen2(file._id, name.encode(), _H5P_DEFAULT), f"open dataset {name}")
        space = lib.H5Dget_space(self._id)
        self.shape = _dims(space)
        lib.H5Sclose(space)
        t = lib.H5Dget_type(self._id)
        self.dtype = _dtype_of(t)
        lib.H5Tclose(t)

    @property
    def ndim(self):
        return len(self.shape)

    def __len__(self):
        return self.shape[0]

    def close(self):
        if self._id is not None:
            _load().H5Dclose(self._id)
            self._id = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def _selection(self, key):
        if key is Ellipsis or key == () or (isinstance(key, slice) and key == slice(None)):
            key = ()
        if not isinstance(key, tuple):
            key = (key,)
        start, count, squeeze = [], [], []
        for axis, n in enumerate(self.shape):
            k = key[axis] if axis < len(key) else slice(None)
            if isinstance(k, (int, np.integer)):
                k = int(k) + (n if k < 0 else 0)
                if not 0 <= k < n:
                    raise IndexError(f"index {k} out of range for axis {axis} of size {n}")
                start.append(k)
                count.append(1)
                squeeze.append(axis)
                continue
            lo, hi, step = k.indices(n)
            if step != 1:
                raise ValueError("only unit-step slices are supported")
            start.append(lo)
            count.append(max(0, hi - lo))
        return start, count, squeeze

    def read(self, key=(), dtype=None):
        lib = _load()
        if self.dtype == np.dtype(object):
            raise TypeError("string datasets are not supported")
        out_dtype = np.dtype(dtype) if dtype is not None else self.dtype
        if not self.shape:
            out = np.empty((), dtype=out_dtype)
            _check(lib.H5Dread(self._id, _native(out_dtype), _H5S_ALL, _H5S_ALL, _H5P_DEFAULT,
                               out.ctypes.data), f"read {self.name}")
            return out
        start, count, squeeze = self._selection(key)
        out = np.empty(count, dtype=out_dtype)
        if out.size == 0:
            return out.squeeze(axis=tuple(squeeze)) if squeeze else out
        rank = len(self.shape)
        c_start = (hsize_t * rank)(*start)
        c_count = (hsize_t * rank)(*count)
        fspace = lib.H5Dget_space(self._id)
        try:
            _check(lib.H5Sselect_hyperslab(fspace, _H5S_SELECT_SET, c_start, None, c_count, None), "select")
            mspace = lib.H5Screate_simple(rank, c_count, None)
            try:
                _check(lib.H5Dread(self._id, _native(out_dtype), mspace, fspace, _H5P_DEFAULT,
                                   out.ctypes.data), f"read {self.name}")
            finally:
                lib.H5Sclose(mspace)
        finally:
            lib.H5Sclose(fspace)
        return out.squeeze(axis=tuple(squeeze)) if squeeze else out

    def __getitem__(self, key):
        return self.read(key)

    def __setitem__(self, key, data):
        self.write(key, data)

    def write(self, key, data):
        """Write ``data`` into the selection ``key`` (slices of unit step)."""
        lib = _load()
        start, count, _ = self._selection(key)
        arr = np.ascontiguousarray(np.asarray(data, dtype=self.dtype)).reshape(count)
        rank = len(self.shape)
        c_start = (hsize_t * rank)(*start)
        c_count = (hsize_t * rank)(*count)
        fspace = lib.H5Dget_space(self._id)
        try:
            _check(lib.H5Sselect_hyperslab(fspace, _H5S_SELECT_SET, c_start, None, c_count, None), "select")
            mspace = lib.H5Screate_simple(rank, c_count, None)
            try:
                _check(lib.H5Dwrite(self._id, _native(arr.dtype), mspace, fspace, _H5P_DEFAULT, arr.ctypes.data),
                       f"write {self.name}")
            finally:
                lib.H5Sclose(mspace)
        finally:
            lib.H5Sclose(fspace)

    def __array__(self, dtype=None, copy=None):
        return self.read(dtype=dtype)


class _Attrs:
    def __init__(self, file):
        self._file = file

    def __contains__(self, name):
        return _load().H5Aexists(self._file._id, name.encode()) > 0

    def __getitem__(self, name):
        lib = _load()
        if name not in self:
            raise KeyError(name)
        aid = _check(lib.H5Aopen(self._file._id, name.encode(), _H5P_DEFAULT), f"open attribute {name}")
        try:
            t = lib.H5Aget_type(aid)
            space = lib.H5Aget_space(aid)
            try:
                shape = _dims(space)
                if lib.H5Tget_class(t) == _H5T_STRING:
                    if lib.H5Tis_variable_str(t) > 0:
                        ptr = ctypes.c_char_p()
                        _check(lib.H5Aread(aid, t, ctypes.byref(ptr)), "read string attribute")
                        value = ptr.value.decode("utf-8")
                        lib.H5free_memory(ptr)
                        return value
                    size = lib.H5Tget_size(t)
                    buf = ctypes.create_string_buffer(size + 1)
                    _check(lib.H5Aread(aid, t, buf), "read string attribute")
                    return np.bytes_(buf.raw[:size].rstrip(b"\0"))
                dtype = _dtype_of(t)
                out = np.empty(shape, dtype=dtype)
                _check(lib.H5Aread(aid, _native(dtype), out.ctypes.data), f"read attribute {name}")
                return out[()] if not shape else out
            finally:
                lib.H5Sclose(space)
                lib.H5Tclose(t)
        finally:
            lib.H5Aclose(aid)

    def __setitem__(self, name, value):
        lib = _load()
        fid = self._file._id
        if name in self:
            lib.H5Adelete(fid, name.encode())
        space = lib.H5Screate(_H5S_SCALAR)
        try:
            if isinstance(value, (bytes, np.bytes_)):
                # fixed-length ASCII, as h5py stores numpy bytes
                raw = bytes(value)
                t = lib.H5Tcopy(_global_id("H5T_C_S1_g"))
                lib.H5Tset_size(t, max(1, len(raw)))
                try:
                    aid = _check(lib.H5Acreate2(fid, name.encode(), t, space, _H5P_DEFAULT, _H5P_DEFAULT), name)
                    buf = ctypes.create_string_buffer(raw, max(1, len(raw)))
                    _check(lib.H5Awrite(aid, t, buf), name)
                    lib.H5Aclose(aid)
                finally:
                    lib.H5Tclose(t)
                return
            if isinstance(value, (str, np.str_)):
                text = str(value)
                t = lib.H5Tcopy(_global_id("H5T_C_S1_g"))
                lib.H5Tset_size(t, _H5T_VARIABLE)
                lib.H5Tset_cset(t, _H5T_CSET_UTF8)
                try:
                    aid = _check(lib.H5Acreate2(fid, name.encode(), t, space, _H5P_DEFAULT, _H5P_DEFAULT), name)
                    ptr = ctypes.c_char_p(text.encode("utf-8"))
                    _check(lib.H5Awrite(aid, t, ctypes.byref(ptr)), name)
                    lib.H5Aclose(aid)
                finally:
                    lib.H5Tclose(t)
                return
            arr = np.asarray(value)
            if arr.shape != ():
                raise TypeError("only scalar attributes are supported")
            if arr.dtype.kind == "b":
                arr = arr.astype(np.int8)
            elif arr.dtype.kind in "iu":
                arr = arr.astype(np.int64)
            else:
                arr = arr.astype(np.float64)
            arr = np.ascontiguousarray(arr)
            native = _native(arr.dtype)
            aid = _check(lib.H5Acreate2(fid, name.encode(), native, space, _H5P_DEFAULT, _H5P_DEFAULT), name)
            _check(lib.H5Awrite(aid, native, arr.ctypes.data), name)
            lib.H5Aclose(aid)
        finally:
            lib.H5Sclose(space)


class File:
    """An HDF5 file: mode "r", "r+", "w" or "a"."""

    def __init__(self, path, mode="r"):
        lib = _load()
        self.filename = str(path)
        raw = os.fsencode(self.filename)
        if mode == "r":
            self._id = lib.H5Fopen(raw, _H5F_ACC_RDONLY, _H5P_DEFAULT)
        elif mode == "r+":
            self._id = lib.H5Fopen(raw, _H5F_ACC_RDWR, _H5P_DEFAULT)
        elif mode == "w":
            self._id = lib.H5Fcreate(raw, _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT)
        elif mode == "a":
            if Path(self.filename).exists():
                self._id = lib.H5Fopen(raw, _H5F_ACC_RDWR, _H5P_DEFAULT)
            else:
                self._id = lib.H5Fcreate(raw, _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT)
        else:
            raise ValueError(f"unsupported mode {mode!r}")
        if self._id < 0:
            raise OSError(f"Unable to open file {self.filename!r} (mode {mode!r})")
        self.attrs = _Attrs(self)
        self._datasets = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if getattr(self, "_id", None) is not None and self._id >= 0:
            for ds in self._datasets:
                ds.close()
            self._datasets = []
            _load().H5Fclose(self._id)
            self._id = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __contains__(self, name):
        return _load().H5Lexists(self._id, name.encode(), _H5P_DEFAULT) > 0

    def __getitem__(self, name):
        if name not in self:
            raise KeyError(f"Unable to open object {name!r} in {self.filename}")
        ds = Dataset(self, name)
        self._datasets.append(ds)
        return ds

    def __delitem__(self, name):
        _check(_load().H5Ldelete(self._id, name.encode(), _H5P_DEFAULT), f"delete {name}")

    def __setitem__(self, name, data):
        self.create_dataset(name, data=data)

    def create_dataset(self, name, data=None, compression=None, shuffle=False, chunks=None, compression_opts=None,
                       shape=None, dtype=None):
        """Create (and write) a dataset; with data None, an unwritten one of the given shape and dtype."""
        lib = _load()
        if data is None:
            rank = len(shape)
            native = _native(np.dtype(dtype))
            space = lib.H5Screate_simple(rank, (hsize_t * rank)(*shape), None)
            try:
                did = _check(lib.H5Dcreate2(self._id, name.encode(), native, space, _H5P_DEFAULT, _H5P_DEFAULT,
                                            _H5P_DEFAULT), f"create dataset {name}")
                lib.H5Dclose(did)
            finally:
                lib.H5Sclose(space)
            return self[name]
        arr = np.ascontiguousarray(np.asarray(data))
        if arr.dtype.kind == "b":
            arr = arr.astype(np.int8)
        rank = arr.ndim
        native = _native(arr.dtype)
        if rank == 0:
            space = lib.H5Screate(_H5S_SCALAR)
        else:
            dims = (hsize_t * rank)(*arr.shape)
            space = lib.H5Screate_simple(rank, dims, None)
        dcpl = _H5P_DEFAULT
        if (compression or shuffle or chunks) and rank > 0 and arr.size > 0:
            dcpl = lib.H5Pcreate(_global_id("H5P_CLS_DATASET_CREATE_ID_g"))
            if chunks is None or chunks is True:
                # about 64 KiB chunks that keep whole fastest rows
                target = max(1, 65536 // arr.itemsize)
                chunk = list(arr.shape)
                for axis in range(rank - 1):
                    inner = int(np.prod(chunk[axis + 1:]))
                    chunk[axis] = max(1, min(chunk[axis], target // max(1, inner)))
            else:
                chunk = list(chunks)
            c_chunk = (hsize_t * rank)(*chunk)
            lib.H5Pset_chunk(dcpl, rank, c_chunk)
            if shuffle:
                lib.H5Pset_shuffle(dcpl)
            if compression in ("gzip", 1, True):
                lib.H5Pset_deflate(dcpl, 4 if compression_opts is None else int(compression_opts))
        try:
            did = _check(lib.H5Dcreate2(self._id, name.encode(), native, space, _H5P_DEFAULT, dcpl, _H5P_DEFAULT),
                         f"create dataset {name}")
            try:
                if arr.size:
                    _check(lib.H5Dwrite(did, native, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, arr.ctypes.data),
                           f"write {name}")
            finally:
                lib.H5Dclose(did)
        finally:
            if dcpl != _H5P_DEFAULT:
                lib.H5Pclose(dcpl)
            lib.H5Sclose(space)


def open_file(path, mode="r"):
    """An h5py.File when h5py is importable, else this module's File."""
    if _h5py is not None and not os.environ.get("QPOP_FORCE_H5LITE"):
        return _h5py.File(path, mode)
    return File(path, mode)


def read_dataset(fh, name, key=(), dtype=np.float64):
    """Read ``fh[name][key]`` converted to ``dtype`` (h5py or this module)."""
    ds = fh[name]
    if isinstance(ds, Dataset):
        return ds.read(key, dtype=dtype)
    return np.asarray(ds[key] if key != () else ds[()], dtype=dtype)
