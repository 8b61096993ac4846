"""ctypes binding of ``libusct_h5.so`` (include/usct_h5.h): the reader and
writer of the HDF5 files that carry acquisition datasets.

:class:`Reader` stands where the reference opens ``h5py.File(path, "r")``
(simulation/ringfwi/dataset.py:142-174) and :class:`Writer` where it opens
``h5py.File(path, "w")`` (dataset.py:112-139).  Errors follow h5py's: a
missing file is ``FileNotFoundError``, a damaged or non-HDF5 file ``OSError``,
a missing object or attribute ``KeyError``; HDF5 features outside the subset
the library reads raise ``NotImplementedError``.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libusct_h5.so")

OK, EINVAL, ENOMEM, EIO, ENOTSUP, EFORMAT, ENOENT = range(7)
F32, F64, I32, I64, U8, STR, I8, I16, U16, U32, U64, F16 = range(1, 13)
OTHER = 99
ABSENT, GROUP, DATASET = 0, 1, 2
MAX_RANK = 32

CODE_DTYPE = {F32: np.float32, F64: np.float64, I32: np.int32, I64: np.int64, U8: np.uint8,
              I8: np.int8, I16: np.int16, U16: np.uint16, U32: np.uint32, U64: np.uint64,
              F16: np.float16}
DTYPE_CODE = {np.dtype(v): k for k, v in CODE_DTYPE.items()}
WRITE_CODES = (F32, F64, I32, I64, U8)

# Every symbol declared in include/usct_h5.h.
EXPORTS = (
    "usct_h5_open", "usct_h5_close", "usct_h5_kind", "usct_h5_list", "usct_h5_dataset_info",
    "usct_h5_read", "usct_h5_attr_list", "usct_h5_attr_info", "usct_h5_attr_read",
    "usct_h5_create", "usct_h5_add_group", "usct_h5_add_dataset", "usct_h5_add_attr",
    "usct_h5_finish", "usct_h5_last_error",
)

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run openusct_amd.build.build_h5()")
    lib = ctypes.CDLL(LIB_PATH)
    P, I, L, C = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_char_p
    PP, IP, LP = ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)
    sig = {
        "usct_h5_open": [PP, C], "usct_h5_close": [P], "usct_h5_kind": [P, C, IP],
        "usct_h5_list": [P, C, P, L, LP], "usct_h5_dataset_info": [P, C, IP, LP, IP, IP],
        "usct_h5_read": [P, C, L, L, I, P, L, I], "usct_h5_attr_list": [P, C, P, L, LP],
        "usct_h5_attr_info": [P, C, C, IP, LP, IP], "usct_h5_attr_read": [P, C, C, I, P, L],
        "usct_h5_create": [PP, C], "usct_h5_add_group": [P, C],
        "usct_h5_add_dataset": [P, C, I, I, LP, P, LP, I, I],
        "usct_h5_add_attr": [P, C, C, I, L, P], "usct_h5_finish": [P, I],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.usct_h5_last_error.argtypes = []
    lib.usct_h5_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _check(rc, what=""):
    if rc == OK:
        return
    msg = _lib.usct_h5_last_error().decode("utf-8", "replace")
    if what:
        msg = f"{what}: {msg}"
    if rc == ENOENT:
        raise KeyError(msg)
    if rc == ENOTSUP:
        raise NotImplementedError(msg)
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    raise OSError(msg)


def _b(s):
    return s.encode("utf-8")


def _names(fn, *args):
    need = ctypes.c_int64(0)
    _check(fn(*args, None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(max(int(need.value), 1))
    _check(fn(*args, buf, need.value, ctypes.byref(need)))
    raw = buf.raw[:need.value]
    return [x.decode("utf-8") for x in raw.split(b"\0")[:-1]] if raw else []


class Reader:
    """Read-only handle on an HDF5 file; paths are absolute (``/geometry``)."""

    def __init__(self, path):
        lib = load()
        h = ctypes.c_void_p()
        rc = lib.usct_h5_open(ctypes.byref(h), _b(os.fspath(path)))
        if rc == ENOENT:
            raise FileNotFoundError(lib.usct_h5_last_error().decode("utf-8", "replace"))
        _check(rc)
        self._h = h
        self.path = os.fspath(path)

    def close(self):
        if self._h:
            _lib.usct_h5_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def kind(self, path):
        k = ctypes.c_int(0)
        _check(_lib.usct_h5_kind(self._h, _b(path), ctypes.byref(k)))
        return {ABSENT: None, GROUP: "group", DATASET: "dataset"}[k.value]

    def __contains__(self, path):
        return self.kind(path) is not None

    def members(self, path="/"):
        return _names(_lib.usct_h5_list, self._h, _b(path))

    def info(self, path):
        """(shape, numpy dtype or None, chunked) of a dataset."""
        nd, tc, ch = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        dims = (ctypes.c_int64 * MAX_RANK)()
        _check(_lib.usct_h5_dataset_info(self._h, _b(path), ctypes.byref(nd), dims, ctypes.byref(tc),
                                         ctypes.byref(ch)), path)
        dt = CODE_DTYPE.get(tc.value)
        return tuple(int(dims[i]) for i in range(nd.value)), (np.dtype(dt) if dt else None), bool(ch.value)

    def read(self, path, rows=None, dtype=None, threads=0):
        """The dataset (or rows ``range(*rows)`` of its first axis) as a NumPy
        array of its own element type or of ``dtype``."""
        shape, fdt, _ = self.info(path)
        dt = np.dtype(dtype) if dtype is not None else fdt
        if dt is None or dt not in DTYPE_CODE:
            raise NotImplementedError(f"{path}: element type not readable as {dtype}")
        r0, nr = 0, -1
        if rows is not None:
            if not shape:
                raise ValueError(f"{path} is a scalar")
            r0, r1 = rows
            nr = r1 - r0
            shape = (nr,) + shape[1:]
        out = np.empty(shape, dtype=dt)
        self.read_into(path, out.ctypes.data, out.size, dt, r0, nr, threads)
        return out if shape else out[()]

    def read_into(self, path, ptr, n, dtype, row0=0, nrows=-1, threads=0):
        """Read into memory at address ``ptr`` (``n`` elements of ``dtype``):
        host memory of any kind, e.g. a pinned torch tensor's ``data_ptr()``."""
        code = DTYPE_CODE[np.dtype(dtype)]
        _check(_lib.usct_h5_read(self._h, _b(path), int(row0), int(nrows), code, ctypes.c_void_p(ptr),
                                 int(n), int(threads)), path)

    def attr_names(self, path="/"):
        return _names(_lib.usct_h5_attr_list, self._h, _b(path))

    def attr(self, path, name):
        """An attribute as h5py returns it: ``str`` for strings, a NumPy scalar
        for a scalar, a NumPy array otherwise."""
        tc, n, rank = ctypes.c_int(0), ctypes.c_int64(0), ctypes.c_int(0)
        _check(_lib.usct_h5_attr_info(self._h, _b(path), _b(name), ctypes.byref(tc), ctypes.byref(n),
                                      ctypes.byref(rank)), f"{path}@{name}")
        if tc.value == STR:
            buf = ctypes.create_string_buffer(int(n.value) + 1)
            _check(_lib.usct_h5_attr_read(self._h, _b(path), _b(name), STR, buf, n.value + 1))
            return buf.raw[:n.value].decode("utf-8")
        dt = CODE_DTYPE.get(tc.value)
        if dt is None:
            raise NotImplementedError(f"{path}@{name}: attribute type not readable")
        out = np.empty(max(int(n.value), 1), dtype=dt)
        _check(_lib.usct_h5_attr_read(self._h, _b(path), _b(name), tc.value, out.ctypes.data, out.size))
        return out[0] if rank.value == 0 else out[:n.value]

    def attrs(self, path="/"):
        return {k: self.attr(path, k) for k in self.attr_names(path)}


class Writer:
    """Create an HDF5 file.  Used as a context manager it commits on a clean
    exit and removes the partial file when the block raises."""

    def __init__(self, path):
        lib = load()
        h = ctypes.c_void_p()
        _check(lib.usct_h5_create(ctypes.byref(h), _b(os.fspath(path))))
        self._h = h
        self._keep = []

    def group(self, path):
        _check(_lib.usct_h5_add_group(self._h, _b(path)), path)

    def dataset(self, path, data, chunks=None, gzip=None, threads=0):
        """Store an array (C order) with its own dtype.  ``chunks``: chunk shape
        or None (contiguous); ``gzip``: deflate level 0-9 or None."""
        a = np.asarray(data, order="C")  # keeps 0-d arrays scalar (ascontiguousarray would not)
        if a.dtype == np.bool_:
            a = a.astype(np.uint8)
        code = DTYPE_CODE.get(a.dtype)
        if code not in WRITE_CODES:
            raise TypeError(f"{path}: cannot store dtype {a.dtype}")
        dims = (ctypes.c_int64 * max(a.ndim, 1))(*a.shape)
        ch = None
        if chunks is not None:
            if len(chunks) != a.ndim:
                raise ValueError(f"{path}: chunk rank {len(chunks)} != data rank {a.ndim}")
            ch = (ctypes.c_int64 * a.ndim)(*chunks)
        lvl = -1 if gzip is None else int(gzip)
        if lvl >= 0 and ch is None:
            raise ValueError(f"{path}: compression needs chunks")
        _check(_lib.usct_h5_add_dataset(self._h, _b(path), code, a.ndim, dims,
                                        ctypes.c_void_p(a.ctypes.data if a.size else None), ch, lvl,
                                        int(threads)), path)

    def attr(self, path, name, value):
        """str -> variable-length UTF-8 string; Python int -> int64 scalar;
        float -> float64 scalar; NumPy scalars and arrays keep their type."""
        if isinstance(value, (str, np.str_)):
            _check(_lib.usct_h5_add_attr(self._h, _b(path), _b(name), STR, -1,
                                         ctypes.c_char_p(_b(str(value)))), f"{path}@{name}")
            return
        if isinstance(value, bool):
            value = np.uint8(value)
        elif isinstance(value, int):
            value = np.int64(value)
        elif isinstance(value, float):
            value = np.float64(value)
        a = np.asarray(value)
        if a.dtype == np.bool_:
            a = a.astype(np.uint8)
        code = DTYPE_CODE.get(a.dtype)
        if code not in WRITE_CODES:
            raise TypeError(f"{path}@{name}: cannot store dtype {a.dtype}")
        a = np.asarray(a, order="C")
        n = -1 if a.ndim == 0 else a.size
        if a.ndim > 1:
            raise ValueError(f"{path}@{name}: attributes are scalars or 1-D arrays")
        _check(_lib.usct_h5_add_attr(self._h, _b(path), _b(name), code, n, ctypes.c_void_p(a.ctypes.data)),
               f"{path}@{name}")

    def close(self, commit=True):
        if self._h:
            h, self._h = self._h, None
            _check(_lib.usct_h5_finish(h, 1 if commit else 0))

    def __enter__(self):
        return self

    def __exit__(self, exc_type, *_):
        self.close(commit=exc_type is None)
