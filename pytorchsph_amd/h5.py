"""A small HDF5 writer and reader over the HDF5 C library (ctypes), with the
part of h5py's interface that the reference's export code uses
(simulationBase.py:442-492, 742-780): File(path, "w" | "r"), .attrs[key] =
value, create_group, create_dataset(name, data=array), and reading back
group[name][()] / .attrs / .keys().

Layout choices follow h5py's defaults so h5py readers see the same objects:
contiguous datasets, str attributes as variable-length UTF-8 scalars, bool
as an enum {FALSE, TRUE} over int8, Python int / float as int64 / float64
scalars, numpy arrays with their own dtype.  The library is found through
SPH_HDF5_LIB, else the image's libhdf5.so (conda or system path); without it
every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_CANDIDATES = ("/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
               "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so", "libhdf5.so")

hid_t = C.c_int64
herr_t = C.c_int
hsize_t = C.c_uint64
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5T_VARIABLE = C.c_size_t(-1).value
H5T_CSET_UTF8 = 1
H5T_STR_NULLTERM = 0
H5S_SCALAR = 0
# H5T_class_t
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_ENUM = 0, 1, 3, 8
H5T_SGN_NONE = 0
H5O_TYPE_GROUP, H5O_TYPE_DATASET = 0, 1
H5_INDEX_NAME, H5_ITER_INC = 0, 0
H5O_INFO_BASIC, H5O_INFO_NUM_ATTRS = 0x0001, 0x0004

_lib = None
_types: dict = {}


def lib() -> C.CDLL:
    """The HDF5 C library, loaded once; raises if it is not installed."""
    global _lib
    if _lib is not None:
        return _lib
    paths = [os.environ["SPH_HDF5_LIB"]] if os.environ.get("SPH_HDF5_LIB") else list(_CANDIDATES)
    err = None
    for p in paths:
        try:
            L = C.CDLL(p)
            break
        except OSError as e:
            err = e
    else:
        raise RuntimeError(f"HDF5 C library not found ({err}); set SPH_HDF5_LIB")
    if L.H5open() < 0:
        raise RuntimeError("H5open failed")
    # failures are reported through return codes; no error stacks on stderr
    L.H5Eset_auto2.argtypes = [hid_t, C.c_void_p, C.c_void_p]
    L.H5Eset_auto2(0, None, None)
    sig = {
        "H5Fcreate": ([C.c_char_p, C.c_uint, hid_t, hid_t], hid_t),
        "H5Fopen": ([C.c_char_p, C.c_uint, hid_t], hid_t),
        "H5Fclose": ([hid_t], herr_t),
        "H5Fflush": ([hid_t, C.c_int], herr_t),
        "H5Gcreate2": ([hid_t, C.c_char_p, hid_t, hid_t, hid_t], hid_t),
        "H5Gopen2": ([hid_t, C.c_char_p, hid_t], hid_t),
        "H5Gclose": ([hid_t], herr_t),
        "H5Screate": ([C.c_int], hid_t),
        "H5Screate_simple": ([C.c_int, C.POINTER(hsize_t), C.c_void_p], hid_t),
        "H5Sclose": ([hid_t], herr_t),
        "H5Sget_simple_extent_ndims": ([hid_t], C.c_int),
        "H5Sget_simple_extent_dims": ([hid_t, C.POINTER(hsize_t), C.c_void_p], C.c_int),
        "H5Dcreate2": ([hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t], hid_t),
        "H5Dopen2": ([hid_t, C.c_char_p, hid_t], hid_t),
        "H5Dwrite": ([hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p], herr_t),
        "H5Dread": ([hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p], herr_t),
        "H5Dget_space": ([hid_t], hid_t),
        "H5Dget_type": ([hid_t], hid_t),
        "H5Dclose": ([hid_t], herr_t),
        "H5Acreate2": ([hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t], hid_t),
        "H5Aopen": ([hid_t, C.c_char_p, hid_t], hid_t),
        "H5Awrite": ([hid_t, hid_t, C.c_void_p], herr_t),
        "H5Aread": ([hid_t, hid_t, C.c_void_p], herr_t),
        "H5Aget_space": ([hid_t], hid_t),
        "H5Aget_type": ([hid_t], hid_t),
        "H5Aclose": ([hid_t], herr_t),
        "H5Aexists": ([hid_t, C.c_char_p], C.c_int),
        "H5Adelete": ([hid_t, C.c_char_p], herr_t),
        "H5Oget_info2": ([hid_t, C.c_void_p, C.c_uint], herr_t),
        "H5Aget_name_by_idx": ([hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p,
                                C.c_size_t, hid_t], C.c_ssize_t),
        "H5Tcopy": ([hid_t], hid_t),
        "H5Tclose": ([hid_t], herr_t),
        "H5Tset_size": ([hid_t, C.c_size_t], herr_t),
        "H5Tset_cset": ([hid_t, C.c_int], herr_t),
        "H5Tset_strpad": ([hid_t, C.c_int], herr_t),
        "H5Tget_class": ([hid_t], C.c_int),
        "H5Tget_size": ([hid_t], C.c_size_t),
        "H5Tget_sign": ([hid_t], C.c_int),
        "H5Tis_variable_str": ([hid_t], C.c_int),
        "H5Tenum_create": ([hid_t], hid_t),
        "H5Tenum_insert": ([hid_t, C.c_char_p, C.c_void_p], herr_t),
        "H5Lexists": ([hid_t, C.c_char_p, hid_t], C.c_int),
        "H5Gget_info": ([hid_t, C.c_void_p], herr_t),
        "H5Lget_name_by_idx": ([hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p,
                                C.c_size_t, hid_t], C.c_ssize_t),
        "H5Oget_info_by_name2": ([hid_t, C.c_char_p, C.c_void_p, C.c_uint, hid_t], herr_t),
        "H5free_memory": ([C.c_void_p], herr_t),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res
    for key, sym in (("f4", "H5T_NATIVE_FLOAT_g"), ("f8", "H5T_NATIVE_DOUBLE_g"),
                     ("i1", "H5T_NATIVE_INT8_g"), ("u1", "H5T_NATIVE_UINT8_g"),
                     ("i2", "H5T_NATIVE_INT16_g"), ("u2", "H5T_NATIVE_UINT16_g"),
                     ("i4", "H5T_NATIVE_INT32_g"), ("u4", "H5T_NATIVE_UINT32_g"),
                     ("i8", "H5T_NATIVE_INT64_g"), ("u8", "H5T_NATIVE_UINT64_g"),
                     ("s1", "H5T_C_S1_g")):
        _types[key] = hid_t.in_dll(L, sym).value
    _lib = L
    return L


def _check(rc, what):
    if rc < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return rc


def _mem_type(dt: np.dtype) -> int:
    lib()
    key = dt.kind + str(dt.itemsize)
    if key not in _types:
        raise TypeError(f"HDF5 export: unsupported dtype {dt}")
    return _types[key]


def _bool_type() -> int:
    L = lib()
    t = _check(L.H5Tenum_create(_types["i1"]), "H5Tenum_create")
    for name, v in ((b"FALSE", 0), (b"TRUE", 1)):
        val = C.c_int8(v)
        _check(L.H5Tenum_insert(t, name, C.byref(val)), "H5Tenum_insert")
    return t


def _vstr_type() -> int:
    L = lib()
    t = _check(L.H5Tcopy(_types["s1"]), "H5Tcopy")
    _check(L.H5Tset_size(t, H5T_VARIABLE), "H5Tset_size")
    _check(L.H5Tset_cset(t, H5T_CSET_UTF8), "H5Tset_cset")
    return t


def _space(shape) -> int:
    L = lib()
    if shape == ():
        return _check(L.H5Screate(H5S_SCALAR), "H5Screate")
    dims = (hsize_t * len(shape))(*shape)
    return _check(L.H5Screate_simple(len(shape), dims, None), "H5Screate_simple")


def _read_space(sp) -> tuple:
    L = lib()
    nd = L.H5Sget_simple_extent_ndims(sp)
    if nd <= 0:
        return ()
    dims = (hsize_t * nd)()
    L.H5Sget_simple_extent_dims(sp, dims, None)
    return tuple(int(d) for d in dims)


def _np_of(t) -> np.dtype | str:
    L = lib()
    cls = L.H5Tget_class(t)
    size = L.H5Tget_size(t)
    if cls == H5T_FLOAT:
        return np.dtype(f"f{size}")
    if cls == H5T_INTEGER:
        return np.dtype(("u" if L.H5Tget_sign(t) == H5T_SGN_NONE else "i") + str(size))
    if cls == H5T_ENUM and size == 1:
        return np.dtype(bool)
    if cls == H5T_STRING:
        return "vstr" if L.H5Tis_variable_str(t) > 0 else f"S{size}"
    raise TypeError(f"HDF5 read: unsupported type class {cls}")


class _Attrs:
    """The attribute set of an object, as h5py's .attrs."""

    def __init__(self, oid: int):
        self._id = oid

    def __setitem__(self, key: str, value) -> None:
        L = lib()
        name = key.encode()
        if L.H5Aexists(self._id, name) > 0:
            _check(L.H5Adelete(self._id, name), "H5Adelete")
        if isinstance(value, str):
            buf = C.c_char_p(value.encode())
            mem, own, shape, data = _vstr_type(), True, (), C.byref(buf)
        elif isinstance(value, (bool, np.bool_)):
            arr = np.asarray(value, dtype=np.int8)
            mem, own, shape, data = _bool_type(), True, (), arr.ctypes.data
        else:
            if isinstance(value, (int, np.integer)) and not isinstance(value, np.ndarray):
                arr = np.asarray(value, dtype=np.int64 if isinstance(value, int) else value.dtype)
            elif isinstance(value, float):
                arr = np.asarray(value, dtype=np.float64)
            else:
                arr = np.asarray(value)
                arr = np.ascontiguousarray(arr) if arr.ndim else arr.copy()
            mem, own, shape, data = _mem_type(arr.dtype), False, arr.shape, arr.ctypes.data
        sp = a = -1
        try:
            sp = _space(shape)
            a = _check(L.H5Acreate2(self._id, name, mem, sp, H5P_DEFAULT, H5P_DEFAULT),
                       "H5Acreate2")
            _check(L.H5Awrite(a, mem, data), "H5Awrite")
        finally:
            if a >= 0:
                L.H5Aclose(a)
            if sp >= 0:
                L.H5Sclose(sp)
            if own:
                L.H5Tclose(mem)

    def __getitem__(self, key: str):
        L = lib()
        a = L.H5Aopen(self._id, key.encode(), H5P_DEFAULT)
        if a < 0:
            raise KeyError(key)
        t = L.H5Aget_type(a)
        sp = L.H5Aget_space(a)
        shape = _read_space(sp)
        dt = _np_of(t)
        try:
            if dt == "vstr":
                out = C.c_char_p()
                mt = _vstr_type()
                try:
                    _check(L.H5Aread(a, mt, C.byref(out)), "H5Aread")
                finally:
                    L.H5Tclose(mt)
                val = out.value.decode() if out.value is not None else ""
                L.H5free_memory(C.cast(out, C.c_void_p))
                return val
            if isinstance(dt, str):
                raise TypeError("fixed-length string attributes are not read here")
            if dt == np.dtype(bool):
                arr = np.zeros(shape, dtype=np.int8)
                mt = _bool_type()
                try:
                    _check(L.H5Aread(a, mt, arr.ctypes.data), "H5Aread")
                finally:
                    L.H5Tclose(mt)
                arr = arr.astype(bool)
            else:
                arr = np.zeros(shape, dtype=dt)
                _check(L.H5Aread(a, _mem_type(dt), arr.ctypes.data), "H5Aread")
            return arr[()] if shape == () else arr
        finally:
            L.H5Sclose(sp)
            L.H5Tclose(t)
            L.H5Aclose(a)

    def keys(self) -> list[str]:
        L = lib()
        # H5O_info_t (1.10): fileno, addr, type + rc, four times, num_attrs
        info = (C.c_uint64 * 64)()
        _check(L.H5Oget_info2(self._id, info, H5O_INFO_BASIC | H5O_INFO_NUM_ATTRS),
               "H5Oget_info2")
        n = int(info[7])
        out = []
        for k in range(max(n, 0)):
            buf = C.create_string_buffer(1024)
            L.H5Aget_name_by_idx(self._id, b".", H5_INDEX_NAME, H5_ITER_INC, k, buf, 1024,
                                 H5P_DEFAULT)
            out.append(buf.value.decode())
        return out

    def __contains__(self, key: str) -> bool:
        return lib().H5Aexists(self._id, key.encode()) > 0


class Dataset:
    def __init__(self, did: int):
        self._id = did
        self.attrs = _Attrs(did)

    @property
    def shape(self) -> tuple:
        L = lib()
        sp = L.H5Dget_space(self._id)
        try:
            return _read_space(sp)
        finally:
            L.H5Sclose(sp)

    @property
    def dtype(self) -> np.dtype:
        L = lib()
        t = L.H5Dget_type(self._id)
        try:
            return _np_of(t)
        finally:
            L.H5Tclose(t)

    def __getitem__(self, idx):
        L = lib()
        dt = self.dtype
        if dt == np.dtype(bool):
            arr = np.zeros(self.shape, dtype=np.int8)
            mt = _bool_type()
            try:
                if arr.size:
                    _check(L.H5Dread(self._id, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                     arr.ctypes.data), "H5Dread")
            finally:
                L.H5Tclose(mt)
            arr = arr.astype(bool)
        else:
            arr = np.zeros(self.shape, dtype=dt)
            if arr.size:
                _check(L.H5Dread(self._id, _mem_type(dt), H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                 arr.ctypes.data), "H5Dread")
        return arr if isinstance(idx, tuple) and idx == () else arr[idx]

    def close(self) -> None:
        if self._id >= 0:
            lib().H5Dclose(self._id)
            self._id = -1


class Group:
    def __init__(self, gid: int):
        self._id = gid
        self.attrs = _Attrs(gid)

    def create_group(self, name: str) -> "Group":
        L = lib()
        return Group(_check(L.H5Gcreate2(self._id, name.encode(), H5P_DEFAULT, H5P_DEFAULT,
                                         H5P_DEFAULT), f"H5Gcreate2 {name}"))

    def create_dataset(self, name: str, data=None) -> None:
        """Writes `data` as a contiguous dataset; read it back with
        group[name][()]."""
        L = lib()
        arr = np.asarray(data)
        if arr.dtype == np.dtype(bool):
            mem, own = _bool_type(), True
            arr = np.ascontiguousarray(arr, dtype=np.int8)
        else:
            arr = np.ascontiguousarray(arr)
            mem, own = _mem_type(arr.dtype), False
        sp = d = -1
        try:
            sp = _space(arr.shape)
            d = _check(L.H5Dcreate2(self._id, name.encode(), mem, sp, H5P_DEFAULT, H5P_DEFAULT,
                                    H5P_DEFAULT), f"H5Dcreate2 {name}")
            if arr.size:
                _check(L.H5Dwrite(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data),
                       f"H5Dwrite {name}")
        finally:
            if d >= 0:
                L.H5Dclose(d)
            if sp >= 0:
                L.H5Sclose(sp)
            if own:
                L.H5Tclose(mem)

    def keys(self) -> list[str]:
        L = lib()
        # H5G_info_t: storage_type (int), nlinks (hsize_t), max_corder (int64), mounted (hbool)
        info = (C.c_uint64 * 4)()
        _check(L.H5Gget_info(self._id, info), "H5Gget_info")
        n = int(info[1])
        out = []
        for k in range(n):
            buf = C.create_string_buffer(1024)
            L.H5Lget_name_by_idx(self._id, b".", H5_INDEX_NAME, H5_ITER_INC, k, buf, 1024,
                                 H5P_DEFAULT)
            out.append(buf.value.decode())
        return out

    def __contains__(self, name: str) -> bool:
        return lib().H5Lexists(self._id, name.encode(), H5P_DEFAULT) > 0

    def __getitem__(self, name: str):
        L = lib()
        if name not in self:
            raise KeyError(name)
        g = L.H5Gopen2(self._id, name.encode(), H5P_DEFAULT) if self._is_group(name) else -1
        if g >= 0:
            return Group(g)
        return Dataset(_check(L.H5Dopen2(self._id, name.encode(), H5P_DEFAULT), "H5Dopen2"))

    def _is_group(self, name: str) -> bool:
        L = lib()
        # H5O_info_t begins with fileno (unsigned long), addr (haddr_t), type (H5O_type_t)
        info = (C.c_uint64 * 64)()
        if L.H5Oget_info_by_name2(self._id, name.encode(), info, H5O_INFO_BASIC,
                                  H5P_DEFAULT) < 0:
            return False
        return C.cast(info, C.POINTER(C.c_int))[4] == H5O_TYPE_GROUP

    def close(self) -> None:
        if self._id >= 0:
            lib().H5Gclose(self._id)
            self._id = -1


class File(Group):
    """h5py.File(path, 'w') / h5py.File(path, 'r')."""

    def __init__(self, path, mode: str = "r"):
        L = lib()
        p = os.fspath(path).encode()
        if mode == "w":
            fid = L.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
        elif mode == "r":
            fid = L.H5Fopen(p, H5F_ACC_RDONLY, H5P_DEFAULT)
        else:
            raise ValueError("mode must be 'w' or 'r'")
        super().__init__(_check(fid, f"open {path}"))
        self.filename = os.fspath(path)

    def flush(self) -> None:
        lib().H5Fflush(self._id, 0)

    def close(self) -> None:
        if self._id >= 0:
            lib().H5Fclose(self._id)
            self._id = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
