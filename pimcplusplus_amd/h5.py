"""HDF5 files in the layout of PIMC++'s IO layer, over the HDF5 C library (ctypes).

PIMC++ stores its pair-action fits through IOSectionClass (Common/IO/IOHDF5.cc,
Common/IO/IOVarHDF5.h), which maps onto HDF5 as follows:

* a section is a group; sections that share a name carry a ``_<n>`` suffix, which the
  reader strips back into (name, number) (IOHDF5.cc:54-79, NumExtension :83-95);
* a variable is a dataset: a scalar is a one-element rank-1 dataset (NewIOVar0HDF5,
  IOVarHDF5.h:325-345; VarRead asserts the single element, :562-568), a string a
  fixed-length C string, an array its blitz (row-major) shape with the element type
  H5T_NATIVE_DOUBLE or H5T_NATIVE_INT (IOVarHDF5.h:195-240, 577-601).

This module reads and writes that subset: groups, and datasets of doubles, ints and
fixed-length strings.  The HDF5 library is found through PIMC_HDF5_LIBRARY, the loader's
search path, or a conda prefix; without it every call raises H5Error.
"""

from __future__ import annotations

import ctypes
import ctypes.util
import os

import numpy as np

hid_t = ctypes.c_int64
hsize_t = ctypes.c_ulonglong
herr_t = ctypes.c_int

H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0x0000, 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5I_GROUP, H5I_DATASET = 2, 5
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5_INDEX_NAME, H5_ITER_INC = 0, 0


class H5Error(RuntimeError):
    pass


_lib = None
_types: dict[str, int] = {}


def _candidates() -> list[str]:
    out = []
    env = os.environ.get("PIMC_HDF5_LIBRARY")
    if env:
        out.append(env)
    found = ctypes.util.find_library("hdf5")
    if found:
        out.append(found)
    for prefix in (os.environ.get("CONDA_PREFIX"), "/opt/conda", "/usr", "/usr/local"):
        if prefix:
            for d in ("lib", "lib64", "lib/x86_64-linux-gnu", "lib/x86_64-linux-gnu/hdf5/serial"):
                out.append(os.path.join(prefix, d, "libhdf5.so"))
    return out


def library() -> ctypes.CDLL:
    """The HDF5 C library with the signatures this module uses."""
    global _lib
    if _lib is not None:
        return _lib
    lib = None
    for path in _candidates():
        try:
            lib = ctypes.CDLL(path)
            break
        except OSError:
            continue
    if lib is None:
        raise H5Error("the HDF5 C library (libhdf5.so) was not found; set PIMC_HDF5_LIBRARY")
    sig = {
        "H5open": (herr_t, []),
        "H5Eset_auto2": (herr_t, [hid_t, ctypes.c_void_p, ctypes.c_void_p]),
        "H5Fopen": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
        "H5Fcreate": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]),
        "H5Fclose": (herr_t, [hid_t]),
        "H5Gopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Gcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t]),
        "H5Gclose": (herr_t, [hid_t]),
        "H5Dopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Dcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dclose": (herr_t, [hid_t]),
        "H5Dget_space": (hid_t, [hid_t]),
        "H5Dget_type": (hid_t, [hid_t]),
        "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Screate_simple": (hid_t, [ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Sclose": (herr_t, [hid_t]),
        "H5Tcopy": (hid_t, [hid_t]),
        "H5Tset_size": (herr_t, [hid_t, ctypes.c_size_t]),
        "H5Tget_size": (ctypes.c_size_t, [hid_t]),
        "H5Tget_class": (ctypes.c_int, [hid_t]),
        "H5Tis_variable_str": (ctypes.c_int, [hid_t]),
        "H5Tclose": (herr_t, [hid_t]),
        "H5Lget_name_by_idx": (ctypes.c_ssize_t, [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t,
                                                  ctypes.c_char_p, ctypes.c_size_t, hid_t]),
        "H5Lexists": (ctypes.c_int, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Oopen": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Oclose": (herr_t, [hid_t]),
        "H5Iget_type": (ctypes.c_int, [hid_t]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if lib.H5open() < 0:
        raise H5Error("H5open failed")
    lib.H5Eset_auto2(0, None, None)   # errors come back as return codes, not printed
    for t in ("H5T_NATIVE_DOUBLE_g", "H5T_NATIVE_INT_g", "H5T_C_S1_g"):
        _types[t] = hid_t.in_dll(lib, t).value
    _lib = lib
    return lib


def _check(v: int, what: str) -> int:
    if v < 0:
        raise H5Error(f"HDF5: {what} failed")
    return v


def split_section_name(name: str) -> tuple[str, int]:
    """(section name, number) of a group name: a trailing _<n> is the number, else 0
    (IOTreeHDF5Class::StripName, IOHDF5.cc:54-79)."""
    pos = name.rfind("_")
    if pos <= 0:
        return name, 0
    tail = name[pos + 1:]
    return (name[:pos], int(tail)) if tail.isdigit() else (name[:pos], -1)


class Group:
    """One open group (a section)."""

    def __init__(self, gid: int, is_file: bool = False):
        self.id = gid
        self._file = is_file

    # ---- structure ----
    def names(self) -> list[str]:
        lib = library()
        out = []
        idx = 0
        buf = ctypes.create_string_buffer(1024)
        while True:
            n = lib.H5Lget_name_by_idx(self.id, b".", H5_INDEX_NAME, H5_ITER_INC, idx, buf, 1024, H5P_DEFAULT)
            if n < 0:
                break
            out.append(buf.value.decode())
            idx += 1
        return out

    def _kind(self, name: str) -> int:
        lib = library()
        if lib.H5Lexists(self.id, name.encode(), H5P_DEFAULT) <= 0:
            return -1
        oid = lib.H5Oopen(self.id, name.encode(), H5P_DEFAULT)
        if oid < 0:
            return -1
        kind = lib.H5Iget_type(oid)
        lib.H5Oclose(oid)
        return kind

    def has_group(self, name: str) -> bool:
        return self._kind(name) == H5I_GROUP

    def has_var(self, name: str) -> bool:
        return self._kind(name) == H5I_DATASET

    def group(self, name: str) -> "Group":
        if not self.has_group(name):
            raise H5Error(f"HDF5: no section '{name}'")
        return Group(_check(library().H5Gopen2(self.id, name.encode(), H5P_DEFAULT), f"open group {name}"))

    def sections(self, name: str) -> list[tuple[int, "Group"]]:
        """The sections called `name` (groups `name` or `name_<n>`), as (number, group) in
        number order (IOSectionClass::CountSections / OpenSection(name, i))."""
        found = []
        for member in self.names():
            base, num = split_section_name(member)
            if base == name and num >= 0 and self.has_group(member):
                found.append((num, member))
        return [(num, self.group(member)) for num, member in sorted(found)]

    # ---- variables ----
    def read(self, name: str):
        """A variable: a float or int for a one-element dataset, a str for a string, else
        a numpy array of its shape (float64 or int32)."""
        lib = library()
        if not self.has_var(name):
            raise H5Error(f"HDF5: no variable '{name}'")
        did = _check(lib.H5Dopen2(self.id, name.encode(), H5P_DEFAULT), f"open dataset {name}")
        sid = tid = -1
        try:
            sid = _check(lib.H5Dget_space(did), "get_space")
            nd = lib.H5Sget_simple_extent_ndims(sid)
            dims = (hsize_t * max(nd, 1))()
            lib.H5Sget_simple_extent_dims(sid, dims, None)
            shape = tuple(int(dims[i]) for i in range(nd))
            tid = _check(lib.H5Dget_type(did), "get_type")
            cls = lib.H5Tget_class(tid)
            count = int(np.prod(shape)) if shape else 1
            if cls == H5T_STRING:
                if lib.H5Tis_variable_str(tid) > 0:
                    raise H5Error(f"HDF5: variable-length string '{name}' is not supported")
                size = lib.H5Tget_size(tid)
                buf = ctypes.create_string_buffer(size * count)
                _check(lib.H5Dread(did, tid, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), f"read {name}")
                vals = [buf.raw[i * size:(i + 1) * size].split(b"\0", 1)[0].decode() for i in range(count)]
                return vals[0] if count == 1 else vals
            if cls == H5T_FLOAT:
                arr, mem = np.empty(shape, dtype=np.float64), _types["H5T_NATIVE_DOUBLE_g"]
            elif cls == H5T_INTEGER:
                arr, mem = np.empty(shape, dtype=np.int32), _types["H5T_NATIVE_INT_g"]
            else:
                raise H5Error(f"HDF5: variable '{name}' has an unsupported type class {cls}")
            _check(lib.H5Dread(did, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data), f"read {name}")
            if shape == (1,):
                return arr.item()
            return arr
        finally:
            if tid >= 0:
                lib.H5Tclose(tid)
            if sid >= 0:
                lib.H5Sclose(sid)
            lib.H5Dclose(did)

    # ---- writing (the same layout) ----
    def new_group(self, name: str) -> "Group":
        lib = library()
        return Group(_check(lib.H5Gcreate2(self.id, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                            f"create group {name}"))

    def new_section(self, name: str, number: int | None = None) -> "Group":
        """A section, with the _<n> suffix when it is one of several (NumExtension)."""
        return self.new_group(name if number is None else f"{name}_{number}")

    def write(self, name: str, value) -> None:
        lib = library()
        owned_type = False
        if isinstance(value, str):
            data = value.encode() + b"\0"
            tid = _check(lib.H5Tcopy(_types["H5T_C_S1_g"]), "copy string type")
            lib.H5Tset_size(tid, len(data))
            owned_type = True
            shape = (1,)
            buf = ctypes.create_string_buffer(data, len(data))
            ptr = buf
        else:
            arr = np.asarray(value)
            if arr.dtype.kind in "iub":
                arr = np.ascontiguousarray(arr, dtype=np.int32)
                tid = _types["H5T_NATIVE_INT_g"]
            else:
                arr = np.ascontiguousarray(arr, dtype=np.float64)
                tid = _types["H5T_NATIVE_DOUBLE_g"]
            shape = arr.shape if arr.ndim else (1,)
            ptr = arr.ctypes.data
        dims = (hsize_t * len(shape))(*shape)
        sid = _check(lib.H5Screate_simple(len(shape), dims, None), "create dataspace")
        try:
            did = _check(lib.H5Dcreate2(self.id, name.encode(), tid, sid, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                         f"create dataset {name}")
            try:
                _check(lib.H5Dwrite(did, tid, H5S_ALL, H5S_ALL, H5P_DEFAULT, ptr), f"write {name}")
            finally:
                lib.H5Dclose(did)
        finally:
            lib.H5Sclose(sid)
            if owned_type:
                lib.H5Tclose(tid)

    def close(self) -> None:
        if self.id >= 0:
            lib = library()
            (lib.H5Fclose if self._file else lib.H5Gclose)(self.id)
            self.id = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def open_file(path: str) -> Group:
    """The root section of an existing file, read-only."""
    lib = library()
    fid = lib.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, H5P_DEFAULT)
    if fid < 0:
        raise H5Error(f"HDF5: cannot open {path}")
    return Group(fid, is_file=True)


def new_file(path: str) -> Group:
    """The root section of a new file (an existing one is replaced)."""
    lib = library()
    fid = lib.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
    if fid < 0:
        raise H5Error(f"HDF5: cannot create {path}")
    return Group(fid, is_file=True)
