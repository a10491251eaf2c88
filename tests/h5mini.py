"""A small ctypes reader and writer for HDF5 files (libhdf5 1.10's C API), for
tests: write catalogues the way a two_pcf user would hand them to the hdf5
reader, and read back what the hdf5 writers produced.  No h5py here.

The library is found as the host layer finds it: $TWOPCF_HDF5_LIB,
libhdf5.so, libhdf5.so.103, /opt/conda/lib/libhdf5.so.
"""
import ctypes as C
import os

import numpy as np

_LIB = None
hid_t = C.c_int64
hsize_t = C.c_ulonglong


class _GInfo(C.Structure):
    _fields_ = [("storage_type", C.c_int), ("nlinks", hsize_t), ("max_corder", C.c_int64), ("mounted", C.c_bool)]


def lib():
    """The loaded libhdf5, or None if there is none."""
    global _LIB
    if _LIB is not None:
        return _LIB or None
    for name in (os.environ.get("TWOPCF_HDF5_LIB"), "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"):
        if not name:
            continue
        try:
            L = C.CDLL(name)
            break
        except OSError:
            continue
    else:
        _LIB = False
        return None
    for f in ("H5Fopen", "H5Fcreate", "H5Dopen2", "H5Dcreate2", "H5Dget_space", "H5Dget_type", "H5Screate_simple",
              "H5Gopen2", "H5Aopen", "H5Aget_type", "H5Aget_space", "H5Tcopy", "H5Acreate2", "H5Screate"):
        getattr(L, f).restype = hid_t
    L.H5Tget_size.restype = C.c_size_t
    L.H5Lget_name_by_idx.restype = C.c_ssize_t
    L.H5Aget_name_by_idx.restype = C.c_ssize_t
    L.H5Sget_simple_extent_npoints.restype = C.c_int64
    assert L.H5open() >= 0
    L.H5Eset_auto2(hid_t(0), None, None)  # no error stack printing on expected misses
    _LIB = L
    return L


def _type(name):
    return hid_t.in_dll(lib(), name + "_g").value


def _native(a):
    if a.dtype == np.float64:
        return _type("H5T_NATIVE_DOUBLE")
    if a.dtype == np.float32:
        return _type("H5T_NATIVE_FLOAT")
    if a.dtype == np.int32:
        return _type("H5T_NATIVE_INT")
    if a.dtype in (np.int64, np.uint64):
        return _type("H5T_NATIVE_LLONG")
    raise TypeError(a.dtype)


def write(path, datasets):
    """A new file with one 1-d dataset per (name, array) at its root, each
    stored in the array's own native type."""
    L = lib()
    f = L.H5Fcreate(path.encode(), C.c_uint(2), hid_t(0), hid_t(0))
    assert f >= 0, path
    for name, a in datasets.items():
        a = np.ascontiguousarray(a)
        dims = (hsize_t * 1)(a.size)
        sp = L.H5Screate_simple(1, dims, None)
        t = _native(a)
        d = L.H5Dcreate2(hid_t(f), name.encode(), hid_t(t), hid_t(sp), hid_t(0), hid_t(0), hid_t(0))
        assert d >= 0, name
        assert L.H5Dwrite(hid_t(d), hid_t(t), hid_t(0), hid_t(0), hid_t(0), a.ctypes.data_as(C.c_void_p)) >= 0
        L.H5Dclose(hid_t(d))
        L.H5Sclose(hid_t(sp))
    L.H5Fclose(hid_t(f))


class File:
    """Read-only view of an HDF5 file: groups, datasets (as float64 arrays of
    their stored shape), attributes."""

    def __init__(self, path):
        self.L = lib()
        self.f = self.L.H5Fopen(path.encode(), C.c_uint(0), hid_t(0))
        if self.f < 0:
            raise OSError(path)

    def close(self):
        self.L.H5Fclose(hid_t(self.f))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _group(self, group):
        g = self.L.H5Gopen2(hid_t(self.f), group.encode(), hid_t(0))
        if g < 0:
            raise KeyError(group)
        return g

    def members(self, group="/"):
        """Names of the links in a group, in name order."""
        L = self.L
        g = self._group(group)
        info = _GInfo()
        assert L.H5Gget_info(hid_t(g), C.byref(info)) >= 0
        out = []
        buf = C.create_string_buffer(256)
        for k in range(info.nlinks):
            n = L.H5Lget_name_by_idx(hid_t(g), b".", 0, 0, hsize_t(k), buf, C.c_size_t(256), hid_t(0))
            assert n >= 0
            out.append(buf.value.decode())
        L.H5Gclose(hid_t(g))
        return out

    def dataset(self, group, name):
        """(values as float64, shape, stored element size in bytes)."""
        L = self.L
        g = self._group(group)
        d = L.H5Dopen2(hid_t(g), name.encode(), hid_t(0))
        if d < 0:
            L.H5Gclose(hid_t(g))
            raise KeyError(f"{group}/{name}")
        sp = L.H5Dget_space(hid_t(d))
        nd = L.H5Sget_simple_extent_ndims(hid_t(sp))
        dims = (hsize_t * max(nd, 1))()
        L.H5Sget_simple_extent_dims(hid_t(sp), dims, None)
        shape = tuple(int(dims[k]) for k in range(nd))
        t = L.H5Dget_type(hid_t(d))
        size = int(L.H5Tget_size(hid_t(t)))
        L.H5Tclose(hid_t(t))
        out = np.empty(shape, np.float64)
        assert L.H5Dread(hid_t(d), hid_t(_type("H5T_NATIVE_DOUBLE")), hid_t(0), hid_t(0), hid_t(0),
                         out.ctypes.data_as(C.c_void_p)) >= 0
        L.H5Sclose(hid_t(sp))
        L.H5Dclose(hid_t(d))
        L.H5Gclose(hid_t(g))
        return out, shape, size

    def attr_names(self, group):
        L = self.L
        g = self._group(group)
        out = []
        buf = C.create_string_buffer(256)
        k = 0
        while L.H5Aget_name_by_idx(hid_t(g), b".", 0, 0, hsize_t(k), buf, C.c_size_t(256), hid_t(0)) >= 0:
            out.append(buf.value.decode())
            k += 1
        L.H5Gclose(hid_t(g))
        return sorted(out)

    def attr(self, group, name, kind):
        """kind: 'int', 'double' or 'str' (read in the attribute's own string
        type; returns (bytes, stored size))."""
        L = self.L
        g = self._group(group)
        a = L.H5Aopen(hid_t(g), name.encode(), hid_t(0))
        if a < 0:
            L.H5Gclose(hid_t(g))
            raise KeyError(f"{group}@{name}")
        if kind == "str":
            t = L.H5Aget_type(hid_t(a))
            size = int(L.H5Tget_size(hid_t(t)))
            buf = C.create_string_buffer(size + 1)
            assert L.H5Aread(hid_t(a), hid_t(t), buf) >= 0
            L.H5Tclose(hid_t(t))
            v = (buf.raw[:size], size)
        elif kind == "int":
            x = C.c_int()
            assert L.H5Aread(hid_t(a), hid_t(_type("H5T_NATIVE_INT")), C.byref(x)) >= 0
            v = x.value
        else:
            x = C.c_double()
            assert L.H5Aread(hid_t(a), hid_t(_type("H5T_NATIVE_DOUBLE")), C.byref(x)) >= 0
            v = x.value
        L.H5Aclose(hid_t(a))
        L.H5Gclose(hid_t(g))
        return v

    def attr_size(self, group, name):
        """Stored size in bytes of an attribute's element type."""
        L = self.L
        g = self._group(group)
        a = L.H5Aopen(hid_t(g), name.encode(), hid_t(0))
        t = L.H5Aget_type(hid_t(a))
        size = int(L.H5Tget_size(hid_t(t)))
        L.H5Tclose(hid_t(t))
        L.H5Aclose(hid_t(a))
        L.H5Gclose(hid_t(g))
        return size


def tree(path):
    """Everything an hdf5 file of the two_pcf writers holds, for comparing two
    files: {member of the root: {"datasets": {name: (values as float64,
    shape, stored element size)}, "attrs": {name: (stored bytes, size)}}},
    members and names in name order; a root member that is not a group is
    listed as a dataset of the root ("/")."""
    out = {}
    with File(path) as f:
        for m in f.members("/"):
            try:
                names = f.members(m)
            except KeyError:
                out.setdefault("/", {"datasets": {}, "attrs": {}})["datasets"][m] = f.dataset("/", m)
                continue
            out[m] = {"datasets": {n: f.dataset(m, n) for n in names},
                      "attrs": {a: f.attr(m, a, "str") for a in f.attr_names(m)}}
    return out
