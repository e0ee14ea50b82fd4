"""Read-back of HDF5 files for the tests, through the HDF5 C library (ctypes): the group
listing, and each dataset's class, element size, shape and values."""
import ctypes
import os

import numpy as np

_CANDIDATES = (os.environ.get("AMCAFE_HDF5_LIB"), "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103")


def hdf5():
    for n in _CANDIDATES:
        if not n:
            continue
        try:
            L = ctypes.CDLL(n)
        except OSError:
            continue
        L.H5open()
        return L
    return None


H5T_INTEGER, H5T_FLOAT = 0, 1


class File:
    def __init__(self, path):
        L = self.L = hdf5()
        i64 = ctypes.c_int64
        L.H5Fopen.restype = i64
        L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, i64]
        for f in ("H5Dopen2", "H5Gopen2", "H5Aopen"):
            getattr(L, f).restype = i64
            getattr(L, f).argtypes = [i64, ctypes.c_char_p, i64]
        for f in ("H5Dget_space", "H5Dget_type", "H5Aget_space", "H5Aget_type"):
            getattr(L, f).restype = i64
            getattr(L, f).argtypes = [i64]
        L.H5Tget_class.argtypes = [i64]
        L.H5Tget_size.argtypes = [i64]
        L.H5Tget_size.restype = ctypes.c_size_t
        L.H5Sget_simple_extent_ndims.argtypes = [i64]
        L.H5Sget_simple_extent_dims.argtypes = [i64, ctypes.c_void_p, ctypes.c_void_p]
        L.H5Dread.argtypes = [i64, i64, i64, i64, i64, ctypes.c_void_p]
        L.H5Aread.argtypes = [i64, i64, ctypes.c_void_p]
        for f in ("H5Dclose", "H5Sclose", "H5Tclose", "H5Fclose", "H5Gclose", "H5Aclose"):
            getattr(L, f).argtypes = [i64]
        L.H5Gget_info.argtypes = [i64, ctypes.c_void_p]
        self.f = L.H5Fopen(path.encode(), 0, 0)
        assert self.f >= 0, path
        self.nat_int = i64.in_dll(L, "H5T_NATIVE_INT_g").value
        self.nat_float = i64.in_dll(L, "H5T_NATIVE_FLOAT_g").value

    def close(self):
        self.L.H5Fclose(self.f)

    def _shape_type(self, sp, ty):
        L = self.L
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (ctypes.c_ulonglong * max(nd, 1))()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        return tuple(dims[:nd]), L.H5Tget_class(ty), L.H5Tget_size(ty)

    def dataset(self, name):
        """(class, element size, shape, values) of a dataset; values as int32 or float32."""
        L = self.L
        d = L.H5Dopen2(self.f, name.encode(), 0)
        assert d >= 0, name
        sp, ty = L.H5Dget_space(d), L.H5Dget_type(d)
        shape, cls, size = self._shape_type(sp, ty)
        out = np.zeros(shape, np.int32 if cls == H5T_INTEGER else np.float32)
        assert L.H5Dread(d, self.nat_int if cls == H5T_INTEGER else self.nat_float, 0, 0, 0,
                         out.ctypes.data) >= 0
        L.H5Tclose(ty)
        L.H5Sclose(sp)
        L.H5Dclose(d)
        return cls, size, shape, out

    def attribute_int(self, obj, name):
        L = self.L
        a = L.H5Aopen(self.f if obj == "/" else L.H5Gopen2(self.f, obj.encode(), 0), name.encode(), 0)
        assert a >= 0, name
        v = ctypes.c_int()
        assert L.H5Aread(a, self.nat_int, ctypes.byref(v)) >= 0
        L.H5Aclose(a)
        return v.value

    def n_links(self, group):
        L = self.L
        g = L.H5Gopen2(self.f, group.encode(), 0)
        assert g >= 0, group
        info = (ctypes.c_int64 * 4)()  # H5G_info_t: storage_type, nlinks, max_corder, mounted
        assert L.H5Gget_info(g, info) >= 0
        L.H5Gclose(g)
        return ctypes.cast(info, ctypes.POINTER(ctypes.c_ulonglong))[1]
