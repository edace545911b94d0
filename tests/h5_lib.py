"""Minimal ctypes binding of the HDF5 1.10 C library, for the tests of the filter plugin
(falcons_amd/h5z). It drives the plugin the way an application does: the plugin directory is
prepended to HDF5's search path and the library loads the plugin when a dataset names filter 40024."""
import ctypes as C
import os

import numpy as np

HDF5_LIB = os.environ.get("FALCONS_TEST_HDF5_LIB", "/opt/conda/lib/libhdf5.so")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGIN_DIR = os.path.join(ROOT, "falcons_amd", "h5z")
PLUGIN_SO = os.path.join(PLUGIN_DIR, "libh5z_falcons_mi355x.so")
FILTER_ID = 40024

hid_t = C.c_int64
herr_t = C.c_int
hsize_t = C.c_uint64


def available():
    return os.path.exists(HDF5_LIB) and os.path.exists(PLUGIN_SO)


class H5:
    def __init__(self):
        L = C.CDLL(HDF5_LIB)
        self.L = L
        L.H5open.restype = herr_t
        if L.H5open() < 0:
            raise RuntimeError("H5open failed")
        L.H5PLprepend.argtypes = [C.c_char_p]
        L.H5PLprepend.restype = herr_t
        if L.H5PLprepend(PLUGIN_DIR.encode()) < 0:
            raise RuntimeError("H5PLprepend failed")
        L.H5Eset_auto2.argtypes = [hid_t, C.c_void_p, C.c_void_p]
        L.H5Eset_auto2(0, None, None)  # failures are checked through return codes
        L.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, hid_t, hid_t]
        L.H5Fcreate.restype = hid_t
        L.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid_t]
        L.H5Fopen.restype = hid_t
        L.H5Fclose.argtypes = [hid_t]
        L.H5Screate_simple.argtypes = [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]
        L.H5Screate_simple.restype = hid_t
        L.H5Sclose.argtypes = [hid_t]
        L.H5Pcreate.argtypes = [hid_t]
        L.H5Pcreate.restype = hid_t
        L.H5Pclose.argtypes = [hid_t]
        L.H5Pset_chunk.argtypes = [hid_t, C.c_int, C.POINTER(hsize_t)]
        L.H5Pset_chunk.restype = herr_t
        L.H5Pset_filter.argtypes = [hid_t, C.c_int, C.c_uint, C.c_size_t, C.POINTER(C.c_uint)]
        L.H5Pset_filter.restype = herr_t
        L.H5Pget_filter_by_id2.argtypes = [hid_t, C.c_int, C.POINTER(C.c_uint), C.POINTER(C.c_size_t),
                                           C.POINTER(C.c_uint), C.c_size_t, C.c_char_p, C.POINTER(C.c_uint)]
        L.H5Pget_filter_by_id2.restype = herr_t
        L.H5Dcreate2.argtypes = [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]
        L.H5Dcreate2.restype = hid_t
        L.H5Dopen2.argtypes = [hid_t, C.c_char_p, hid_t]
        L.H5Dopen2.restype = hid_t
        L.H5Dclose.argtypes = [hid_t]
        L.H5Dclose.restype = herr_t
        L.H5Fclose.restype = herr_t
        L.H5Dwrite.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]
        L.H5Dwrite.restype = herr_t
        L.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]
        L.H5Dread.restype = herr_t
        L.H5Dget_create_plist.argtypes = [hid_t]
        L.H5Dget_create_plist.restype = hid_t
        L.H5Dget_chunk_storage_size.argtypes = [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]
        L.H5Dget_chunk_storage_size.restype = herr_t
        L.H5Dread_chunk.argtypes = [hid_t, hid_t, C.POINTER(hsize_t), C.POINTER(C.c_uint32), C.c_void_p]
        L.H5Dread_chunk.restype = herr_t
        L.H5Zfilter_avail.argtypes = [C.c_int]
        L.H5Zfilter_avail.restype = C.c_int
        g = lambda name: hid_t.in_dll(L, name).value  # noqa: E731
        self.DCPL = g("H5P_CLS_DATASET_CREATE_ID_g")
        self.TYPES = {np.dtype(np.float32): g("H5T_NATIVE_FLOAT_g"), np.dtype(np.float64): g("H5T_NATIVE_DOUBLE_g"),
                      np.dtype(np.int16): g("H5T_NATIVE_SHORT_g"), np.dtype(np.uint8): g("H5T_NATIVE_UCHAR_g"),
                      np.dtype(np.int32): g("H5T_NATIVE_INT_g"), np.dtype(np.uint16): g("H5T_NATIVE_USHORT_g"),
                      np.dtype(np.int64): g("H5T_NATIVE_LLONG_g"), np.dtype(np.uint64): g("H5T_NATIVE_ULLONG_g"),
                      np.dtype(np.int8): g("H5T_NATIVE_SCHAR_g"), np.dtype(np.uint32): g("H5T_NATIVE_UINT_g")}

    @staticmethod
    def _dims(shape):
        return (hsize_t * len(shape))(*shape)

    def write(self, path, name, data, chunks, cd=None, check=True):
        """Creates `path` with one chunked dataset through filter 40024 (cd_values `cd`, a list of
        unsigned ints, or none) and writes `data`. Returns the lowest status of the write and of the
        closes: chunks pass the filter when HDF5's chunk cache flushes them, at the latest on
        H5Dclose."""
        L = self.L
        f = L.H5Fcreate(path.encode(), 2, 0, 0)  # H5F_ACC_TRUNC
        assert f >= 0
        sp = L.H5Screate_simple(data.ndim, self._dims(data.shape), None)
        dcpl = L.H5Pcreate(self.DCPL)
        assert L.H5Pset_chunk(dcpl, len(chunks), self._dims(chunks)) >= 0
        cdv = (C.c_uint * len(cd))(*cd) if cd else None
        assert L.H5Pset_filter(dcpl, FILTER_ID, 0, len(cd) if cd else 0, cdv) >= 0  # H5Z_FLAG_MANDATORY
        ty = self.TYPES[data.dtype]
        d = L.H5Dcreate2(f, name.encode(), ty, sp, 0, dcpl, 0)
        assert d >= 0, "H5Dcreate2 failed"
        rc = 0
        if data.size:
            buf = np.ascontiguousarray(data)
            rc = L.H5Dwrite(d, ty, 0, 0, 0, buf.ctypes.data)
        rc = min(rc, L.H5Dclose(d))
        L.H5Pclose(dcpl)
        L.H5Sclose(sp)
        rc = min(rc, L.H5Fclose(f))
        if check:
            assert rc >= 0, "writing the dataset failed"
        return rc

    def create_only(self, path, name, shape, dtype, chunks, cd=None):
        """Creates the dataset without writing it (runs the filter's set_local only)."""
        return self.write(path, name, _Empty(shape, dtype), chunks, cd)

    def read(self, path, name, shape, dtype):
        L = self.L
        f = L.H5Fopen(path.encode(), 0, 0)
        d = L.H5Dopen2(f, name.encode(), 0)
        out = np.empty(shape, dtype)
        rc = L.H5Dread(d, self.TYPES[np.dtype(dtype)], 0, 0, 0, out.ctypes.data)
        L.H5Dclose(d)
        L.H5Fclose(f)
        if rc < 0:
            raise RuntimeError("H5Dread failed")
        return out

    def cd_values(self, path, name):
        """The filter's cd_values as stored in the dataset's creation property list."""
        L = self.L
        f = L.H5Fopen(path.encode(), 0, 0)
        d = L.H5Dopen2(f, name.encode(), 0)
        pl = L.H5Dget_create_plist(d)
        n = C.c_size_t(64)
        cd = (C.c_uint * 64)()
        fl = C.c_uint(0)
        rc = L.H5Pget_filter_by_id2(pl, FILTER_ID, C.byref(fl), C.byref(n), cd, 0, None, None)
        L.H5Pclose(pl)
        L.H5Dclose(d)
        L.H5Fclose(f)
        assert rc >= 0
        return [cd[i] for i in range(n.value)]

    def raw_chunk(self, path, name, offset):
        """The stored (filtered) bytes of the chunk at element offset `offset`."""
        L = self.L
        f = L.H5Fopen(path.encode(), 0, 0)
        d = L.H5Dopen2(f, name.encode(), 0)
        off = self._dims(offset)
        sz = hsize_t(0)
        assert L.H5Dget_chunk_storage_size(d, off, C.byref(sz)) >= 0
        buf = np.empty(sz.value, np.uint8)
        mask = C.c_uint32(0)
        rc = L.H5Dread_chunk(d, 0, off, C.byref(mask), buf.ctypes.data)
        L.H5Dclose(d)
        L.H5Fclose(f)
        assert rc >= 0
        return buf.tobytes(), mask.value


class _Empty:
    """Shape and dtype of a dataset that is created and never written."""

    def __init__(self, shape, dtype):
        self.shape, self.dtype, self.ndim, self.size = tuple(shape), np.dtype(dtype), len(shape), 0


def cd_from_bytes(b):
    """Config::save bytes -> cd_values (zero-padded unsigned ints), as H5Z_Falcons.cpp:27-36 stores them."""
    pad = (-len(b)) % 4
    return [int(v) for v in np.frombuffer(bytes(b) + b"\0" * pad, dtype="<u4")]


def cd_to_bytes(cd):
    return np.asarray(cd, dtype="<u4").tobytes()
