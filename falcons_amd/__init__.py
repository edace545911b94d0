"""MI355X implementation of the Falcons interpolation compression path.

The Python surface mirrors pyfalcons (tools/pyfalcons/src/pyfalcons/falcons.pyx,
pyConfig.pyx, pyConfigEnums.py): ``falconsConfig``, ``falcons.compress``,
``falcons.decompress``, ``falcons.verify`` and the enum classes, with the same argument meaning and
error behaviour (ValueError where the C++ API raises std::invalid_argument). Every call goes
through the C-ABI of ``libfalcons_mi355x.so`` (declared in include/falcons_mi355x.h), whose
kernels run on the GPU. There is no CPU fallback: when the library is not built, importing this
package fails.

Beyond pyfalcons, arrays already on the GPU (torch tensors on a ROCm device) are compressed in
place on the device, and ``Context.compress_ex`` exposes the pre-zstd buffer and device timings.
"""
import ctypes as C
import os
from enum import IntEnum

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfalcons_mi355x.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `make -C falcons_amd/csrc` "
                      "(or __graft_entry__.build())")

# Independent calls (the batched and serving forms, several contexts) overlap on the device only
# as far as the runtime gives their streams hardware queues; its default is 4. Raised to 16 (unless
# FALCONS_KEEP_HW_QUEUES is set); the setting takes effect when this process initialises the GPU
# runtime, so it is made before the library loads.
if not os.environ.get("FALCONS_KEEP_HW_QUEUES"):
    try:
        _hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    except ValueError:
        _hwq = 4
    if _hwq < 16:
        os.environ["GPU_MAX_HW_QUEUES"] = "16"

_L = C.CDLL(LIB_PATH)

FZ_OK, FZ_ERR_INVALID_ARGUMENT, FZ_ERR_UNSUPPORTED, FZ_ERR_GPU, FZ_ERR_RUNTIME = range(5)


class fz_config(C.Structure):
    _fields_ = [
        ("N", C.c_int32), ("dims", C.c_uint64 * 4), ("num", C.c_uint64),
        ("cmpr_algo", C.c_int32), ("eb_mode", C.c_int32),
        ("abs_eb", C.c_double), ("rel_eb", C.c_double), ("psnr_eb", C.c_double), ("l2norm_eb", C.c_double),
        ("lorenzo", C.c_int32), ("lorenzo2", C.c_int32), ("regression", C.c_int32), ("regression2", C.c_int32),
        ("openmp", C.c_int32), ("quantbin_cnt", C.c_int32), ("block_size", C.c_int32),
        ("pred_dim", C.c_int32), ("data_type", C.c_int32),
        ("interp_algo", C.c_int32), ("interp_direction", C.c_int32), ("interp_anchor_stride", C.c_int32),
        ("interp_alpha", C.c_double), ("interp_beta", C.c_double),
    ]


class fz_opts(C.Structure):
    _fields_ = [
        ("simd", C.c_int32), ("copy_back", C.c_int32),
        ("prezstd", C.POINTER(C.c_uint8)), ("prezstd_cap", C.c_size_t), ("prezstd_len", C.c_size_t),
        ("t_device_ms", C.c_double), ("t_zstd_ms", C.c_double), ("t_total_ms", C.c_double),
        ("ncodes", C.c_uint64), ("nunpred", C.c_uint64), ("prezstd_bytes", C.c_uint64),
        ("method", C.c_int32), ("use_delta", C.c_int32),
    ]


_u8p = C.POINTER(C.c_uint8)
_L.fz_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
_L.fz_ctx_create.restype = C.c_int
_L.fz_ctx_destroy.argtypes = [C.c_void_p]
_L.fz_last_error.argtypes = [C.c_void_p]
_L.fz_last_error.restype = C.c_char_p
_L.fz_opts_init.argtypes = [C.POINTER(fz_opts)]
_L.fz_config_init.argtypes = [C.POINTER(fz_config), C.c_int, C.POINTER(C.c_uint64)]
_L.fz_config_size_est.argtypes = [C.POINTER(fz_config)]
_L.fz_config_size_est.restype = C.c_size_t
_L.fz_config_save.argtypes = [C.POINTER(fz_config), _u8p]
_L.fz_config_save.restype = C.c_size_t
_L.fz_read_config.argtypes = [_u8p, C.c_size_t, C.POINTER(fz_config), C.c_char_p, C.c_size_t]
_L.fz_read_config.restype = C.c_int
_L.fz_check_prezstd.argtypes = [C.POINTER(fz_config), C.c_int, _u8p, C.c_size_t, C.c_char_p, C.c_size_t]
_L.fz_check_prezstd.restype = C.c_int
_L.fz_config_set_dims.argtypes = [C.POINTER(fz_config), C.c_int, C.POINTER(C.c_uint64)]
_L.fz_config_set_dims.restype = C.c_int
_L.fz_config_load.argtypes = [_u8p, C.c_size_t, C.POINTER(fz_config), C.c_char_p, C.c_size_t]
_L.fz_config_load.restype = C.c_int
_L.fz_compress_size_bound.argtypes = [C.POINTER(fz_config), C.c_int]
_L.fz_compress_size_bound.restype = C.c_size_t
_L.fz_compress.argtypes = [C.c_void_p, C.POINTER(fz_config), C.c_int, C.c_void_p, C.c_int, _u8p, C.c_size_t,
                           C.POINTER(C.c_size_t), C.POINTER(fz_opts)]
_L.fz_compress.restype = C.c_int
_L.fz_decompress.argtypes = [C.c_void_p, _u8p, C.c_size_t, C.c_int, C.c_void_p, C.c_size_t, C.c_int,
                             C.POINTER(fz_config), C.POINTER(fz_opts)]
_L.fz_decompress.restype = C.c_int
_L.fz_prezstd_bound.argtypes = [C.POINTER(fz_config), C.c_int]
_L.fz_prezstd_bound.restype = C.c_size_t
_L.fz_compress_device.argtypes = [C.c_void_p, C.POINTER(fz_config), C.c_int, C.c_void_p, C.c_void_p, C.c_size_t,
                                  C.POINTER(C.c_size_t), C.POINTER(fz_opts)]
_L.fz_compress_device.restype = C.c_int
_L.fz_compress_device_noalloc.argtypes = _L.fz_compress_device.argtypes
_L.fz_compress_device_noalloc.restype = C.c_int
_L.fz_decompress_device.argtypes = [C.c_void_p, C.POINTER(fz_config), C.c_int, _u8p, C.c_void_p, C.c_size_t,
                                    C.c_void_p, C.c_size_t, C.POINTER(fz_opts)]
_L.fz_decompress_device.restype = C.c_int
_L.fz_device_alloc.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]
_L.fz_device_alloc.restype = C.c_int
_L.fz_device_free.argtypes = [C.c_void_p, C.c_void_p]
_L.fz_device_free.restype = C.c_int
_L.fz_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_L.fz_memcpy.restype = C.c_int
_L.fz_synchronize.argtypes = [C.c_void_p]
_L.fz_synchronize.restype = C.c_int
_L.fz_ctx_stream.argtypes = [C.c_void_p]
_L.fz_ctx_stream.restype = C.c_void_p
_L.fz_zstd_version.argtypes = []
_L.fz_zstd_version.restype = C.c_uint
_L.fz_profile_enable.argtypes = [C.c_void_p, C.c_int]
_L.fz_profile_enable.restype = C.c_int
_L.fz_profile_reset.argtypes = [C.c_void_p]
_L.fz_profile_reset.restype = C.c_int
_L.fz_profile_count.argtypes = [C.c_void_p]
_L.fz_profile_count.restype = C.c_int
_L.fz_profile_entry.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_uint64),
                                C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
_L.fz_profile_entry.restype = C.c_int


class fz_batch_item(C.Structure):
    _fields_ = [("conf", fz_config), ("dtype", C.c_int32), ("status", C.c_int32), ("d_data", C.c_void_p),
                ("d_pre", C.c_void_p), ("pre_cap", C.c_size_t), ("pre_len", C.c_size_t)]


class fz_batch_dec_item(C.Structure):
    _fields_ = [("conf", fz_config), ("dtype", C.c_int32), ("status", C.c_int32), ("h_pre", C.c_void_p),
                ("d_pre", C.c_void_p), ("pre_len", C.c_size_t), ("d_out", C.c_void_p),
                ("out_cap_elems", C.c_size_t)]


class fz_request(C.Structure):
    _fields_ = [("op", C.c_int32), ("dtype", C.c_int32), ("conf", fz_config), ("src", C.c_void_p),
                ("src_len", C.c_size_t), ("dst", C.c_void_p), ("dst_cap", C.c_size_t), ("tag", C.c_uint64)]


class fz_result(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("op", C.c_int32), ("status", C.c_int32), ("out_len", C.c_size_t),
                ("field_bytes", C.c_size_t), ("t_submit_ms", C.c_double), ("t_start_ms", C.c_double),
                ("t_done_ms", C.c_double), ("t_device_ms", C.c_double), ("t_zstd_ms", C.c_double),
                ("msg", C.c_char * 160)]


class fz_server_stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("submitted", "completed", "failed", "in_flight", "max_in_flight",
                                          "pending", "max_pending", "held_bytes", "max_held_bytes",
                                          "budget_bytes")]


_L.fz_tune_report.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
_L.fz_tune_report.restype = C.c_int
_L.fz_set_workers.argtypes = [C.c_void_p, C.c_int]
_L.fz_set_workers.restype = C.c_int
_L.fz_compress_batch.argtypes = [C.c_void_p, C.POINTER(fz_batch_item), C.c_int, C.POINTER(fz_opts)]
_L.fz_compress_batch.restype = C.c_int
_L.fz_decompress_batch.argtypes = [C.c_void_p, C.POINTER(fz_batch_dec_item), C.c_int, C.POINTER(fz_opts)]
_L.fz_decompress_batch.restype = C.c_int
_L.fz_server_create.argtypes = [C.c_int, C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]
_L.fz_server_create.restype = C.c_int
_L.fz_server_submit.argtypes = [C.c_void_p, C.POINTER(fz_request)]
_L.fz_server_submit.restype = C.c_int
_L.fz_server_poll.argtypes = [C.c_void_p, C.POINTER(fz_result), C.c_int]
_L.fz_server_poll.restype = C.c_int
_L.fz_server_get_stats.argtypes = [C.c_void_p, C.POINTER(fz_server_stats)]
_L.fz_server_get_stats.restype = C.c_int
_L.fz_server_destroy.argtypes = [C.c_void_p]

EXPORTED_SYMBOLS = ("fz_ctx_create", "fz_ctx_destroy", "fz_last_error", "fz_opts_init", "fz_config_init",
                    "fz_config_size_est", "fz_config_save", "fz_read_config", "fz_config_set_dims",
                    "fz_config_load", "fz_compress_size_bound", "fz_config_on_gpu",
                    "fz_compress", "fz_decompress", "fz_prezstd_bound", "fz_compress_device",
                    "fz_compress_device_noalloc",
                    "fz_decompress_device", "fz_device_alloc", "fz_device_free", "fz_memcpy", "fz_synchronize",
                    "fz_ctx_stream",
                    "fz_zstd_version", "fz_profile_enable", "fz_profile_reset", "fz_profile_count",
                    "fz_profile_entry", "fz_set_workers", "fz_compress_batch", "fz_decompress_batch",
                    "fz_server_create", "fz_server_submit", "fz_server_poll", "fz_server_get_stats",
                    "fz_server_destroy", "fz_tune_report", "fz_check_prezstd")


def zstd_version():
    """ZSTD_versionNumber() of the libzstd the host stage uses (0 if none loads)."""
    return int(_L.fz_zstd_version())


class falconsErrorBoundMode(IntEnum):
    ABS = 0
    REL = 1
    PSNR = 2
    L2NORM = 3
    ABS_AND_REL = 4
    ABS_OR_REL = 5


class falconsAlgorithm(IntEnum):
    LORENZO_REG = 0
    INTERP_LORENZO = 1
    INTERP = 2
    NOPRED = 3
    LOSSLESS = 4


class falconsInterpAlgorithm(IntEnum):
    LINEAR = 0
    CUBIC = 1


class FalconsError(RuntimeError):
    pass


class FalconsUnsupported(FalconsError):
    pass


def _raise(rc, msg):
    if rc == FZ_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if rc == FZ_ERR_UNSUPPORTED:
        raise FalconsUnsupported(msg)
    raise FalconsError(msg)


class falconsConfig:
    """Falcons::Config (include/Falcons/utils/Config.hpp) as seen from Python."""

    def __init__(self, *args):
        self.conf = fz_config()
        self.setDims(*args if args else (1,))

    def setDims(self, *args):
        if not args:
            raise ValueError("At least one dimension required")
        for a in args:
            if int(a) != a or int(a) < 1:
                raise ValueError(f"Dimension must be positive integer, got {a}")
        keep = {k: getattr(self.conf, k) for k, _ in fz_config._fields_ if k not in ("N", "dims", "num",
                                                                                   "block_size", "pred_dim")}
        fresh = self.conf.N == 0
        dims = (C.c_uint64 * 4)(*[int(a) for a in args][:4])
        _L.fz_config_init(C.byref(self.conf), len(args), dims)
        if not fresh:
            for k, v in keep.items():
                setattr(self.conf, k, v)
        return self.num_elements

    def copy(self):
        other = falconsConfig.__new__(falconsConfig)
        other.conf = fz_config()
        C.pointer(other.conf)[0] = self.conf
        return other

    @property
    def dims(self):
        return tuple(int(self.conf.dims[i]) for i in range(self.conf.N))

    @property
    def num_elements(self):
        return int(self.conf.num)

    @property
    def ndim(self):
        return int(self.conf.N)

    def _prop(name, cast=float):
        def g(self):
            return cast(getattr(self.conf, name))

        def s(self, v):
            setattr(self.conf, name, int(v) if cast is not float else float(v))

        return property(g, s)

    absErrorBound = _prop("abs_eb")
    relErrorBound = _prop("rel_eb")
    psnrErrorBound = _prop("psnr_eb")
    l2normErrorBound = _prop("l2norm_eb")
    errorBoundMode = _prop("eb_mode", int)
    cmprAlgo = _prop("cmpr_algo", int)
    interpAlgo = _prop("interp_algo", int)
    interpDirection = _prop("interp_direction", int)
    interpAnchorStride = _prop("interp_anchor_stride", int)
    interpAlpha = _prop("interp_alpha")
    interpBeta = _prop("interp_beta")
    lorenzo = _prop("lorenzo", bool)
    lorenzo2 = _prop("lorenzo2", bool)
    regression = _prop("regression", bool)
    regression2 = _prop("regression2", bool)
    openmp = _prop("openmp", bool)
    quantbinCnt = _prop("quantbin_cnt", int)
    blockSize = _prop("block_size", int)
    dataType = _prop("data_type", int)
    del _prop

    def size_est(self):
        return int(_L.fz_config_size_est(C.byref(self.conf)))

    def save(self):
        buf = (C.c_uint8 * 256)()
        n = _L.fz_config_save(C.byref(self.conf), buf)
        return bytes(buf[:n])


def check_prezstd(pre, config, dtype):
    """fz_check_prezstd: raises ValueError unless the pre-zstd buffer is well formed for config (the
    stream's configuration, e.g. from read_config) and dtype. Host only; needs no GPU."""
    b = np.ascontiguousarray(np.frombuffer(bytes(pre), dtype=np.uint8) if not isinstance(pre, np.ndarray) else pre,
                             dtype=np.uint8)
    err = C.create_string_buffer(256)
    rc = _L.fz_check_prezstd(C.byref(config.conf), _DT[np.dtype(dtype)], b.ctypes.data_as(_u8p), b.size, err, 256)
    if rc:
        _raise(rc, err.value.decode())


def read_config(stream):
    """Falcons_read_config: parse and validate a stream header (ValueError when invalid)."""
    b = np.frombuffer(bytes(stream), dtype=np.uint8) if not isinstance(stream, np.ndarray) else stream
    b = np.ascontiguousarray(b, dtype=np.uint8)
    cfg = falconsConfig(1)
    err = C.create_string_buffer(256)
    rc = _L.fz_read_config(b.ctypes.data_as(_u8p), b.size, C.byref(cfg.conf), err, 256)
    if rc:
        raise ValueError(err.value.decode())
    return cfg


# element types, numbered as Falcons_FLOAT .. Falcons_INT64 (Config.hpp:29-38)
_DT = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.uint8): 2, np.dtype(np.int8): 3,
       np.dtype(np.uint16): 4, np.dtype(np.int16): 5, np.dtype(np.uint32): 6, np.dtype(np.int32): 7,
       np.dtype(np.uint64): 8, np.dtype(np.int64): 9}
_DT_OF = {v: k for k, v in _DT.items()}


def _is_device_tensor(x):
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


_DL_CPU, _DL_CUDA, _DL_ROCM = 1, 2, 10  # DLDeviceType (dlpack.h)
_DL_INT, _DL_UINT, _DL_FLOAT = 0, 1, 2  # DLDataTypeCode kDLInt, kDLUInt, kDLFloat


class _DLTensor(C.Structure):  # DLTensor (dlpack.h)
    _fields_ = [("data", C.c_void_p), ("device_type", C.c_int32), ("device_id", C.c_int32), ("ndim", C.c_int32),
                ("code", C.c_uint8), ("bits", C.c_uint8), ("lanes", C.c_uint16),
                ("shape", C.POINTER(C.c_int64)), ("strides", C.POINTER(C.c_int64)), ("byte_offset", C.c_uint64)]


class _DLManagedTensor(C.Structure):
    _fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", C.c_void_p), ("deleter", C.c_void_p)]


_DL_DELETER = C.CFUNCTYPE(None, C.c_void_p)
_capi = C.pythonapi
_capi.PyCapsule_GetPointer.restype = C.c_void_p
_capi.PyCapsule_GetPointer.argtypes = [C.py_object, C.c_char_p]
_capi.PyCapsule_SetName.restype = C.c_int
_capi.PyCapsule_SetName.argtypes = [C.py_object, C.c_char_p]
_capi.PyCapsule_IsValid.restype = C.c_int
_capi.PyCapsule_IsValid.argtypes = [C.py_object, C.c_char_p]
_capi.PyCapsule_New.restype = C.py_object
_capi.PyCapsule_New.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]


class DLDeviceArray:
    """A GPU array taken from a DLPack capsule: its device pointer, shape and dtype, with no copy and
    no other GPU runtime involved. Owns the capsule (consumed, as the protocol asks) and calls the
    producer's deleter on release()."""

    def __init__(self, producer, stream=None):
        # stream: the consuming stream (an integer handle); the producer orders its pending writes
        # before it. None leaves the producer's default (the legacy default stream).
        cap = producer.__dlpack__() if stream is None else producer.__dlpack__(stream=stream)
        if not _capi.PyCapsule_IsValid(cap, b"dltensor"):
            raise ValueError("not an unconsumed DLPack capsule")
        mt = C.cast(_capi.PyCapsule_GetPointer(cap, b"dltensor"), C.POINTER(_DLManagedTensor))
        _capi.PyCapsule_SetName(cap, b"used_dltensor")
        self._cap, self._mt = cap, mt
        t = mt.contents.dl_tensor
        self.device_type, self.device = int(t.device_type), int(t.device_id)
        self.shape = tuple(int(t.shape[i]) for i in range(t.ndim))
        self.size = int(np.prod(self.shape)) if self.shape else 1
        kinds = {_DL_FLOAT: "f", _DL_INT: "i", _DL_UINT: "u"}
        if t.code not in kinds or t.lanes != 1 or np.dtype(f"{kinds[t.code]}{t.bits // 8}") not in _DT:
            self.release()
            raise TypeError("Unsupported dtype: float32/64 and 8- to 64-bit integer arrays are compressed")
        self.dtype = np.dtype(f"{kinds[t.code]}{t.bits // 8}")
        if t.strides:
            want, st = 1, []
            for d in reversed(self.shape):
                st.append(want)
                want *= d
            if any(int(t.strides[i]) != st[len(st) - 1 - i] for i in range(t.ndim) if self.shape[i] > 1):
                self.release()
                raise ValueError("DLPack array must be C-contiguous")
        self.ptr = int(t.data or 0) + int(t.byte_offset)
        self.nbytes = self.size * self.dtype.itemsize

    def release(self):
        mt, self._mt = getattr(self, "_mt", None), None
        if mt and mt.contents.deleter:
            _DL_DELETER(mt.contents.deleter)(C.cast(mt, C.c_void_p))

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def from_dlpack(x, stream=None):
    """Any array that speaks DLPack (__dlpack__ / __dlpack_device__), without a copy: a GPU array
    becomes a DLDeviceArray (device pointer, shape, dtype), a host array a numpy array. numpy
    arrays and torch tensors pass through unchanged. `stream` is the stream that will read the GPU
    array (the library's calls pass their context's stream)."""
    if isinstance(x, (np.ndarray, DLDeviceArray)) or type(x).__module__.startswith("torch"):
        return x
    if not (hasattr(x, "__dlpack__") and hasattr(x, "__dlpack_device__")):
        return x
    dev_type, _ = x.__dlpack_device__()
    if int(dev_type) in (_DL_CUDA, _DL_ROCM):
        return DLDeviceArray(x, stream)
    if int(dev_type) == _DL_CPU:
        return np.from_dlpack(x)
    raise TypeError(f"unsupported DLPack device type {int(dev_type)}")


def _device_view(x):
    """(pointer, element count, dtype code) of a device array, or None for host arrays."""
    if isinstance(x, DLDeviceArray):
        return x.ptr, x.size, _DT[x.dtype]
    if _is_device_tensor(x):
        import torch
        dt = {torch.float32: 0, torch.float64: 1}.get(x.dtype)
        if dt is None or not x.is_contiguous():
            raise TypeError("device input must be a contiguous float32/float64 tensor")
        return x.data_ptr(), x.numel(), dt
    return None


# DLManagedTensors handed out by DeviceArray.__dlpack__, by address, with what they point into;
# the consumer's call of the deleter removes an entry (a capsule that is never consumed keeps its
# few hundred bytes until the process ends).
_DL_EXPORTS = {}


def _dl_export_deleter(addr):
    _DL_EXPORTS.pop(int(addr or 0), None)


_DL_EXPORT_DELETER = _DL_DELETER(_dl_export_deleter)


class DeviceArray:
    """A typed view of a DeviceBuffer that other libraries can take through DLPack (ROCm device,
    no copy). The buffer must not be freed while a consumer uses it."""

    def __init__(self, buf, shape, dtype):
        dtype = np.dtype(dtype)
        if dtype not in _DT:
            raise TypeError(f"Unsupported dtype: {dtype}")
        n = int(np.prod(shape))
        if n * dtype.itemsize > buf.nbytes:
            raise ValueError("shape larger than the device buffer")
        self.buf, self.shape, self.dtype = buf, tuple(int(s) for s in shape), dtype

    def __dlpack_device__(self):
        return (_DL_ROCM, int(self.buf.ctx.device))

    def __dlpack__(self, stream=None, **kw):
        # the library's work on this buffer is finished before any consumer stream reads it
        rc = _L.fz_synchronize(self.buf.ctx.h) if getattr(self.buf.ctx, "h", None) else 0
        if rc:
            _raise(rc, self.buf.ctx._err())
        shp = (C.c_int64 * max(1, len(self.shape)))(*self.shape)
        mt = _DLManagedTensor()
        t = mt.dl_tensor
        t.data = self.buf.ptr.value if isinstance(self.buf.ptr, C.c_void_p) else self.buf.ptr
        t.device_type, t.device_id, t.ndim = _DL_ROCM, int(self.buf.ctx.device), len(self.shape)
        t.code = {"f": _DL_FLOAT, "i": _DL_INT, "u": _DL_UINT}[self.dtype.kind]
        t.bits, t.lanes = 8 * self.dtype.itemsize, 1
        t.shape, t.strides, t.byte_offset = C.cast(shp, C.POINTER(C.c_int64)), None, 0
        mt.deleter = C.cast(_DL_EXPORT_DELETER, C.c_void_p)
        _DL_EXPORTS[C.addressof(mt)] = (mt, shp, self.buf)
        return _capi.PyCapsule_New(C.addressof(mt), b"dltensor", None)


class DeviceBuffer:
    """Device memory owned by a Context (freed with it or by free())."""

    def __init__(self, ctx, nbytes):
        p = C.c_void_p()
        rc = _L.fz_device_alloc(ctx.h, int(nbytes), C.byref(p))
        if rc:
            _raise(rc, ctx._err())
        self.ctx, self.ptr, self.nbytes = ctx, p, int(nbytes)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        rc = _L.fz_memcpy(self.ctx.h, self.ptr, C.c_void_p(a.ctypes.data), a.nbytes, 0)
        if rc:
            _raise(rc, self.ctx._err())

    def download(self, nbytes=None, dtype=np.uint8):
        n = self.nbytes if nbytes is None else int(nbytes)
        out = np.empty(n // np.dtype(dtype).itemsize, dtype=dtype)
        rc = _L.fz_memcpy(self.ctx.h, C.c_void_p(out.ctypes.data), self.ptr, out.nbytes, 1)
        if rc:
            _raise(rc, self.ctx._err())
        return out

    def as_array(self, shape, dtype):
        """This buffer as a typed array other libraries can take through DLPack."""
        return DeviceArray(self, shape, dtype)

    def free(self):
        if self.ptr and self.ctx.h:
            _L.fz_device_free(self.ctx.h, self.ptr)
        self.ptr = None


class Context:
    """One GPU: stream, device workspace. Methods are not re-entrant on one context."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = _L.fz_ctx_create(int(device), C.byref(h))
        if rc:
            raise FalconsError(f"cannot create a GPU context on device {device}")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            _L.fz_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self):
        return _L.fz_last_error(self.h).decode()

    def stream_handle(self):
        """The context's HIP stream as an integer (the `stream` a DLPack producer is given)."""
        return int(_L.fz_ctx_stream(self.h) or 0)

    def compress_ex(self, data, config, simd=1, want_prezstd=False, copy_back=True):
        """Compress a numpy array (host) or a torch GPU tensor (device, overwritten in place).
        Returns (stream bytes as uint8 ndarray, fz_opts with timings, pre-zstd bytes or None)."""
        data = from_dlpack(data, self.stream_handle())
        dv = _device_view(data)
        on_dev = dv is not None
        if on_dev:
            ptr, size, dt = dv
            nbytes = size * _DT_OF[dt].itemsize
        else:
            if data.dtype not in _DT:
                raise TypeError(f"Unsupported dtype: {data.dtype}")
            if not data.flags["C_CONTIGUOUS"] or not data.flags["WRITEABLE"]:
                raise ValueError("compression needs a C-contiguous writable array")
            dt, ptr, nbytes, size = _DT[data.dtype], data.ctypes.data, data.nbytes, data.size
        if size == 0:
            raise ValueError("Cannot compress an empty array")
        if config.num_elements != size:
            config.setDims(*data.shape)
        cap = int(_L.fz_compress_size_bound(C.byref(config.conf), dt))
        out = np.empty(cap, dtype=np.uint8)
        opts = fz_opts()
        _L.fz_opts_init(C.byref(opts))
        opts.simd = int(simd)
        opts.copy_back = 1 if copy_back else 0
        pre = None
        if want_prezstd:
            pre = np.empty(nbytes * 3 + (4 << 20), dtype=np.uint8)
            opts.prezstd = pre.ctypes.data_as(_u8p)
            opts.prezstd_cap = pre.size
        olen = C.c_size_t(0)
        rc = _L.fz_compress(self.h, C.byref(config.conf), dt, C.c_void_p(ptr), 1 if on_dev else 0,
                            out.ctypes.data_as(_u8p), cap, C.byref(olen), C.byref(opts))
        if rc:
            _raise(rc, self._err())
        if want_prezstd:
            pre = pre[: opts.prezstd_len].tobytes()
        return out[: olen.value], opts, pre

    def decompress_ex(self, stream, dtype, out=None):
        """Decompress into a new numpy array, or into `out` (numpy array or torch GPU tensor)."""
        b = np.ascontiguousarray(np.frombuffer(bytes(stream), dtype=np.uint8)
                                 if not isinstance(stream, np.ndarray) else stream, dtype=np.uint8)
        dt = _DT[np.dtype(dtype)]
        cfg = falconsConfig(1)
        opts = fz_opts()
        _L.fz_opts_init(C.byref(opts))
        if out is None:
            hdr = read_config(b)
            out = np.empty(hdr.num_elements, dtype=dtype)
        out = from_dlpack(out, self.stream_handle())
        dv = _device_view(out)
        on_dev = dv is not None
        ptr, cap = (dv[0], dv[1]) if on_dev else (out.ctypes.data, out.size)
        rc = _L.fz_decompress(self.h, b.ctypes.data_as(_u8p), b.size, dt, C.c_void_p(ptr), cap, 1 if on_dev else 0,
                              C.byref(cfg.conf), C.byref(opts))
        if rc:
            _raise(rc, self._err())
        return out, cfg, opts


    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def compress_device(self, config, dtype, d_data, d_pre, simd=1):
        """fz_compress_device: field in d_data (DeviceBuffer, overwritten with the reconstruction) -> pre-zstd
        buffer in d_pre. Returns (pre-zstd length, fz_opts). config is updated with the absolute bound."""
        dt = _DT[np.dtype(dtype)]
        opts = fz_opts()
        _L.fz_opts_init(C.byref(opts))
        opts.simd = int(simd)
        n = C.c_size_t(0)
        rc = _L.fz_compress_device(self.h, C.byref(config.conf), dt, d_data.ptr, d_pre.ptr, d_pre.nbytes, C.byref(n),
                                   C.byref(opts))
        if rc:
            _raise(rc, self._err())
        return n.value, opts

    def compress_device_noalloc(self, config, dtype, d_data, d_pre, simd=1):
        """fz_compress_device_noalloc: as compress_device, with the workspace in the library's own
        device array instead of allocated memory (fields up to 256 MiB of workspace)."""
        dt = _DT[np.dtype(dtype)]
        opts = fz_opts()
        _L.fz_opts_init(C.byref(opts))
        opts.simd = int(simd)
        n = C.c_size_t(0)
        rc = _L.fz_compress_device_noalloc(self.h, C.byref(config.conf), dt, d_data.ptr, d_pre.ptr, d_pre.nbytes,
                                           C.byref(n), C.byref(opts))
        if rc:
            _raise(rc, self._err())
        return n.value, opts

    def decompress_device(self, config, dtype, h_pre, d_pre, d_out, simd=1):
        """fz_decompress_device: pre-zstd buffer (host copy h_pre, device copy d_pre or None) -> field in d_out."""
        dt = _DT[np.dtype(dtype)]
        b = np.ascontiguousarray(np.frombuffer(bytes(h_pre), dtype=np.uint8) if not isinstance(h_pre, np.ndarray)
                                 else h_pre, dtype=np.uint8)
        opts = fz_opts()
        _L.fz_opts_init(C.byref(opts))
        opts.simd = int(simd)
        es = np.dtype(dtype).itemsize
        rc = _L.fz_decompress_device(self.h, C.byref(config.conf), dt, b.ctypes.data_as(_u8p),
                                     d_pre.ptr if d_pre is not None else None, b.size, d_out.ptr, d_out.nbytes // es,
                                     C.byref(opts))
        if rc:
            _raise(rc, self._err())
        return opts

    def prezstd_bound(self, config, dtype):
        return int(_L.fz_prezstd_bound(C.byref(config.conf), _DT[np.dtype(dtype)]))

    def profile(self, on=True):
        """Switch the per-launch HIP-event profiler on or off (see kernel_times)."""
        rc = _L.fz_profile_enable(self.h, 1 if on else 0)
        if rc:
            _raise(rc, self._err())

    def profile_reset(self):
        _L.fz_profile_reset(self.h)

    def kernel_times(self):
        """{kernel: (launches, total ms, algorithmic bytes)} since the last profile_reset."""
        out = {}
        name = C.create_string_buffer(128)
        ln, ms, by = C.c_uint64(), C.c_double(), C.c_uint64()
        for i in range(_L.fz_profile_count(self.h)):
            if _L.fz_profile_entry(self.h, i, name, 128, C.byref(ln), C.byref(ms), C.byref(by)) == 0:
                out[name.value.decode()] = (ln.value, ms.value, by.value)
        return out

    def synchronize(self):
        rc = _L.fz_synchronize(self.h)
        if rc:
            _raise(rc, self._err())

    def tune_report(self):
        """Ratios of the candidates of the last ALGO_INTERP_LORENZO parameter search on this context."""
        buf = (C.c_double * 16)()
        n = _L.fz_tune_report(self.h, buf, 16)
        return [float(buf[i]) for i in range(min(n, 16))]

    def set_workers(self, n):
        rc = _L.fz_set_workers(self.h, int(n))
        if rc:
            _raise(rc, self._err())

    def compress_batch(self, items, simd=1):
        """fz_compress_batch. items: sequence of (config, dtype, d_data, d_pre) with DeviceBuffers; every
        config is updated as compress_device updates it. Returns (pre-zstd lengths, statuses, wall ms);
        raises on the first failure after the whole batch ran."""
        arr = (fz_batch_item * len(items))()
        for it, (cfg, dtype, d_data, d_pre) in zip(arr, items):
            it.conf = cfg.conf
            it.dtype = _DT[np.dtype(dtype)]
            it.d_data, it.d_pre, it.pre_cap = d_data.ptr, d_pre.ptr, d_pre.nbytes
        opts = fz_opts()
        _L.fz_opts_init(C.byref(opts))
        opts.simd = int(simd)
        rc = _L.fz_compress_batch(self.h, arr, len(items), C.byref(opts))
        for it, (cfg, _, _, _) in zip(arr, items):
            cfg.conf = it.conf
        if rc:
            _raise(rc, self._err())
        return [int(it.pre_len) for it in arr], [int(it.status) for it in arr], opts.t_total_ms

    def decompress_batch(self, items):
        """fz_decompress_batch. items: sequence of (config, dtype, h_pre ndarray, d_pre DeviceBuffer or None,
        d_out DeviceBuffer). Returns (statuses, wall ms)."""
        arr = (fz_batch_dec_item * len(items))()
        keep = []
        for it, (cfg, dtype, h_pre, d_pre, d_out) in zip(arr, items):
            b = np.ascontiguousarray(h_pre, dtype=np.uint8)
            keep.append(b)
            it.conf = cfg.conf
            it.dtype = _DT[np.dtype(dtype)]
            it.h_pre, it.pre_len = b.ctypes.data, b.size
            it.d_pre = d_pre.ptr if d_pre is not None else None
            it.d_out, it.out_cap_elems = d_out.ptr, d_out.nbytes // np.dtype(dtype).itemsize
        opts = fz_opts()
        _L.fz_opts_init(C.byref(opts))
        rc = _L.fz_decompress_batch(self.h, arr, len(items), C.byref(opts))
        if rc:
            _raise(rc, self._err())
        return [int(it.status) for it in arr], opts.t_total_ms


class Server:
    """fz_server: a request loop on one GPU with `slots` workers and an HBM budget in bytes (0 = 80% of the
    free device memory). submit() queues a compress (numpy field -> stream) or decompress (stream -> numpy
    field) request and returns at once; poll() hands back finished requests in completion order."""

    def __init__(self, device=0, slots=8, hbm_budget=0):
        h = C.c_void_p()
        rc = _L.fz_server_create(int(device), int(slots), int(hbm_budget), C.byref(h))
        if rc:
            raise FalconsError(f"cannot create a server on device {device}")
        self.h = h
        self._live = {}  # tag -> (request, buffers) kept alive until polled

    def submit_compress(self, tag, data, config):
        a = np.ascontiguousarray(data)
        dt = _DT[a.dtype]
        cfg = config.copy()
        if cfg.num_elements != a.size:
            cfg.setDims(*a.shape)
        out = np.empty(int(_L.fz_compress_size_bound(C.byref(cfg.conf), dt)), dtype=np.uint8)
        r = fz_request(op=0, dtype=dt, conf=cfg.conf, src=a.ctypes.data, src_len=a.nbytes,
                       dst=out.ctypes.data, dst_cap=out.size, tag=int(tag))
        self._submit(tag, r, (a, out))

    def submit_decompress(self, tag, stream, dtype, num):
        b = np.ascontiguousarray(np.frombuffer(bytes(stream), dtype=np.uint8)
                                 if not isinstance(stream, np.ndarray) else stream, dtype=np.uint8)
        out = np.empty(int(num), dtype=dtype)
        r = fz_request(op=1, dtype=_DT[np.dtype(dtype)], src=b.ctypes.data, src_len=b.size, dst=out.ctypes.data,
                       dst_cap=out.nbytes, tag=int(tag))
        self._submit(tag, r, (b, out))

    def _submit(self, tag, r, bufs):
        if int(tag) in self._live:
            raise ValueError(f"tag {tag} is still in flight")
        self._live[int(tag)] = (r, bufs)
        rc = _L.fz_server_submit(self.h, C.byref(r))
        if rc:
            del self._live[int(tag)]
            _raise(rc, "request rejected (malformed, or larger than the HBM budget)")

    def poll(self, timeout_ms=-1):
        """Next finished request as a dict (tag, op, status, output array, times in ms), or None."""
        res = fz_result()
        if _L.fz_server_poll(self.h, C.byref(res), int(timeout_ms)) == 0:
            return None
        _, bufs = self._live.pop(int(res.tag))
        out = bufs[1]
        if res.op == 0 and res.status == FZ_OK:
            out = out[:res.out_len]
        return {"tag": int(res.tag), "op": int(res.op), "status": int(res.status), "msg": res.msg.decode(),
                "out": out, "field_bytes": int(res.field_bytes), "t_submit_ms": res.t_submit_ms,
                "t_start_ms": res.t_start_ms, "t_done_ms": res.t_done_ms, "t_device_ms": res.t_device_ms,
                "t_zstd_ms": res.t_zstd_ms}

    def stats(self):
        st = fz_server_stats()
        _L.fz_server_get_stats(self.h, C.byref(st))
        return {k: int(getattr(st, k)) for k, _ in fz_server_stats._fields_}

    def close(self):
        if self.h:
            _L.fz_server_destroy(self.h)
            self.h = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class falcons:
    """pyfalcons.falcons: compress / decompress / verify (falcons.pyx:31-333)."""

    # the element types pyfalcons compresses (falcons.pyx:22-28)
    _SUPPORTED_DTYPES = (np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.int32), np.dtype(np.int64))

    @staticmethod
    def compress(data, config, copy=False):
        data = from_dlpack(data, default_context().stream_handle())
        dv = _device_view(data)
        if dv is not None:
            out, _, _ = default_context().compress_ex(data, config)
            return out, dv[1] * _DT_OF[dv[2]].itemsize / float(out.size)
        if data.dtype not in falcons._SUPPORTED_DTYPES:
            raise TypeError(f"Unsupported dtype: {data.dtype}. Supported: float32, float64, int32, int64")
        if data.size == 0:
            raise ValueError("Cannot compress an empty array")
        if not isinstance(config, falconsConfig):
            raise TypeError(f"config must be falconsConfig, got {type(config)}")
        if copy:
            data = np.array(data, order="C", copy=True)
        else:
            if not data.flags["C_CONTIGUOUS"]:
                raise ValueError("copy=False requires a C-contiguous array; pass copy=True")
            if not data.flags["WRITEABLE"]:
                raise ValueError("copy=False requires a writable array; pass copy=True")
        out, _, _ = default_context().compress_ex(data, config)
        return out, data.nbytes / float(out.size)

    @staticmethod
    def decompress(compressed, dtype, shape, config=None):
        compressed = np.asarray(compressed)
        if compressed.dtype != np.uint8:
            raise TypeError(f"Compressed data must be uint8, got {compressed.dtype}")
        if np.dtype(dtype) not in falcons._SUPPORTED_DTYPES:
            raise TypeError(f"Unsupported dtype: {dtype}")
        shape = tuple(int(s) for s in shape)
        if any(s <= 0 for s in shape):
            raise ValueError("Shape dimensions must be positive integers")
        hdr = read_config(compressed)
        if int(np.prod(shape)) != hdr.num_elements:
            raise ValueError("Requested shape does not match the compressed element count")
        out = np.empty(hdr.num_elements, dtype=dtype)
        out, cfg, _ = default_context().decompress_ex(compressed, dtype, out)
        return out.reshape(shape), cfg

    @staticmethod
    def decompress_into(compressed, out):
        """Decompresses into an existing array: a DLPack producer or torch tensor on the GPU (the
        field is written in HBM, no host copy), or a host array. Its element count must be the
        stream's; its dtype gives the value type. Returns (out, config)."""
        compressed = np.asarray(compressed)
        if compressed.dtype != np.uint8:
            raise TypeError(f"Compressed data must be uint8, got {compressed.dtype}")
        out = from_dlpack(out, default_context().stream_handle())
        dv = _device_view(out)
        if dv is not None:
            dt, n = _DT_OF[dv[2]].type, dv[1]
        else:
            dt = out.dtype.type if out.dtype in falcons._SUPPORTED_DTYPES else None
            n = out.size
            if not out.flags["C_CONTIGUOUS"] or not out.flags["WRITEABLE"]:
                raise ValueError("the output array must be C-contiguous and writable")
        if dt is None:
            raise TypeError(f"Unsupported dtype: {out.dtype}")
        hdr = read_config(compressed)
        if n != hdr.num_elements:
            raise ValueError("Output size does not match the compressed element count")
        _, cfg, _ = default_context().decompress_ex(compressed, dt, out)
        return out, cfg

    @staticmethod
    def verify(src_data, dec_data):
        if np.shape(src_data) != np.shape(dec_data):
            raise ValueError("Shape mismatch between src_data and dec_data")
        if src_data.dtype != dec_data.dtype:
            raise ValueError("Dtype mismatch between src_data and dec_data")
        if src_data.size == 0:
            raise ValueError("Cannot verify empty arrays")
        source = src_data.astype(np.float64)
        data_range = float(np.max(source) - np.min(source))
        diff = source - dec_data.astype(np.float64)
        max_diff = float(np.max(np.abs(diff)))
        mse = float(np.mean(diff ** 2))
        if mse == 0:
            nrmse, psnr = 0.0, float("inf")
        elif data_range == 0:
            nrmse, psnr = float("inf"), -float("inf")
        else:
            nrmse = float(np.sqrt(mse) / data_range)
            psnr = float(-20.0 * np.log10(nrmse))
        return max_diff, psnr, nrmse


__all__ = ["falcons", "from_dlpack", "DLDeviceArray", "DeviceArray", "falconsConfig", "falconsErrorBoundMode", "falconsAlgorithm", "falconsInterpAlgorithm",
           "Context", "DeviceBuffer", "zstd_version", "read_config", "check_prezstd", "FalconsError", "FalconsUnsupported", "LIB_PATH", "EXPORTED_SYMBOLS"]
