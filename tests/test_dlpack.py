"""pyfalcons over device-resident arrays through DLPack, the interface CuPy, JAX, numba and torch
arrays share. The library reads the device pointer, shape and dtype out of the DLPack capsule itself
(no other GPU runtime takes part), and its own device buffers export DLPack, which is the producer
the GPU test uses: in one process every GPU array must come from the HIP runtime the library uses."""
import ctypes as C

import numpy as np
import pytest

import falcons_amd as fa
from falcons_amd import synth
from oracle_lib import Oracle


class Producer:
    """A foreign array: only __dlpack__ and __dlpack_device__."""

    def __init__(self, inner, device=None):
        self.inner, self.device = inner, device

    def __dlpack__(self, **kw):
        return self.inner.__dlpack__(**kw)

    def __dlpack_device__(self):
        return self.device or self.inner.__dlpack_device__()


class FakeBuffer:
    """Stands in for a DeviceBuffer when only the capsule is inspected (nothing touches the GPU)."""

    class _Ctx:
        device = 0

    def __init__(self, ptr, nbytes):
        self.ptr, self.nbytes, self.ctx = C.c_void_p(ptr), nbytes, self._Ctx()


def test_host_producer_becomes_numpy_without_copy():
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    b = fa.from_dlpack(Producer(a))
    assert isinstance(b, np.ndarray) and b.shape == (3, 4)
    assert b.__array_interface__["data"][0] == a.__array_interface__["data"][0]


def test_passthrough_and_rejects():
    a = np.ones(4, np.float32)
    assert fa.from_dlpack(a) is a
    obj = object()
    assert fa.from_dlpack(obj) is obj
    with pytest.raises(TypeError):
        fa.from_dlpack(Producer(a, device=(4, 0)))  # kDLOpenCL


def test_capsule_fields_read_without_copy():
    a = np.arange(60, dtype=np.float64).reshape(3, 4, 5)
    v = fa.DLDeviceArray(a)
    assert (v.ptr, v.shape, v.dtype, v.size, v.nbytes) == (a.ctypes.data, (3, 4, 5), np.float64, 60, 480)
    assert v.device_type == 1
    v.release()
    v.release()  # the deleter runs once
    with pytest.raises(ValueError):
        fa.DLDeviceArray(a[:, ::2])
    with pytest.raises(TypeError):  # element types outside Falcons_FLOAT .. Falcons_INT64
        fa.DLDeviceArray(np.arange(6, dtype=np.float16))
    v = fa.DLDeviceArray(np.arange(6, dtype=np.int32))  # integer fields are compressed too
    assert v.dtype == np.int32
    v.release()


def test_device_array_exports_a_rocm_capsule():
    arr = fa.DeviceArray(FakeBuffer(0x7F0000001000, 4096), (16, 8), np.float32)
    assert arr.__dlpack_device__() == (10, 0)  # kDLROCM
    v = fa.from_dlpack(arr)
    assert isinstance(v, fa.DLDeviceArray)
    assert (v.ptr, v.shape, v.dtype, v.device_type) == (0x7F0000001000, (16, 8), np.float32, 10)
    with pytest.raises(ValueError):
        fa.DeviceArray(FakeBuffer(0x1000, 64), (16, 8), np.float32)


def test_consumer_stream_reaches_the_producer():
    seen = {}

    class Recorder(Producer):
        def __dlpack__(self, **kw):
            seen.update(kw)
            return self.inner.__dlpack__(**kw)

    arr = fa.DeviceArray(FakeBuffer(0x7F0000003000, 256), (64,), np.float32)
    v = fa.from_dlpack(Recorder(arr), stream=0x1234)
    assert seen == {"stream": 0x1234} and v.ptr == 0x7F0000003000
    seen.clear()
    fa.from_dlpack(Recorder(arr))
    assert seen == {}  # no stream: the producer's default


def test_exported_capsule_outlives_its_producer():
    import gc
    before = len(fa._DL_EXPORTS)
    v = fa.from_dlpack(fa.DeviceArray(FakeBuffer(0x7F0000002000, 1024), (256,), np.float32))
    gc.collect()  # the producer object is gone; the capsule's tensor must still be readable
    assert (v.ptr, v.shape) == (0x7F0000002000, (256,))
    assert len(fa._DL_EXPORTS) == before + 1
    v.release()
    assert len(fa._DL_EXPORTS) == before


@pytest.mark.gpu
def test_device_arrays_compress_and_decompress_into():
    shape = (48, 40, 36)
    data = synth.power_law(shape, -3.0, seed=21, dtype=np.float32)
    cfg = fa.falconsConfig(*shape)
    cfg.cmprAlgo = fa.falconsAlgorithm.INTERP
    cfg.errorBoundMode = fa.falconsErrorBoundMode.REL
    cfg.relErrorBound = 1e-3
    orc = Oracle()
    orc.L.fo_set_simd(1)
    o_stream, _ = orc.compress(data, orc.config(shape, cmpr_algo=2, eb_mode=1, rel_eb=1e-3))
    o_dec, _ = orc.decompress(o_stream, np.float32)

    ctx = fa.default_context()
    src = ctx.alloc(data.nbytes)
    dst = ctx.alloc(data.nbytes)
    try:
        src.upload(data)
        print("[dlpack] compress from a device array", flush=True)
        stream, ratio = fa.falcons.compress(Producer(src.as_array(shape, np.float32)), cfg)
        assert bytes(stream) == o_stream and ratio > 1
        print("[dlpack] decompress into a device array", flush=True)
        out, back = fa.falcons.decompress_into(stream, Producer(dst.as_array(shape, np.float32)))
        assert isinstance(out, fa.DLDeviceArray) and out.ptr == dst.ptr.value and back.dims == shape
        assert np.array_equal(dst.download(data.nbytes, np.uint32), o_dec.reshape(-1).view(np.uint32))
        with pytest.raises(ValueError):
            fa.falcons.decompress_into(stream, dst.as_array((10,), np.float32))
    finally:
        src.free()
        dst.free()
