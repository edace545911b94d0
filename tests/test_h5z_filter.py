"""HDF5 filter 40024 (falcons_amd/h5z/h5z_falcons.c) driven through libhdf5 the way an application
uses the reference's plugin (tools/H5Z-Falcons/src/H5Z_Falcons.cpp): the plugin is found on HDF5's
plugin path, set_local stores the chunk's configuration as Config::save bytes in cd_values, and each
chunk is stored as one Falcons stream.

Pinned against the oracle: the stored cd_values equal the oracle's Config::save of the expected
configuration, every stored chunk equals the oracle's stream for that chunk, and reading the dataset
back gives the oracle's decompression of each chunk. The CPU tests reach only code that returns
before any GPU call (set_local, pass-through of chunks under 20 values, the integer-type error)."""
import ctypes as C
import os

import numpy as np
import pytest

import falcons_amd as fa
from falcons_amd import synth
from h5_lib import H5, PLUGIN_SO, available, cd_from_bytes, cd_to_bytes
from oracle_lib import Oracle, dtype_code

pytestmark = pytest.mark.skipif(not available(), reason="libhdf5 or the filter plugin is not built")


@pytest.fixture(scope="module")
def h5():
    return H5()


@pytest.fixture(scope="module")
def orc():
    o = Oracle()
    o.L.fo_set_simd(1)
    return o


def oracle_conf(orc, chunk, data_type, **kw):
    return orc.config(chunk, data_type=data_type, **kw)


def oracle_cd(orc, conf):
    b = np.zeros(256, np.uint8)
    n = orc.L.fo_config_save(C.byref(conf), b.ctypes.data_as(C.POINTER(C.c_uint8)))
    return cd_from_bytes(b[:n].tobytes())


def user_cd(**kw):
    """cd_values a user builds from a saved configuration (its dims are replaced per chunk)."""
    c = fa.falconsConfig(512, 512, 512)
    for k, v in kw.items():
        setattr(c, k, v)
    return cd_from_bytes(c.save())


class H5Z_class2_t(C.Structure):
    _fields_ = [("version", C.c_int), ("id", C.c_int), ("encoder_present", C.c_uint),
                ("decoder_present", C.c_uint), ("name", C.c_char_p), ("can_apply", C.c_void_p),
                ("set_local", C.c_void_p), ("filter", C.c_void_p)]


def test_plugin_info():
    L = C.CDLL(PLUGIN_SO)
    L.H5PLget_plugin_type.restype = C.c_int
    L.H5PLget_plugin_info.restype = C.POINTER(H5Z_class2_t)
    assert L.H5PLget_plugin_type() == 0  # H5PL_TYPE_FILTER
    info = L.H5PLget_plugin_info().contents
    assert (info.version, info.id, info.encoder_present, info.decoder_present) == (1, 40024, 1, 1)
    assert info.set_local and info.filter and not info.can_apply


def test_filter_found_on_plugin_path(h5):
    assert h5.L.H5Zfilter_avail(40024) > 0


SET_LOCAL = [
    # dataset shape, chunk, dtype, user cd_values (None: the default configuration), oracle fields
    ((40, 48, 64), (20, 48, 64), np.float32, None, {}),
    ((3, 40, 50), (1, 40, 50), np.float64, dict(cmprAlgo=fa.falconsAlgorithm.INTERP,
                                                 errorBoundMode=fa.falconsErrorBoundMode.REL, relErrorBound=1e-3),
     dict(cmpr_algo=2, eb_mode=1, rel_eb=1e-3)),
    ((100,), (50,), np.int16, None, {}),
    ((8, 8), (8, 4), np.uint8, dict(errorBoundMode=fa.falconsErrorBoundMode.ABS, absErrorBound=0.5),
     dict(abs_eb=0.5)),
    ((64, 64), (64, 64), np.int32, dict(errorBoundMode=fa.falconsErrorBoundMode.PSNR, psnrErrorBound=80.0),
     dict(eb_mode=2, psnr_eb=80.0)),
]
DTYPE_CODE = {np.float32: 0, np.float64: 1, np.uint8: 2, np.int16: 5, np.int32: 7}  # Config.hpp:29-38


@pytest.mark.parametrize("shape,chunk,dtype,user,fields", SET_LOCAL)
def test_set_local_cd_values(h5, orc, tmp_path, shape, chunk, dtype, user, fields):
    path = str(tmp_path / "s.h5")
    h5.create_only(path, "d", shape, dtype, chunk, cd=user_cd(**user) if user else None)
    got = h5.cd_values(path, "d")
    want = oracle_cd(orc, oracle_conf(orc, chunk, DTYPE_CODE[dtype], **fields))
    assert got == want
    back = fa.falconsConfig(1)
    b = np.frombuffer(cd_to_bytes(got), np.uint8).copy()
    err = C.create_string_buffer(160)
    assert fa._L.fz_config_load(b.ctypes.data_as(C.POINTER(C.c_uint8)), b.size, C.byref(back.conf), err, 160) == 0
    assert back.dims == tuple(d for d in chunk if d > 1) and back.dataType == DTYPE_CODE[dtype]


def test_chunks_under_20_values_pass_through(h5, tmp_path):
    path = str(tmp_path / "small.h5")
    data = np.arange(19, dtype=np.float32) * 0.25
    h5.write(path, "d", data, (19,))
    raw, mask = h5.raw_chunk(path, "d", (0,))
    assert raw == data.tobytes() and mask == 0
    assert np.array_equal(h5.read(path, "d", data.shape, np.float32), data)


def test_config_load_and_set_dims():
    c = fa.falconsConfig(100, 200, 300)
    c.errorBoundMode = fa.falconsErrorBoundMode.ABS_AND_REL
    c.absErrorBound, c.relErrorBound = 1e-2, 1e-4
    b = np.frombuffer(c.save(), np.uint8).copy()
    back = fa.falconsConfig(1)
    err = C.create_string_buffer(160)
    assert fa._L.fz_config_load(b.ctypes.data_as(C.POINTER(C.c_uint8)), b.size, C.byref(back.conf), err, 160) == 0
    assert back.dims == (100, 200, 300) and back.absErrorBound == 1e-2 and back.relErrorBound == 1e-4
    assert fa._L.fz_config_load(b.ctypes.data_as(C.POINTER(C.c_uint8)), 5, C.byref(back.conf), err, 160) == 1
    assert err.value.decode() == "Truncated Falcons configuration"
    dims = (C.c_uint64 * 5)(7, 1, 9, 1, 11)
    assert fa._L.fz_config_set_dims(C.byref(back.conf), 5, dims) == 0
    assert back.dims == (7, 9, 11) and back.num_elements == 693 and back.blockSize == 6
    assert back.absErrorBound == 1e-2  # other fields kept
    dims = (C.c_uint64 * 5)(2, 2, 2, 2, 2)
    assert fa._L.fz_config_set_dims(C.byref(back.conf), 5, dims) == 1
    assert back.dims == (7, 9, 11)


def _chunk_offsets(shape, chunk):
    grids = [range(0, s, c) for s, c in zip(shape, chunk)]
    return [tuple(o) for o in np.array(np.meshgrid(*grids, indexing="ij")).reshape(len(shape), -1).T]


def _padded_chunk(data, off, chunk):
    out = np.zeros(chunk, data.dtype)  # HDF5 starts a partial chunk from the fill value (0)
    sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(off, chunk, data.shape))
    part = data[sl]
    out[tuple(slice(0, n) for n in part.shape)] = part
    return out, sl, part.shape


GPU_CASES = [
    # name, dataset, chunk, dtype, user cd fields, oracle fields
    ("f32-3d-rel-interp", synth.power_law((40, 48, 64), -3.0, seed=4, dtype=np.float32), (20, 48, 64),
     dict(cmprAlgo=fa.falconsAlgorithm.INTERP, errorBoundMode=fa.falconsErrorBoundMode.REL, relErrorBound=1e-3),
     dict(cmpr_algo=2, eb_mode=1, rel_eb=1e-3)),
    ("f64-2d-chunks-abs-interp-lorenzo", synth.sinusoids((3, 40, 50), seed=5).astype(np.float64), (1, 40, 50),
     dict(errorBoundMode=fa.falconsErrorBoundMode.ABS, absErrorBound=1e-4), dict(abs_eb=1e-4)),
    ("f32-default-config-small-chunks", synth.sinusoids((16, 16, 16), seed=6), (8, 8, 8), None, {}),
    ("f32-2d-partial-edge-chunks", synth.power_law((40, 70), -2.5, seed=7, dtype=np.float32), (32, 32),
     dict(cmprAlgo=fa.falconsAlgorithm.INTERP, errorBoundMode=fa.falconsErrorBoundMode.REL, relErrorBound=1e-4),
     dict(cmpr_algo=2, eb_mode=1, rel_eb=1e-4)),
    # integer chunks, as the reference filter compresses them (H5Z_Falcons.cpp:196-226)
    ("i16-2d-abs-interp", (np.cumsum(np.arange(64 * 96) % 7 - 3).reshape(64, 96) + 1000).astype(np.int16), (32, 48),
     dict(cmprAlgo=fa.falconsAlgorithm.INTERP, errorBoundMode=fa.falconsErrorBoundMode.ABS, absErrorBound=2.0),
     dict(cmpr_algo=2, eb_mode=0, abs_eb=2.0)),
    ("u8-3d-default-config", (np.arange(24 * 30 * 28) % 251).astype(np.uint8).reshape(24, 30, 28), (12, 30, 28),
     None, {}),
    ("i64-3d-rel", (np.cumsum(np.arange(20 * 24 * 26) % 11 - 5) * 1000003).astype(np.int64).reshape(20, 24, 26),
     (20, 24, 13), dict(errorBoundMode=fa.falconsErrorBoundMode.REL, relErrorBound=1e-3),
     dict(eb_mode=1, rel_eb=1e-3)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,data,chunk,user,fields", GPU_CASES, ids=[c[0] for c in GPU_CASES])
def test_filter_chunks_match_oracle(h5, orc, tmp_path, name, data, chunk, user, fields):
    print(f"[h5z] {name} dataset {data.shape} chunk {chunk}", flush=True)
    path = str(tmp_path / "f.h5")
    h5.write(path, "d", data, chunk, cd=user_cd(**user) if user else None)
    code = dtype_code(data.dtype)
    want_read = np.empty_like(data)
    for off in _chunk_offsets(data.shape, chunk):
        raw, mask = h5.raw_chunk(path, "d", off)
        assert mask == 0
        block, sl, part = _padded_chunk(data, off, chunk)
        conf = oracle_conf(orc, chunk, code, **fields)
        stream, _ = orc.compress(block, conf)
        assert raw == stream, f"chunk {off}: stored bytes differ from the oracle's stream"
        dec, _ = orc.decompress(stream, data.dtype)
        want_read[sl] = dec.reshape(chunk)[tuple(slice(0, n) for n in part)]
    got = h5.read(path, "d", data.shape, data.dtype)
    assert np.array_equal(got.view(np.uint8), want_read.view(np.uint8))
