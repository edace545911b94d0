"""The C-ABI library without a GPU: it loads, exports every entry point the header declares, and its
host-side functions (Config save / size_est / read_config / size bound, Falcons_read_config checks,
pyfalcons argument validation) agree byte for byte and message for message with the oracle and the
reference's behaviour. Nothing here launches a kernel."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import falcons_amd as fa
from oracle_lib import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "falcons_mi355x.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fz_\w+)\s*\(", text)))


def test_header_symbols_exported():
    names = declared()
    assert len(names) >= 19
    lib = C.CDLL(fa.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(fa.EXPORTED_SYMBOLS) == names


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", fa.LIB_PATH], capture_output=True, text=True, check=True)
    syms = {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}
    for n in declared():
        assert n in syms, n


def test_import_fails_loudly_without_library(tmp_path):
    pkg = tmp_path / "falcons_amd"
    pkg.mkdir()
    for f in ("__init__.py", "synth.py"):
        (pkg / f).write_text(open(os.path.join(ROOT, "falcons_amd", f)).read())
    r = subprocess.run([sys.executable, "-c", "import falcons_amd"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0 and "ImportError" in r.stderr and "libfalcons_mi355x.so" in r.stderr


CFGS = [
    dict(shape=(512, 512, 512), eb_mode=0, abs_eb=1e-3),
    dict(shape=(512, 512, 512), eb_mode=1, rel_eb=1e-3),
    dict(shape=(33, 40, 65), eb_mode=2, psnr_eb=80.0, cmpr_algo=1),
    dict(shape=(300, 257), eb_mode=3, l2norm_eb=0.5),
    dict(shape=(5000,), eb_mode=4, abs_eb=1e-2, rel_eb=1e-4),
    dict(shape=(17, 1, 3, 4), eb_mode=5, abs_eb=2e-2, rel_eb=1e-3, cmpr_algo=0),
    dict(shape=(1 << 40,), eb_mode=0, abs_eb=1.0, cmpr_algo=4),
    dict(shape=(70000, 3), eb_mode=0, abs_eb=1e-6, lorenzo=0, regression=1, openmp=1),
]


def _pair(orc, spec):
    spec = dict(spec)
    shape = spec.pop("shape")
    g = fa.falconsConfig(*shape)
    o = orc.config(shape)
    for k, v in spec.items():
        setattr(g.conf, k, v)
        setattr(o, k, v)
    return g, o


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.mark.parametrize("spec", CFGS)
def test_config_save_matches_oracle(orc, spec):
    g, o = _pair(orc, spec)
    gb = (C.c_uint8 * 256)()
    ob = np.zeros(256, np.uint8)
    gl = fa._L.fz_config_save(C.byref(g.conf), gb)
    ol = orc.L.fo_config_save(C.byref(o), ob.ctypes.data_as(C.POINTER(C.c_uint8)))
    assert gl == ol and bytes(gb)[:gl] == ob[:ol].tobytes()
    assert fa._L.fz_config_size_est(C.byref(g.conf)) >= gl
    for dt in (0, 1):
        assert fa._L.fz_compress_size_bound(C.byref(g.conf), dt) == orc.L.fo_size_bound(C.byref(o), dt)
    back = fa.read_config(bytes(gb)[:gl] + b"\0")
    assert back.dims == g.dims and back.conf.eb_mode == g.conf.eb_mode
    assert back.conf.cmpr_algo == g.conf.cmpr_algo


def test_header_layout_512_abs():
    # Config::save (Config.hpp:316-362): 44 bytes for a 512^3 ABS field
    g = fa.falconsConfig(512, 512, 512)
    g.errorBoundMode = fa.falconsErrorBoundMode.ABS
    g.absErrorBound = 1e-3
    b = g.save()
    # [u8 size][u32 magic][u32 data version][i8 N][u8 bit width][3 x 10-bit dims in 4 bytes][u64 num]
    # [u8 algo][u8 eb mode][f64 abs eb][u8 flags][u8 data type][i32 quantbinCnt][i32 blockSize][u8 predDim]
    assert len(b) == 44 and b[0] == 44
    assert int.from_bytes(b[1:5], "little") == 0xF342F310
    assert b[9] == 3 and b[10] == 10
    dims = int.from_bytes(b[11:15], "little")
    assert [(dims >> (10 * i)) & 1023 for i in range(3)] == [512, 512, 512]
    assert int.from_bytes(b[15:23], "little") == 512 ** 3
    assert b[23] == 1 and b[24] == 0  # default ALGO_INTERP_LORENZO, EB_ABS
    assert np.frombuffer(b[25:33], np.float64)[0] == 1e-3
    assert int.from_bytes(b[35:39], "little") == 65536


@pytest.mark.parametrize("blob,msg", [
    (b"", "Truncated Falcons configuration"),
    (bytes([44]) + b"\0" * 10, "Truncated Falcons configuration"),
    (bytes([5, 5, 0, 0, 0]), "Truncated Falcons configuration field"),
])
def test_read_config_truncation_messages(blob, msg):
    with pytest.raises(ValueError, match=msg):
        fa.read_config(blob)


def test_read_config_rejects_bad_magic_and_missing_payload():
    g = fa.falconsConfig(10, 10, 10)
    g.absErrorBound = 1e-3
    b = bytearray(g.save())
    with pytest.raises(ValueError, match="Missing Falcons compressed payload"):
        fa.read_config(bytes(b))
    bad = bytearray(b) + b"\0"
    bad[1] ^= 0xFF
    with pytest.raises(ValueError, match="magic number mismatch"):
        fa.read_config(bytes(bad))
    dims = bytearray(b) + b"\0"
    dims[9] = 7  # N out of range
    with pytest.raises(ValueError, match="Invalid Falcons dimensions"):
        fa.read_config(bytes(dims))


def test_pyfalcons_config_validation():
    c = fa.falconsConfig(4, 5, 6)
    assert c.num_elements == 120 and c.dims == (4, 5, 6)
    with pytest.raises(ValueError):
        c.setDims()
    with pytest.raises(ValueError):
        c.setDims(4, 0)
    c.relErrorBound = 1e-4
    c.setDims(8, 8)
    assert c.relErrorBound == 1e-4 and c.num_elements == 64


def test_pyfalcons_decompress_validation():
    with pytest.raises(TypeError):
        fa.falcons.decompress(np.zeros(10, np.int8), np.float32, (10,))
    with pytest.raises(TypeError):  # pyfalcons takes float32, float64, int32 and int64 (falcons.pyx:22-28)
        fa.falcons.decompress(np.zeros(10, np.uint8), np.int16, (10,))
    with pytest.raises(ValueError):
        fa.falcons.decompress(np.zeros(10, np.uint8), np.float32, (0, 3))
    g = fa.falconsConfig(10, 10, 10)
    g.absErrorBound = 1e-3
    stream = np.frombuffer(g.save() + b"\0" * 16, np.uint8)
    with pytest.raises(ValueError, match="shape"):
        fa.falcons.decompress(stream, np.float32, (10, 10, 9))


def test_pyfalcons_verify():
    a = np.linspace(0, 1, 1000, dtype=np.float32)
    b = a + np.float32(1e-3)
    md, psnr, nrmse = fa.falcons.verify(a, b)
    assert abs(md - 1e-3) < 1e-6 and abs(nrmse - 1e-3) < 1e-5 and abs(psnr - 60.0) < 0.01
    assert fa.falcons.verify(a, a.copy())[1] == float("inf")
    with pytest.raises(ValueError):
        fa.falcons.verify(a, b[:10])


def test_zstd_version_recorded():
    assert fa.zstd_version() == Oracle().L.fo_zstd_version()
    assert fa.zstd_version() >= 10400
