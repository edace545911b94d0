"""The reference's own callers, unchanged, over the MI355X library.

include/Falcons/api/falcons.hpp is the front a maintainer puts ahead of Falcons' include directory:
it defines Falcons_compress / Falcons_decompress (falcons.hpp:54-215) over the C-ABI and keeps
Falcons' own templates, renamed, for the configurations the GPU path does not produce.
`make -C oracle shim` compiles tools/falcons/falcons.cpp and tools/H5Z-Falcons/src/H5Z_Falcons.cpp
from the reference tree as they are against it and links them to libfalcons_mi355x.so
(oracle/_ref/falcons_cli, oracle/_ref/libh5z_falcons_front.so). Without a GPU this checks the build,
the exported symbols, the HDF5 plugin registration and a call that the front hands to Falcons' own
code (an integer field); the GPU tests run the CLI on float and double fields against the oracle.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
CLI = os.path.join(ROOT, "oracle", "_ref", "falcons_cli")
H5Z_FRONT = os.path.join(ROOT, "oracle", "_ref", "libh5z_falcons_front.so")


@pytest.fixture(scope="module")
def built():
    if os.path.isdir(os.path.join(REF, "include")):
        r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "shim"], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    if not (os.path.exists(CLI) and os.path.exists(H5Z_FRONT)):
        pytest.skip("the reference tree is absent and the front builds were not made")
    return True


def test_cli_links_the_mi355x_library(built):
    out = subprocess.run(["ldd", CLI], capture_output=True, text=True).stdout
    assert "libfalcons_mi355x.so" in out
    syms = subprocess.run(["nm", "-u", CLI], capture_output=True, text=True).stdout
    for s in ("fz_compress", "fz_decompress", "fz_ctx_create", "fz_last_error"):
        assert f" {s}\n" in syms or f" {s}@" in syms, s


def test_h5z_front_registers_filter_40024(built):
    class H5ZClass2(C.Structure):
        _fields_ = [("version", C.c_int), ("id", C.c_int), ("encoder_present", C.c_uint),
                    ("decoder_present", C.c_uint), ("name", C.c_char_p), ("can_apply", C.c_void_p),
                    ("set_local", C.c_void_p), ("filter", C.c_void_p)]

    from h5_lib import HDF5_LIB
    if not os.path.exists(HDF5_LIB):
        pytest.skip("libhdf5 not present")
    C.CDLL(HDF5_LIB, mode=C.RTLD_GLOBAL)  # an HDF5 application has the library loaded before its plugins
    L = C.CDLL(H5Z_FRONT)
    L.H5PLget_plugin_type.restype = C.c_int
    L.H5PLget_plugin_info.restype = C.POINTER(H5ZClass2)
    assert L.H5PLget_plugin_type() == 0
    info = L.H5PLget_plugin_info().contents
    assert info.id == 40024 and info.encoder_present == 1 and info.decoder_present == 1
    syms = subprocess.run(["nm", "-u", H5Z_FRONT], capture_output=True, text=True).stdout
    assert "fz_compress" in syms and "fz_decompress" in syms


def test_cli_openmp_mode_runs_the_reference_template(built, tmp_path):
    """The OpenMP stream format (-P) is outside the GPU path: fz_config_on_gpu rules it out, so the
    front hands the call to Falcons' own template without opening a GPU context (none is visible
    here), and the round trip stays within the bound."""
    a = np.cumsum(np.random.default_rng(3).standard_normal((16, 24, 40)), axis=2).astype(np.float32)
    src = tmp_path / "in.dat"
    a.tofile(src)
    z, o = tmp_path / "in.falcons", tmp_path / "out.dat"
    r = subprocess.run([CLI, "-f", "-P", "-i", str(src), "-z", str(z), "-o", str(o), "-3", "40", "24", "16",
                        "-M", "ABS", "1e-2"], capture_output=True, text=True, env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert r.returncode == 0, r.stdout + r.stderr
    dec = np.fromfile(o, dtype=np.float32).reshape(a.shape)
    assert np.abs(dec.astype(np.float64) - a).max() <= 1e-2
    assert z.stat().st_size < a.nbytes


def test_config_on_gpu_rules():
    """fz_config_on_gpu: host-only answer to "may this configuration run on the GPU path"."""
    import falcons_amd as fa
    L = C.CDLL(fa.LIB_PATH)
    L.fz_config_on_gpu.restype = C.c_int
    L.fz_config_on_gpu.argtypes = [C.c_void_p, C.c_int]

    def on(dtype=0, **kw):
        c = fa.falconsConfig(20, 30, 40)
        for k, v in kw.items():
            setattr(c, k, v)
        return L.fz_config_on_gpu(C.byref(c.conf), dtype)

    assert on() == 0 and on(dtype=9) == 0 and on(cmprAlgo=1) == 0
    assert on(dtype=10) != 0 and on(dtype=-1) != 0
    assert on(openmp=True) != 0
    assert on(quantbinCnt=1024) != 0
    assert on(cmprAlgo=0) != 0 and on(cmprAlgo=3) != 0 and on(cmprAlgo=4) != 0
