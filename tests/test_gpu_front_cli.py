"""The falcons CLI (tools/falcons/falcons.cpp, unchanged) built against the API front header and
linked to libfalcons_mi355x.so: its compressed files and decompressed fields on the GPU against the
oracle. The binary is built by `make -C oracle shim` where the reference tree is.

Compared byte for byte: the stream's Config header, the buffer inside its zstd frame (the pre-zstd
buffer), the CLI's decompressed field, and the CLI's decompression of the oracle's stream. The zstd
frame itself is not: like the reference's own CLI (its CMake links the vendored zfse objects into
the executable beside a shared libzstd), this process binds libzstd's calls of its exported FSE_*
functions to the vendored ones, so libzstd 1.4.9 frames the same buffer a few bytes differently
from a process with libzstd alone (the oracle's and the library's)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from falcons_amd import synth
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "oracle", "_ref", "falcons_cli")

CASES = [
    # (shape, dtype, CLI mode, bound, CLI type flags); the CLI's Config keeps ALGO_INTERP_LORENZO
    ((48, 64, 80), np.float32, "REL", 1e-3, ["-f"]),
    ((40, 50, 60), np.float64, "ABS", 1e-6, ["-d"]),
    ((300, 257), np.float32, "ABS", 1e-3, ["-f"]),
    ((24, 36, 40), np.int32, "ABS", 2, ["-I", "32"]),
    ((80, 90), np.uint16, "REL", 1e-3, ["-U", "16"]),  # the CLI sizes its buffer 2 * num * sizeof(T) (falcons.cpp:137)
]


def _field(shape, dt):
    if np.dtype(dt).kind == "f":
        return synth.power_law(shape, -3.0, seed=9, dtype=dt)
    # a smooth integer field centred in the type's range
    info = np.iinfo(dt)
    w = np.cumsum(np.random.default_rng(9).integers(-4, 5, int(np.prod(shape))))
    w = w - (w.min() + w.max()) // 2 + (int(info.min) + int(info.max)) // 2
    return np.clip(w, info.min, info.max).astype(dt).reshape(shape)


@pytest.mark.skipif(not os.path.exists(CLI), reason="front builds not made (no reference tree at build time)")
@pytest.mark.parametrize("case", CASES, ids=["3d-f32-rel", "3d-f64-abs", "2d-f32-abs", "3d-i32-abs", "2d-u16-rel"])
def test_cli_matches_oracle(case, tmp_path):
    shape, dt, mode, eb, tflag = case
    data = _field(shape, dt)
    src, z, o = tmp_path / "in.dat", tmp_path / "in.falcons", tmp_path / "out.dat"
    data.tofile(src)
    dims = [str(s) for s in reversed(shape)]  # the CLI takes the fastest dimension first
    print(f"cli case {shape} {np.dtype(dt).name} {mode} {eb}", flush=True)
    r = subprocess.run([CLI, *tflag, "-i", str(src), "-z", str(z), "-o", str(o), f"-{len(shape)}", *dims,
                        "-M", mode, str(eb)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    orc = Oracle()
    orc.L.fo_set_simd(1)
    oc = orc.config(shape, cmpr_algo=1, eb_mode=0 if mode == "ABS" else 1, abs_eb=eb, rel_eb=eb)
    o_stream, _, _ = orc.compress(data, oc, want_pre=True)
    cli_stream = z.read_bytes()
    h = o_stream[0]
    assert cli_stream[:h] == o_stream[:h], "Config header differs from the oracle's"
    # the buffer inside the zstd frame: the pre-zstd buffer, or the reconstructed field where the
    # lossless form won (ratio below 3, FalconsDispatcher.hpp:121-133; the header says which)
    o_inner = _unzstd(o_stream[h:])
    assert _unzstd(cli_stream[h:]) == o_inner, "the buffer inside the CLI's zstd frame differs from the oracle's"
    o_dec, _ = orc.decompress(o_stream, dt)
    dec = np.fromfile(o, dtype=dt)
    assert np.array_equal(dec.view(np.uint8), o_dec.reshape(-1).view(np.uint8))
    # the CLI decompresses the oracle's stream to the oracle's field
    zo, oo = tmp_path / "oracle.falcons", tmp_path / "oracle.out"
    zo.write_bytes(o_stream)
    r = subprocess.run([CLI, *tflag, "-z", str(zo), "-o", str(oo), f"-{len(shape)}", *dims], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert np.array_equal(np.fromfile(oo, dtype=dt).view(np.uint8), o_dec.reshape(-1).view(np.uint8))


def _unzstd(payload):
    """Lossless_zstd's framing: [size_t n][zstd frame of n bytes]."""
    z = C.CDLL("/opt/conda/lib/libzstd.so.1")
    z.ZSTD_decompress.restype = C.c_size_t
    z.ZSTD_isError.restype = C.c_uint
    n = int.from_bytes(payload[:8], "little")
    out = C.create_string_buffer(max(n, 1))
    r = z.ZSTD_decompress(out, n, payload[8:], len(payload) - 8)
    assert not z.ZSTD_isError(r) and r == n
    return out.raw[:n]
