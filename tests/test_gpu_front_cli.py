"""The falcons CLI (tools/falcons/falcons.cpp, unchanged) built against the API front header and
linked to libfalcons_mi355x.so: its compressed files and decompressed fields on the GPU against the
oracle, byte for byte. The binary is built by `make -C oracle shim` where the reference tree is."""
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
    # (shape, dtype, CLI mode, bound, CLI type flag); the CLI's Config keeps ALGO_INTERP_LORENZO
    ((48, 64, 80), np.float32, "REL", 1e-3, "-f"),
    ((40, 50, 60), np.float64, "ABS", 1e-6, "-d"),
    ((300, 257), np.float32, "ABS", 1e-3, "-f"),
]


@pytest.mark.skipif(not os.path.exists(CLI), reason="front builds not made (no reference tree at build time)")
@pytest.mark.parametrize("case", CASES, ids=["3d-f32-rel", "3d-f64-abs", "2d-f32-abs"])
def test_cli_matches_oracle(case, tmp_path):
    shape, dt, mode, eb, tflag = case
    data = synth.power_law(shape, -3.0, seed=9, dtype=dt)
    src, z, o = tmp_path / "in.dat", tmp_path / "in.falcons", tmp_path / "out.dat"
    data.tofile(src)
    dims = [str(s) for s in reversed(shape)]  # the CLI takes the fastest dimension first
    print(f"cli case {shape} {np.dtype(dt).name} {mode} {eb}", flush=True)
    r = subprocess.run([CLI, tflag, "-i", str(src), "-z", str(z), "-o", str(o), f"-{len(shape)}", *dims,
                        "-M", mode, str(eb)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    orc = Oracle()
    orc.L.fo_set_simd(1)
    oc = orc.config(shape, cmpr_algo=1, eb_mode=0 if mode == "ABS" else 1, abs_eb=eb, rel_eb=eb)
    o_stream, _ = orc.compress(data, oc)
    assert z.read_bytes() == o_stream, "the CLI's compressed file differs from the oracle's stream"
    o_dec, _ = orc.decompress(o_stream, dt)
    dec = np.fromfile(o, dtype=dt)
    assert np.array_equal(dec.view(np.uint8), o_dec.reshape(-1).view(np.uint8))
