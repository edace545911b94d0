"""ALGO_INTERP_LORENZO on the GPU against the oracle: the parameter search's candidate ratios (the zstd
sizes of every candidate's tuning buffer), the chosen parameters and the resulting stream must be the
oracle's, bit for bit, for 3-D (profiled and unprofiled sampling), 2-D (eight candidates), 1-D, f32 and
f64, with and without the anchor grid."""
import ctypes as C

import numpy as np
import pytest

from falcons_amd import synth
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import falcons_amd
    c = falcons_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def oracle():
    o = Oracle()
    o.L.fo_set_simd(1)
    return o


def bump_field(shape, dtype):
    """Mostly constant with one varying corner: few sampled blocks vary by more than the bound, so the
    sampler takes its unprofiled (regular stride) branch."""
    d = np.zeros(shape, dtype=np.float64)
    sl = tuple(slice(0, s // 4) for s in shape)
    d[sl] = synth.sinusoids(tuple(s // 4 for s in shape), seed=3, dtype=np.float64)
    return d.astype(dtype)


CASES = [
    # name, data factory, eb_mode, eb, anchor stride
    ("3d-f32-power", lambda: synth.power_law((96, 100, 104), -3.0, seed=21, dtype=np.float32), 0, 1e-3, -1),
    ("3d-f64-rel", lambda: synth.power_law((100, 96, 90), -3.0, seed=22, dtype=np.float64), 1, 1e-5, -1),
    ("3d-f32-sin", lambda: synth.sinusoids((100, 100, 100), seed=23, dtype=np.float32), 0, 1e-2, -1),
    ("3d-f32-noanchor", lambda: synth.power_law((96, 96, 96), -3.0, seed=24, dtype=np.float32), 0, 1e-3, 0),
    ("3d-f32-bump", lambda: bump_field((128, 120, 112), np.float32), 0, 1e-3, -1),
    ("2d-f32", lambda: synth.power_law((600, 700), -3.0, seed=25, dtype=np.float32), 0, 1e-3, -1),
    ("2d-f64", lambda: synth.power_law((700, 640), -3.0, seed=26, dtype=np.float64), 1, 1e-4, -1),
    ("1d-f32", lambda: synth.sinusoids((1_500_000,), seed=27, dtype=np.float32), 0, 5e-2, -1),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_tuner_matches_oracle(ctx, oracle, case):
    import falcons_amd
    _, make, mode, eb, anchor = case
    data = make()
    shape = data.shape
    c = falcons_amd.falconsConfig(*shape)
    oc = oracle.config(shape)
    for k, v in dict(cmpr_algo=1, eb_mode=mode, abs_eb=eb, rel_eb=eb, interp_anchor_stride=anchor).items():
        setattr(c.conf, k, v)
        setattr(oc, k, v)
    o_stream, o_mut, o_pre = oracle.compress(data, oc, want_pre=True)
    buf = (C.c_double * 16)()
    n = oracle.L.fo_tune_report(buf, 16)
    o_ratios = [float(buf[i]) for i in range(n)]
    assert n >= 4, "the case was meant to run the parameter search"
    g_data = data.copy()
    g_stream, _, g_pre = ctx.compress_ex(g_data, c, want_prezstd=True)
    assert ctx.tune_report() == o_ratios
    assert g_pre == o_pre
    assert bytes(g_stream) == o_stream
    assert np.array_equal(g_data.view(np.uint8), o_mut.view(np.uint8))


def test_tuner_device_resident(ctx, oracle):
    """fz_compress_device accepts ALGO_INTERP_LORENZO: the search runs on the field in HBM."""
    import falcons_amd
    data = synth.power_law((96, 100, 104), -3.0, seed=31, dtype=np.float32)
    c = falcons_amd.falconsConfig(*data.shape)
    c.cmprAlgo = falcons_amd.falconsAlgorithm.INTERP_LORENZO
    c.errorBoundMode = falcons_amd.falconsErrorBoundMode.REL
    c.relErrorBound = 1e-3
    oc = oracle.config(data.shape, cmpr_algo=1, eb_mode=1, rel_eb=1e-3)
    _, _, o_pre = oracle.compress(data, oc, want_pre=True)
    d = ctx.alloc(data.nbytes)
    p = ctx.alloc(ctx.prezstd_bound(c, np.float32))
    d.upload(data)
    run = c.copy()
    n, _ = ctx.compress_device(run, np.float32, d, p)
    assert p.download(n).tobytes() == o_pre
    assert run.cmprAlgo == falcons_amd.falconsAlgorithm.INTERP
    d.free()
    p.free()


def test_tuner_1d_low_ratio_goes_to_lorenzo(ctx, oracle):
    """1-D fields whose best interpolation ratio is below 50 go on to the Lorenzo/regression test in the
    reference (FalconsAlgoInterp.hpp:479-497), which this path does not build: both sides refuse."""
    import falcons_amd
    data = synth.sinusoids((1_500_000,), seed=28, dtype=np.float32)
    c = falcons_amd.falconsConfig(*data.shape)
    c.cmprAlgo = falcons_amd.falconsAlgorithm.INTERP_LORENZO
    c.absErrorBound = 1e-4
    oc = oracle.config(data.shape, cmpr_algo=1, eb_mode=0, abs_eb=1e-4)
    with pytest.raises(ValueError):
        oracle.compress(data, oc)
    buf = (C.c_double * 16)()
    n = oracle.L.fo_tune_report(buf, 16)
    with pytest.raises(falcons_amd.FalconsUnsupported):
        ctx.compress_ex(data.copy(), c)
    assert ctx.tune_report() == [float(buf[i]) for i in range(n)]
