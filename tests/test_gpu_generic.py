"""Integer element types and 4-D fields on the GPU against the oracle, byte for byte.

These take the reference's generic traversal (interpolation_1d for 1-D and 2-D fields,
interpolation_1d_fastest_dim_first for 3-D and 4-D ones, IDH:374-528, 2314-2519) with the integral
quantizer for integers (LinearQuantizer.hpp:190-248) and the double compare for 4-D floats. Every case
compares the buffer handed to zstd, the final stream, the reconstructed input and the decompressed
field with the oracle; the cases cover the ten Falcons element types, 1-D to 4-D, linear and cubic,
several axis orders, anchors on and off, REL bounds on integer ranges, values at the type limits
(escapes and clamped predictions), and the parameter search (ALGO_INTERP_LORENZO).
"""
import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

INT_TYPES = [np.uint8, np.int8, np.uint16, np.int16, np.uint32, np.int32, np.uint64, np.int64]


@pytest.fixture(scope="module")
def ctx():
    import falcons_amd
    c = falcons_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc():
    o = Oracle()
    o.L.fo_set_simd(1)
    return o


def int_field(shape, dt, seed, spread=3, extremes=False):
    """A smooth integer field: a random walk in the wide type, centred in the type's range."""
    info = np.iinfo(dt)
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    w = np.cumsum(rng.integers(-spread, spread + 1, n)).astype(np.int64)
    w -= (int(w.min()) + int(w.max())) // 2
    mid = (int(info.min) + int(info.max)) // 2
    if info.bits < 64:
        a = np.clip(w + mid, int(info.min), int(info.max)).astype(dt)
    else:
        a = np.array([min(max(int(x) + mid, int(info.min)), int(info.max)) for x in w], dtype=dt)
    if extremes:  # type limits scattered in: escapes and clamped predictions around them
        idx = rng.choice(n, size=max(1, n // 50), replace=False)
        a.reshape(-1)[idx] = np.where(rng.random(idx.size) < 0.5, info.min, info.max).astype(dt)
    return a.reshape(shape)


def float_field(shape, dt, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape)
    for ax in range(len(shape)):
        a = np.cumsum(a, axis=ax)
    return (a / np.abs(a).max()).astype(dt)


def _case(ctx, orc, data, algo=2, mode=0, eb=1.0, interp=1, direction=0, anchor=None):
    import falcons_amd
    shape = data.shape
    c = falcons_amd.falconsConfig(*shape)
    c.cmprAlgo = algo
    c.errorBoundMode = mode
    c.absErrorBound = eb
    c.relErrorBound = eb
    c.interpAlgo = interp
    c.interpDirection = direction
    kw = dict(cmpr_algo=algo, eb_mode=mode, abs_eb=eb, rel_eb=eb, interp_algo=interp, interp_direction=direction)
    if anchor is not None:
        c.interpAnchorStride = anchor
        kw["interp_anchor_stride"] = anchor
    oc = orc.config(shape, **kw)
    print(f"generic case {shape} {data.dtype} algo {algo} mode {mode} eb {eb} interp {interp} dir {direction} "
          f"anchor {anchor}", flush=True)
    o_stream, o_mut, o_pre = orc.compress(data, oc, want_pre=True)
    g = data.copy()
    g_stream, opts, g_pre = ctx.compress_ex(g, c, want_prezstd=True)
    assert g_pre == o_pre, "pre-zstd buffer differs from the oracle"
    assert bytes(g_stream) == o_stream, "stream differs from the oracle"
    assert np.array_equal(g.view(np.uint8), o_mut.view(np.uint8)), "reconstructed input differs"
    dec, _, _ = ctx.decompress_ex(g_stream, data.dtype)
    o_dec, _ = orc.decompress(o_stream, data.dtype)
    assert np.array_equal(dec.reshape(shape).view(np.uint8), o_dec.reshape(shape).view(np.uint8))
    return opts


SHAPES = [(3000,), (70, 90), (20, 30, 40), (9, 10, 11, 12)]


@pytest.mark.parametrize("dt", INT_TYPES, ids=[np.dtype(t).name for t in INT_TYPES])
@pytest.mark.parametrize("shape", SHAPES, ids=["1d", "2d", "3d", "4d"])
@pytest.mark.parametrize("interp", [0, 1], ids=["linear", "cubic"])
def test_integer_fields_abs(ctx, orc, dt, shape, interp):
    data = int_field(shape, dt, seed=len(shape) * 10 + interp)
    _case(ctx, orc, data, eb=2.0, interp=interp)


@pytest.mark.parametrize("dt", INT_TYPES, ids=[np.dtype(t).name for t in INT_TYPES])
def test_integer_extremes_and_rel(ctx, orc, dt):
    data = int_field((24, 31, 29), dt, seed=7, spread=40, extremes=True)
    _case(ctx, orc, data, mode=1, eb=1e-3, interp=1)
    _case(ctx, orc, data, eb=0.5, interp=0)  # bound below one unit: everything codes exactly


@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("interp", [0, 1], ids=["linear", "cubic"])
@pytest.mark.parametrize("direction", [0, 7, 23])
def test_4d_float(ctx, orc, dt, interp, direction):
    data = float_field((9, 13, 10, 11), dt, seed=direction + 3 * interp)
    _case(ctx, orc, data, eb=1e-3, interp=interp, direction=direction)


@pytest.mark.parametrize("dt", [np.float32, np.int32], ids=["f32", "i32"])
def test_4d_anchor_grid(ctx, orc, dt):
    """Every axis above the default anchor stride of 16: the anchor grid is stored losslessly."""
    data = float_field((20, 18, 19, 33), dt, 11) if dt == np.float32 else int_field((20, 18, 19, 33), dt, 11)
    _case(ctx, orc, data, mode=1, eb=1e-3, interp=1)
    _case(ctx, orc, data, mode=1, eb=1e-3, interp=0, anchor=0)


@pytest.mark.parametrize("shape,direction", [((70, 90), 1), ((20, 30, 40), 5), ((20, 30, 40), 3)])
def test_integer_directions(ctx, orc, shape, direction):
    _case(ctx, orc, int_field(shape, np.int16, 5), eb=1.0, interp=1, direction=direction)


@pytest.mark.parametrize("dt,shape", [(np.int32, (40, 48, 44)), (np.uint16, (20, 20, 20, 20)),
                                      (np.float32, (20, 20, 20, 20)), (np.int64, (300, 280))],
                         ids=["i32-3d", "u16-4d", "f32-4d", "i64-2d"])
def test_parameter_search(ctx, orc, dt, shape):
    """ALGO_INTERP_LORENZO: the candidates' ratios and the chosen settings match the oracle's search."""
    import ctypes
    data = float_field(shape, dt, 21) if np.dtype(dt).kind == "f" else int_field(shape, dt, 21, spread=5)
    _case(ctx, orc, data, algo=1, mode=1, eb=1e-3)
    r = np.zeros(16)
    n = orc.L.fo_tune_report(r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 16)
    g = ctx.tune_report()
    assert n > 0 and list(g) == list(r[:n])
