"""GPU path against the oracle on the same seeded inputs: the buffer handed to zstd, the final stream
and the decompressed field must be identical bit for bit."""
import numpy as np
import pytest

from falcons_amd import synth
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import falcons_amd
    return falcons_amd.Context(0)


@pytest.fixture(scope="module")
def oracle():
    return Oracle()


def _cfg_pair(oracle, shape, **kw):
    import falcons_amd
    c = falcons_amd.falconsConfig(*shape)
    oc = oracle.config(shape)
    for k, v in kw.items():
        setattr(c.conf, k, v)
        setattr(oc, k, v)
    return c, oc


CASES = [
    # shape, dtype, algo(2 interp), interp_algo, direction, eb_mode, eb, anchor
    ((33, 33, 33), np.float32, 1, 0, 0, 1e-3, -1),
    ((33, 33, 33), np.float32, 0, 0, 0, 1e-3, -1),
    ((17, 40, 65), np.float32, 1, 5, 0, 1e-3, -1),
    ((17, 40, 65), np.float32, 0, 5, 1, 1e-3, -1),
    ((64, 48, 100), np.float32, 1, 0, 1, 1e-4, -1),
    ((64, 48, 100), np.float64, 1, 0, 0, 1e-6, -1),
    ((20, 30, 31), np.float64, 0, 5, 0, 1e-5, -1),
    ((5, 6, 7), np.float32, 1, 0, 0, 1e-3, -1),
    ((2, 3, 5), np.float32, 0, 0, 0, 1e-3, -1),
    ((300, 257), np.float32, 1, 0, 0, 1e-3, -1),
    ((300, 257), np.float32, 0, 1, 1, 1e-3, -1),
    ((129, 130), np.float64, 1, 1, 0, 1e-7, -1),
    ((5000,), np.float32, 1, 0, 0, 1e-3, -1),
    ((4097,), np.float64, 0, 0, 1, 1e-4, -1),
    ((40, 40, 40), np.float32, 1, 0, 0, 1e-3, 0),
    ((32, 32, 32), np.float32, 1, 0, 0, 1e-3, -1),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{np.dtype(c[1]).name}-a{c[2]}d{c[3]}m{c[4]}" for c in CASES])
@pytest.mark.parametrize("simd", [1, 0])
def test_stream_parity(ctx, oracle, case, simd):
    shape, dtype, ialgo, direction, mode, eb, anchor = case
    data = synth.sinusoids(shape, seed=7, dtype=dtype)
    kw = dict(cmpr_algo=2, eb_mode=mode, abs_eb=eb, rel_eb=eb, interp_algo=ialgo, interp_direction=direction,
              interp_anchor_stride=anchor)
    c, oc = _cfg_pair(oracle, shape, **kw)
    oracle.L.fo_set_simd(simd)
    o_stream, o_mut, o_pre = oracle.compress(data, oc, want_pre=True)
    g_data = data.copy()
    g_stream, opts, g_pre = ctx.compress_ex(g_data, c, simd=simd, want_prezstd=True)
    assert len(g_pre) == len(o_pre)
    assert g_pre == o_pre
    assert bytes(g_stream) == o_stream
    assert np.array_equal(g_data.view(np.uint8), o_mut.view(np.uint8))
    dec, _, _ = ctx.decompress_ex(g_stream, dtype)
    o_dec, _ = oracle.decompress(o_stream, dtype)
    assert np.array_equal(dec.reshape(shape).view(np.uint8), o_dec.reshape(shape).view(np.uint8))
    bound = eb if mode == 0 else eb * float(data.max() - data.min())
    assert np.abs(dec.reshape(shape).astype(np.float64) - data.astype(np.float64)).max() <= bound


# Direction-0 3D/2D fields run each level's y and x passes as one launch over tiles of whole rows (ty
# rows from the row width and element size). Shapes that reach the tiling's edges:
TILE_CASES = [
    # shape, dtype, interp_algo
    ((20, 66, 100), np.float32, 0),   # linear even ny: the last tile takes the end row with row ny-3 (34 rows)
    ((20, 34, 100), np.float32, 0),   # ... one tile of ty + 2 rows
    ((18, 70, 100), np.float32, 0),   # linear even ny, last tile of 6 rows
    ((16, 66, 200), np.float64, 0),   # float64 tiles
    ((30, 54, 600), np.float64, 0),   # 4-row tiles; the last merged to 6
    ((66, 1500), np.float32, 0),      # 2D: level-1 rows too wide for LDS (separate passes), coarser levels tiled
    ((98, 700), np.float32, 1),       # 2D cubic, 12-row tiles
    ((30, 40, 1), np.float32, 1),     # one column
    ((9, 1, 50), np.float32, 0),      # one row a plane
    ((1, 70, 64), np.float32, 1),     # one plane
]


@pytest.mark.parametrize("case", TILE_CASES, ids=[f"{c[0]}-{np.dtype(c[1]).name}-a{c[2]}" for c in TILE_CASES])
def test_level_tile_edges(ctx, oracle, case):
    shape, dtype, ialgo = case
    data = synth.sinusoids(shape, seed=11, dtype=dtype)
    kw = dict(cmpr_algo=2, eb_mode=0, abs_eb=1e-3, interp_algo=ialgo, interp_direction=0, interp_anchor_stride=-1)
    c, oc = _cfg_pair(oracle, shape, **kw)
    o_stream, o_mut, o_pre = oracle.compress(data, oc, want_pre=True)
    g_data = data.copy()
    g_stream, _, g_pre = ctx.compress_ex(g_data, c, want_prezstd=True)
    assert g_pre == o_pre
    assert bytes(g_stream) == o_stream
    assert np.array_equal(g_data.view(np.uint8), o_mut.view(np.uint8))
    dec, _, _ = ctx.decompress_ex(g_stream, dtype)
    o_dec, _ = oracle.decompress(o_stream, dtype)
    assert np.array_equal(dec.reshape(shape).view(np.uint8), o_dec.reshape(shape).view(np.uint8))


def _window_cases():
    from test_oracle_traversal import WINDOW
    return [(s, d, c) for s, d in WINDOW for c in (0, 1)]


@pytest.mark.parametrize("shape,direction,cubic", _window_cases())
@pytest.mark.parametrize("simd", [1, 0])
def test_compare_class_window_fields(ctx, oracle, shape, direction, cubic, simd):
    """Fields whose first level-1 targets all sit at error exactly float(eb): the vector compare class
    accepts them and the scalar one escapes them, so any misplaced class changes the stream."""
    from test_oracle_traversal import _window_field
    data = _window_field(shape, direction)
    kw = dict(cmpr_algo=2, eb_mode=0, abs_eb=1e-3, interp_algo=cubic, interp_direction=direction,
              interp_anchor_stride=4)
    c, oc = _cfg_pair(oracle, shape, **kw)
    oracle.L.fo_set_simd(simd)
    try:
        o_stream, o_mut, o_pre = oracle.compress(data, oc, want_pre=True)
    finally:
        oracle.L.fo_set_simd(1)
    g = data.copy()
    g_stream, _, g_pre = ctx.compress_ex(g, c, simd=simd, want_prezstd=True)
    assert g_pre == o_pre
    assert bytes(g_stream) == o_stream
    assert np.array_equal(g.view(np.uint32), o_mut.view(np.uint32))
