"""fz_compress_device_noalloc through the C-ABI against fz_compress_device.

The allocation-free entry point runs the same pipeline with its workspace in the library's own
device array. Two or more calls in a row (the later ones on smaller fields, so the array holds the
previous call's contents) must give the same pre-zstd bytes and the same reconstructed field as
fz_compress_device, and the oracle's bytes. Sizes beyond the array and ALGO_INTERP_LORENZO are
refused before any GPU work.
"""
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


def _cfg(shape, mode, eb):
    import falcons_amd
    c = falcons_amd.falconsConfig(*shape)
    c.cmprAlgo = falcons_amd.falconsAlgorithm.INTERP
    c.errorBoundMode = mode
    if mode == 0:
        c.absErrorBound = eb
    else:
        c.relErrorBound = eb
    return c


def _both(ctx, data, mode, eb):
    """(pre-zstd bytes, reconstructed field) from fz_compress_device_noalloc and fz_compress_device."""
    out = []
    for fn in (ctx.compress_device_noalloc, ctx.compress_device):
        cfg = _cfg(data.shape, mode, eb)
        d = ctx.alloc(data.nbytes)
        p = ctx.alloc(ctx.prezstd_bound(cfg, data.dtype))
        d.upload(data)
        n, _ = fn(cfg, data.dtype, d, p)
        out.append((p.download(n).tobytes(), d.download(data.nbytes).view(data.dtype)))
        d.free()
        p.free()
    return out


CASES = [
    # (shape, dtype, error bound mode, bound): each call after the first is on a smaller field
    ((120, 130, 140), np.float32, 1, 1e-3),
    ((60, 70, 80), np.float32, 1, 1e-3),
    ((50, 40, 30), np.float64, 0, 1e-6),
    ((33, 17, 9), np.float32, 1, 1e-2),
]


def test_noalloc_matches_device_entry_point_calls_in_a_row(ctx):
    orc = Oracle()
    orc.L.fo_set_simd(1)
    for k, (shape, dt, mode, eb) in enumerate(CASES):
        print(f"noalloc case {k}: {shape} {np.dtype(dt).name} mode {mode} eb {eb}", flush=True)
        data = synth.power_law(shape, -3.0, seed=30 + k, dtype=dt)
        (na_pre, na_field), (dv_pre, dv_field) = _both(ctx, data, mode, eb)
        assert na_pre == dv_pre, f"case {k}: pre-zstd bytes differ from fz_compress_device"
        assert np.array_equal(na_field.view(np.uint8), dv_field.view(np.uint8)), f"case {k}: field differs"
        oc = orc.config(shape, cmpr_algo=2, eb_mode=mode, abs_eb=eb, rel_eb=eb)
        _, o_mut, o_pre = orc.compress(data, oc, want_pre=True)
        assert na_pre == o_pre, f"case {k}: pre-zstd bytes differ from the oracle"
        assert np.array_equal(na_field.view(np.uint8), o_mut.reshape(-1).view(np.uint8))


def test_noalloc_largest_field(ctx):
    """176^3 floats: close to the largest field the scratch array serves (5,592,405 elements)."""
    shape = (176, 176, 176)
    data = synth.power_law(shape, -3.0, seed=40, dtype=np.float32)
    (na_pre, na_field), (dv_pre, dv_field) = _both(ctx, data, 1, 1e-4)
    assert na_pre == dv_pre
    assert np.array_equal(na_field.view(np.uint8), dv_field.view(np.uint8))


def test_noalloc_refuses_beyond_limits(ctx):
    import falcons_amd
    d = ctx.alloc(1 << 12)
    p = ctx.alloc(1 << 12)
    try:
        big = _cfg((200, 200, 200), 1, 1e-3)  # 8 M floats: beyond the array
        with pytest.raises(ValueError, match="scratch array"):
            ctx.compress_device_noalloc(big, np.float32, d, p)
        lor = _cfg((64, 64, 64), 0, 1e-3)
        lor.cmprAlgo = falcons_amd.falconsAlgorithm.INTERP_LORENZO
        with pytest.raises(Exception, match="libzstd"):
            ctx.compress_device_noalloc(lor, np.float32, d, p)
    finally:
        d.free()
        p.free()
