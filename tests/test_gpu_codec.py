"""GPU path against the oracle on inputs that exercise each entropy-coder branch (RLE+FSE and order-1
tANS, with and without the delta transform, one and several tANS segments, escapes and NaN/Inf) and the
device-resident entry points. Every comparison is bit for bit."""
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


def quad3(shape=(64, 64, 64), steps=False):
    z = np.arange(shape[0])[:, None, None] + 0 * np.arange(shape[1])[None, :, None] + \
        ((np.arange(shape[2]) / 8.0) ** 2)[None, None, :]
    if steps:
        z = z.copy()
        z[::7] += 0.5
    return z.astype(np.float32)


def spiky(shape, dtype, seed=9):
    d = synth.power_law(shape, -3.0, seed=seed, dtype=dtype)
    rng = np.random.default_rng(seed)
    flat = d.reshape(-1)
    idx = rng.choice(flat.size, size=max(1, flat.size // 500), replace=False)
    flat[idx] = rng.uniform(-1e6, 1e6, size=idx.size).astype(dtype)
    return d


def nan_inf(shape, dtype):
    d = synth.sinusoids(shape, seed=4, dtype=dtype)
    flat = d.reshape(-1)
    flat[17] = np.nan
    flat[flat.size // 2] = np.inf
    flat[flat.size // 3] = -np.inf
    flat[-5] = np.nan
    return d


def check(ctx, oracle, data, method=None, delta=None, **kw):
    import falcons_amd
    shape = data.shape
    c = falcons_amd.falconsConfig(*shape)
    oc = oracle.config(shape)
    kw.setdefault("cmpr_algo", 2)
    for k, v in kw.items():
        setattr(c.conf, k, v)
        setattr(oc, k, v)
    print(f"case {shape} {data.dtype} {kw}", flush=True)
    o_stream, o_mut, o_pre = oracle.compress(data.copy(), oc, want_pre=True)
    g = data.copy()
    g_stream, opts, g_pre = ctx.compress_ex(g, c, want_prezstd=True)
    if method is not None:
        assert opts.method == method
    if delta is not None:
        assert opts.use_delta == delta
    assert g_pre == o_pre
    assert bytes(g_stream) == o_stream
    assert np.array_equal(g.view(np.uint8), o_mut.view(np.uint8))
    dec, _, _ = ctx.decompress_ex(g_stream, data.dtype)
    o_dec, _ = oracle.decompress(o_stream, data.dtype)
    assert np.array_equal(dec.view(np.uint8), o_dec.reshape(-1).view(np.uint8))
    # compression wrote back the reconstruction except at escapes and on the last pass of level 1
    gm, src = g.reshape(-1), data.reshape(-1)
    same = (gm.view(np.uint8).reshape(gm.size, -1) == dec.view(np.uint8).reshape(gm.size, -1)).all(1)
    kept = (gm.view(np.uint8).reshape(gm.size, -1) == src.view(np.uint8).reshape(gm.size, -1)).all(1)
    assert np.all(same | kept)
    return opts


def test_o1_tans_no_delta(ctx, oracle):
    check(ctx, oracle, synth.power_law((96, 96, 96), -3.0, seed=5), method=1, delta=0, eb_mode=1, rel_eb=1e-4)


def test_o1_tans_delta(ctx, oracle):
    check(ctx, oracle, quad3(), method=1, delta=1, eb_mode=0, abs_eb=2e-3, interp_algo=0)


def test_rle_fse_delta(ctx, oracle):
    check(ctx, oracle, quad3(), method=0, delta=1, eb_mode=0, abs_eb=2e-2, interp_algo=0)


def test_o1_tans_delta_cubic(ctx, oracle):
    check(ctx, oracle, quad3(steps=True) ** 1.5, method=1, eb_mode=0, abs_eb=1e-5, interp_algo=1)


def test_o1_tans_1d(ctx, oracle):
    x = np.arange(100000, dtype=np.float64)
    check(ctx, oracle, ((x / 1000) ** 2).astype(np.float32), method=1, eb_mode=0, abs_eb=1e-6, interp_algo=0)


def test_o1_tans_two_segments(ctx, oracle):
    # 264*256*256 - anchors > 16 Mi codes: two order-1 tANS segments, sampled histograms
    check(ctx, oracle, synth.power_law((264, 256, 256), -3.0, seed=6), method=1, eb_mode=1, rel_eb=1e-3)


def test_escapes_f32(ctx, oracle):
    opts = check(ctx, oracle, spiky((80, 72, 90), np.float32), eb_mode=0, abs_eb=1e-3)
    assert opts.nunpred > 100


def test_escapes_f64(ctx, oracle):
    opts = check(ctx, oracle, spiky((60, 70, 81), np.float64), eb_mode=0, abs_eb=1e-7, interp_algo=0,
                 interp_direction=5)
    assert opts.nunpred > 100


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nan_inf(ctx, oracle, dtype):
    check(ctx, oracle, nan_inf((40, 50, 61), dtype), eb_mode=0, abs_eb=1e-3)


def test_bound_edge_f32(ctx, oracle):
    # values far from zero with a bound near one f32 ulp: many points sit at |err| ~ eb
    d = (1000.0 + synth.sinusoids((48, 48, 48), seed=12, dtype=np.float64) * 1e-3).astype(np.float32)
    check(ctx, oracle, d, eb_mode=0, abs_eb=6.1e-5)


@pytest.mark.parametrize("bumps", [2, 3, 4, 5, 6, 7, 9, 12, 15, 20, 23, 31, 45, 64, 100, 150, 250, 400])
def test_rle_fse_short_class_streams(ctx, oracle, bumps):
    """A flat 1-D field with a few bumps: nearly every code is zero, so the RLE+FSE method codes a
    run-class and a value-class stream of about `bumps` bytes each. The FSE block encoder reads each
    state chain's symbols ahead of its steps, in rounds of four with clamped reads near symbol 0;
    these lengths reach its start, its last rounds and its tail at both parities."""
    rng = np.random.default_rng(1000 + bumps)
    d = np.zeros(6000, dtype=np.float32)
    pos = rng.choice(np.arange(8, d.size - 8), size=bumps, replace=False)
    d[pos] = (rng.choice([-1.0, 1.0], size=bumps) * 10.0 ** rng.uniform(-2, 1, size=bumps)).astype(np.float32)
    opts = check(ctx, oracle, d, eb_mode=0, abs_eb=1e-3, interp_algo=0)
    assert opts.method == 0


@pytest.mark.parametrize("density", [0.02, 0.04])
def test_rle_fse_long_class_streams(ctx, oracle, density):
    """A long 1-D field with bumps at a few percent of its points: the RLE+FSE method's run-class and
    value-class streams run to several 64 K-symbol blocks whose FSE bytes exceed the decoder's
    16 KiB LDS window, so each block decodes over several windows."""
    rng = np.random.default_rng(77)
    d = np.zeros(3_000_000, dtype=np.float32)
    k = int(d.size * density)
    pos = rng.choice(d.size, size=k, replace=False)
    d[pos] = (rng.choice([-1.0, 1.0], size=k) * 10.0 ** rng.uniform(-2, 1, size=k)).astype(np.float32)
    opts = check(ctx, oracle, d, eb_mode=0, abs_eb=1e-3, interp_algo=0)
    assert opts.method == 0


def test_device_entry_points(ctx, oracle):
    import falcons_amd
    shape = (70, 64, 96)
    data = synth.power_law(shape, -3.0, seed=8)
    cfg = falcons_amd.falconsConfig(*shape)
    cfg.cmprAlgo = falcons_amd.falconsAlgorithm.INTERP
    cfg.errorBoundMode = falcons_amd.falconsErrorBoundMode.REL
    cfg.relErrorBound = 1e-3
    host = data.copy()
    stream, _, pre = ctx.compress_ex(host, cfg.copy(), want_prezstd=True)
    want, _, _ = ctx.decompress_ex(stream, np.float32)
    d_in = ctx.alloc(data.nbytes)
    d_out = ctx.alloc(data.nbytes)
    d_pre = ctx.alloc(ctx.prezstd_bound(cfg, np.float32))
    try:
        d_in.upload(data)
        run = cfg.copy()
        n, opts = ctx.compress_device(run, np.float32, d_in, d_pre)
        h_pre = d_pre.download(n)
        assert h_pre.tobytes() == pre
        assert np.array_equal(d_in.download(data.nbytes, np.float32).view(np.uint32), host.reshape(-1).view(np.uint32))
        assert run.errorBoundMode == 0 and run.absErrorBound > 0
        for dev in (d_pre, None):
            ctx.decompress_device(run, np.float32, h_pre, dev, d_out)
            assert np.array_equal(d_out.download(data.nbytes, np.float32).view(np.uint32), want.view(np.uint32))
    finally:
        for b in (d_in, d_out, d_pre):
            b.free()


def test_device_undersized_pre_buffer(ctx):
    import falcons_amd
    shape = (40, 40, 40)
    data = synth.sinusoids(shape, seed=2)
    cfg = falcons_amd.falconsConfig(*shape)
    cfg.cmprAlgo = falcons_amd.falconsAlgorithm.INTERP
    cfg.absErrorBound = 1e-3
    d_in = ctx.alloc(data.nbytes)
    d_pre = ctx.alloc(64)
    try:
        d_in.upload(data)
        with pytest.raises(ValueError):
            ctx.compress_device(cfg.copy(), np.float32, d_in, d_pre)
    finally:
        d_in.free()
        d_pre.free()


def test_truncated_stream_raises(ctx):
    import falcons_amd
    shape = (30, 30, 30)
    data = synth.sinusoids(shape, seed=3)
    cfg = falcons_amd.falconsConfig(*shape)
    cfg.absErrorBound = 1e-3
    stream, _, _ = ctx.compress_ex(data.copy(), cfg)
    for cut in (3, 20, len(stream) // 2, len(stream) - 1):
        with pytest.raises(ValueError):
            ctx.decompress_ex(stream[:cut], np.float32)


def test_walk_with_low_context_window(ctx, oracle, monkeypatch, capfd):
    """A smooth 2-D field whose coarse-level chains do not settle: the walk decodes subsequences
    exactly with the tables of a low context window (not 11..14) in LDS. Same bytes as the oracle;
    the decoder's statistics show the path was taken."""
    import re
    monkeypatch.setenv("FALCONS_DECODE_STATS", "1")
    check(ctx, oracle, synth.power_law((4096, 4096), -3.0, seed=3), method=1, eb_mode=1, rel_eb=1e-3)
    err = capfd.readouterr().err
    chains = re.findall(r" (\d+):(\d+)/(\d+)/esc\d+/b[\d.]+/s\d+/h(\d+)", err)
    assert chains, err[-2000:]
    assert any(int(ex) > 0 and int(h) != 11 for _, ex, _, h in chains), chains


SEQ_CASES = [
    ("no-delta", lambda: synth.power_law((96, 96, 96), -3.0, seed=5), dict(eb_mode=1, rel_eb=1e-4)),
    ("delta", lambda: quad3(), dict(eb_mode=0, abs_eb=2e-3, interp_algo=0)),
    ("cubic-steps", lambda: quad3(steps=True) ** 1.5, dict(eb_mode=0, abs_eb=1e-5, interp_algo=1)),
    ("escapes", lambda: spiky((80, 72, 90), np.float32), dict(eb_mode=0, abs_eb=1e-3)),
    ("nan-inf", lambda: nan_inf((40, 50, 61), np.float64), dict(eb_mode=0, abs_eb=1e-3)),
    ("dense-escapes", lambda: synth.power_law((128, 128, 128), -3.0, seed=8), dict(eb_mode=1, rel_eb=2e-5)),
]


@pytest.mark.parametrize("case", SEQ_CASES, ids=[c[0] for c in SEQ_CASES])
def test_o1_sequential_decoder(ctx, oracle, case, monkeypatch):
    """Every order-1 tANS chain through the sequential decoder (FALCONS_O1_SEQUENTIAL=1), which the
    library otherwise keeps for chains with dense escapes: same codes as the oracle."""
    _, make, kw = case
    data = make()
    monkeypatch.setenv("FALCONS_O1_SEQUENTIAL", "1")
    check(ctx, oracle, data, **kw)


PAR_CASES = SEQ_CASES[:5] + [
    ("rel-1e-2", lambda: synth.power_law((100, 100, 100), -3.0, seed=5), dict(eb_mode=1, rel_eb=1e-2)),
    ("rel-1e-3", lambda: synth.power_law((128, 128, 128), -3.0, seed=5), dict(eb_mode=1, rel_eb=1e-3)),
    ("2d", lambda: synth.power_law((1024, 1024), -3.0, seed=6), dict(eb_mode=1, rel_eb=1e-3)),
]


@pytest.mark.parametrize("case", PAR_CASES, ids=[c[0] for c in PAR_CASES])
def test_o1_parallel_decoder_short_chains(ctx, oracle, case, monkeypatch):
    """Chains of at most 2^19 symbols go to the sequential decoder; FALCONS_O1_PARALLEL_SHORT=1 keeps
    them on the parallel one (warm-up, graphs, walk), which then decodes these small fields: same
    codes as the oracle."""
    _, make, kw = case
    data = make()
    monkeypatch.setenv("FALCONS_O1_PARALLEL_SHORT", "1")
    check(ctx, oracle, data, **kw)


def nan_placed(dtype, where):
    """A (41, 50, 61) field (tails of 2 after the last full AVX2 vector) with NaNs placed where the
    AVX2 and scalar data_range differ (Statistic.hpp:38-157)."""
    d = synth.sinusoids((41, 50, 61), seed=4, dtype=dtype)
    flat = d.reshape(-1)
    L = 8 if dtype == np.float32 else 4
    nfull = flat.size // L * L
    if where == "first":
        flat[0] = np.nan
    elif where == "last_lane0":
        flat[nfull - L] = np.nan
    elif where == "lane_history":
        flat[3] = 40.0
        flat[9 * L + 3] = np.nan
    elif where == "last_other_lane":
        flat[L + 2] = 40.0
        flat[nfull - L + 2] = np.nan
    elif where == "tail":
        flat[-1] = np.nan
    return d


@pytest.mark.parametrize("where", ["first", "last_lane0", "lane_history", "last_other_lane", "tail"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("simd", [1, 0])
def test_nan_rel_bound(ctx, oracle, where, dtype, simd):
    """REL bound on fields with NaNs: the value range, and so every code, follows the build's
    data_range (AVX2 lanes or scalar scan)."""
    import falcons_amd
    data = nan_placed(dtype, where)
    shape = data.shape
    c = falcons_amd.falconsConfig(*shape)
    oc = oracle.config(shape, cmpr_algo=2, eb_mode=1, rel_eb=1e-3)
    c.conf.cmpr_algo, c.conf.eb_mode, c.conf.rel_eb = 2, 1, 1e-3
    print(f"case nan {where} {np.dtype(dtype).name} simd {simd}", flush=True)
    oracle.L.fo_set_simd(simd)
    try:
        o_stream, o_mut, o_pre = oracle.compress(data.copy(), oc, want_pre=True)
    finally:
        oracle.L.fo_set_simd(1)
    g = data.copy()
    g_stream, _, g_pre = ctx.compress_ex(g, c, simd=simd, want_prezstd=True)
    assert g_pre == o_pre
    assert bytes(g_stream) == o_stream
    assert np.array_equal(g.view(np.uint8), o_mut.view(np.uint8))
