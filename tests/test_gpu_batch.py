"""Batched and serving forms against the plain per-call path: every item of a batch, and every request
served in a mix, must give the bytes (pre-zstd buffer, stream) and the field that a plain call gives
alone on the same input. Bit for bit."""
import numpy as np
import pytest

from falcons_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import falcons_amd
    c = falcons_amd.Context(0)
    yield c
    c.close()


def chunk_set(count, seed):
    """Uneven 3D chunks (sides log-uniform in [17, 70]) of k^-3 fields, f32 and f64, REL and ABS bounds."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        shape = tuple(int(v) for v in np.floor(np.exp(rng.uniform(np.log(17), np.log(70), 3))))
        dtype = np.float64 if i % 5 == 4 else np.float32
        data = synth.power_law(shape, -3.0, seed=seed * 1000 + i, dtype=dtype)
        if i % 3 == 2:
            data.reshape(-1)[:: 97] *= 50  # spikes: escapes
        mode, eb = (1, [1e-2, 1e-3, 1e-4][i % 3]) if i % 2 == 0 else (0, 10.0 ** -(2 + i % 3))
        out.append((data, dtype, mode, eb))
    return out


def config_for(data, mode, eb):
    import falcons_amd
    c = falcons_amd.falconsConfig(*data.shape)
    c.cmprAlgo = falcons_amd.falconsAlgorithm.INTERP
    c.errorBoundMode = mode
    if mode == 1:
        c.relErrorBound = eb
    else:
        c.absErrorBound = eb
    return c


def test_batch_matches_per_call(ctx):
    chunks = chunk_set(24, seed=41)
    per_call, bufs, items = [], [], []
    for data, dtype, mode, eb in chunks:
        cfg = config_for(data, mode, eb)
        bound = ctx.prezstd_bound(cfg, dtype)
        d1, p1 = ctx.alloc(data.nbytes), ctx.alloc(bound)
        d1.upload(data)
        n1, _ = ctx.compress_device(cfg.copy(), dtype, d1, p1)
        per_call.append((p1.download(n1), d1.download(data.nbytes, dtype)))
        d2, p2 = ctx.alloc(data.nbytes), ctx.alloc(bound)
        d2.upload(data)
        bufs.append((d2, p2))
        items.append((cfg.copy(), dtype, d2, p2))
        d1.free()
        p1.free()
    ctx.set_workers(6)
    lens, status, wall = ctx.compress_batch(items)
    assert status == [0] * len(items) and wall > 0
    for (pre, rec), n, (d2, p2), (_, dtype, _, _) in zip(per_call, lens, bufs, items):
        assert n == pre.size
        assert np.array_equal(p2.download(n), pre)
        assert np.array_equal(d2.download(rec.nbytes, dtype).view(np.uint8), rec.view(np.uint8))
    # decompression of the same buffers, batched, against per-call
    dec_items, outs = [], []
    for (pre, rec), (cfg, dtype, d2, p2) in zip(per_call, items):
        o = ctx.alloc(rec.nbytes)
        outs.append(o)
        dec_items.append((cfg, dtype, pre, p2 if len(outs) % 2 else None, o))
    status, _ = ctx.decompress_batch(dec_items)
    assert status == [0] * len(dec_items)
    for (pre, rec), (cfg, dtype, _, _, o) in zip(per_call, dec_items):
        ref = ctx.alloc(rec.nbytes)
        ctx.decompress_device(cfg, dtype, pre, None, ref)
        got = o.download(rec.nbytes, dtype)
        assert np.array_equal(got.view(np.uint8), ref.download(rec.nbytes, dtype).view(np.uint8))
        ref.free()
    for d2, p2 in bufs:
        d2.free()
        p2.free()
    for o in outs:
        o.free()


def test_decompress_batch_shared_launch_mixed(ctx):
    """The batched decompress decodes the sequential chains of every item in one launch between the
    items' two halves. Items whose chains are long enough for the parallel decoder (160^3, 2^19+
    symbols a chain) run it inside their first half; every item's field must equal a plain call's."""
    shapes = [((40, 50, 60), np.float32, 1, 1e-3), ((160, 160, 160), np.float32, 1, 1e-3),
              ((1400, 1500), np.float32, 0, 1e-3), ((33, 34, 35), np.float64, 1, 1e-5),
              ((150, 170, 160), np.float32, 1, 1e-2), ((20, 21, 22), np.float32, 0, 1e-4)]
    items, refs, bufs = [], [], []
    for k, (shape, dtype, mode, eb) in enumerate(shapes):
        data = synth.power_law(shape, -3.0, seed=70 + k, dtype=dtype)
        cfg = config_for(data, mode, eb)
        d, p = ctx.alloc(data.nbytes), ctx.alloc(ctx.prezstd_bound(cfg, dtype))
        d.upload(data)
        n, _ = ctx.compress_device(cfg, dtype, d, p)
        pre = p.download(n)
        ref = ctx.alloc(data.nbytes)
        ctx.decompress_device(cfg, dtype, pre, None, ref)
        refs.append(ref.download(data.nbytes, dtype))
        o = ctx.alloc(data.nbytes)
        items.append((cfg, dtype, pre, p if k % 2 else None, o))
        bufs += [d, p, ref, o]
    ctx.set_workers(4)
    status, _ = ctx.decompress_batch(items)
    assert status == [0] * len(items)
    for (cfg, dtype, _, _, o), ref in zip(items, refs):
        assert np.array_equal(o.download(ref.nbytes, dtype).view(np.uint8), ref.view(np.uint8))
    for b in bufs:
        b.free()


@pytest.mark.parametrize("slice_bytes", [1 << 20, 12 << 20])
def test_decompress_batch_slice_fallback(ctx, monkeypatch, slice_bytes):
    """Items whose first part does not fit their workspace slice are decompressed by a plain call
    instead (1 MiB: every item; 12 MiB: the larger ones), beside items that go through the shared
    launch: every field equals a plain call's."""
    items, refs, bufs = [], [], []
    for k, shape in enumerate([(24, 30, 36), (96, 100, 110), (40, 41, 42), (128, 120, 100), (20, 20, 20)]):
        data = synth.power_law(shape, -3.0, seed=90 + k, dtype=np.float32)
        cfg = config_for(data, 1, 1e-3)
        d, p = ctx.alloc(data.nbytes), ctx.alloc(ctx.prezstd_bound(cfg, np.float32))
        d.upload(data)
        n, _ = ctx.compress_device(cfg, np.float32, d, p)
        pre = p.download(n)
        ref = ctx.alloc(data.nbytes)
        ctx.decompress_device(cfg, np.float32, pre, None, ref)
        refs.append(ref.download(data.nbytes, np.float32))
        o = ctx.alloc(data.nbytes)
        items.append((cfg, np.float32, pre, p, o))
        bufs += [d, p, ref, o]
    monkeypatch.setenv("FALCONS_BATCH_SLICE_BYTES", str(slice_bytes))
    ctx.set_workers(3)
    status, _ = ctx.decompress_batch(items)
    assert status == [0] * len(items)
    for (_, dtype, _, _, o), ref in zip(items, refs):
        assert np.array_equal(o.download(ref.nbytes, dtype).view(np.uint8), ref.view(np.uint8))
    for b in bufs:
        b.free()


def test_batch_reports_item_errors(ctx):
    import falcons_amd
    data = synth.power_law((20, 21, 22), -3.0, seed=3, dtype=np.float32)
    cfg = config_for(data, 0, 1e-3)
    d, p, small = ctx.alloc(data.nbytes), ctx.alloc(ctx.prezstd_bound(cfg, np.float32)), ctx.alloc(64)
    d.upload(data)
    d2 = ctx.alloc(data.nbytes)
    d2.upload(data)
    with pytest.raises(ValueError):
        ctx.compress_batch([(cfg.copy(), np.float32, d, p), (cfg.copy(), np.float32, d2, small)])
    for b in (d, p, small, d2):
        b.free()


def serve_inputs(seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(20):
        if i % 4 == 3:
            shape = (int(rng.integers(200, 700)), int(rng.integers(200, 700)))
            dtype = np.float32
        else:
            shape = tuple(int(v) for v in rng.integers(24, 90, 3))
            dtype = np.float64 if i % 5 == 0 else np.float32
        data = synth.power_law(shape, -3.0, seed=seed * 100 + i, dtype=dtype)
        mode = int(i % 2)
        eb = float(10 ** rng.uniform(-5, -2))
        reqs.append((data, dtype, mode, eb))
    return reqs


@pytest.mark.parametrize("budget_mb", [0, 400])
def test_server_mix_matches_alone(ctx, budget_mb):
    import falcons_amd
    reqs = serve_inputs(7)
    alone = []
    for data, dtype, mode, eb in reqs:
        cfg = config_for(data, mode, eb)
        stream, _, _ = ctx.compress_ex(data.copy(), cfg.copy())
        alone.append(np.asarray(stream, dtype=np.uint8).copy())
    srv = falcons_amd.Server(0, slots=6, hbm_budget=budget_mb << 20)
    try:
        expect = {}
        for k, ((data, dtype, mode, eb), stream) in enumerate(zip(reqs, alone)):
            # interleave: a compress of field k and a decompress of stream k
            srv.submit_compress(2 * k, data, config_for(data, mode, eb))
            srv.submit_decompress(2 * k + 1, stream, dtype, data.size)
            expect[2 * k] = stream
            ref, _, _ = ctx.decompress_ex(stream, dtype)
            expect[2 * k + 1] = ref
        got = {}
        while len(got) < len(expect):
            r = srv.poll(timeout_ms=60000)
            assert r is not None, "server stalled"
            assert r["status"] == 0, r["msg"]
            assert r["t_submit_ms"] <= r["t_start_ms"] <= r["t_done_ms"]
            got[r["tag"]] = r["out"]
        for tag, ref in expect.items():
            out = got[tag]
            if tag % 2 == 0:
                assert np.array_equal(out, ref), f"request {tag}"
            else:
                assert np.array_equal(out.view(np.uint8), ref.view(np.uint8)), f"request {tag}"
        st = srv.stats()
        assert st["completed"] == len(expect) and st["failed"] == 0
        assert st["max_held_bytes"] <= st["budget_bytes"]
        if budget_mb == 0:
            assert st["max_in_flight"] > 1
    finally:
        srv.close()


def test_server_rejects_oversized(ctx):
    import falcons_amd
    srv = falcons_amd.Server(0, slots=2, hbm_budget=1 << 20)
    try:
        data = synth.power_law((64, 64, 64), -3.0, seed=1, dtype=np.float32)
        with pytest.raises(ValueError):
            srv.submit_compress(1, data, config_for(data, 0, 1e-3))
    finally:
        srv.close()
