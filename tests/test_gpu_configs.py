"""The benchmark workloads themselves, on the GPU, against the oracle on the same bytes.

Each case compares the buffer handed to zstd, the final stream and the decompressed field with the
oracle bit for bit, at the workload's real size:
  128^3 float32 sinusoids, ABS 1e-3, ALGO_INTERP
  512^3 float32 k^-3 field, REL 1e-3 and REL 1e-4, ALGO_INTERP
  256x384x384 float64 k^-5/3 field with a radial tanh front, ABS 1e-6, ALGO_INTERP_LORENZO (the
      parameter search's candidate ratios are compared too)
  a sample of the 4096-chunk mix (the largest chunks and an even spread) through the batched entry
      points
  a sample of the serving mix through the request loop
"""
import ctypes

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


def _pair(oracle, shape, algo, mode, eb):
    import falcons_amd
    c = falcons_amd.falconsConfig(*shape)
    c.cmprAlgo = algo
    c.errorBoundMode = mode
    if mode == 0:
        c.absErrorBound = eb
    else:
        c.relErrorBound = eb
    oc = oracle.config(shape, cmpr_algo=int(algo), eb_mode=int(mode), abs_eb=eb, rel_eb=eb)
    return c, oc


def _field_case(ctx, oracle, data, algo, mode, eb):
    shape = data.shape
    c, oc = _pair(oracle, shape, algo, mode, eb)
    o_stream, o_mut, o_pre = oracle.compress(data, oc, want_pre=True)
    g = data.copy()
    g_stream, opts, g_pre = ctx.compress_ex(g, c, want_prezstd=True)
    assert len(g_pre) == len(o_pre)
    assert g_pre == o_pre, "pre-zstd buffer differs from the oracle"
    assert bytes(g_stream) == o_stream, "stream differs from the oracle"
    assert np.array_equal(g.view(np.uint8), o_mut.view(np.uint8)), "reconstructed input differs"
    del g, o_mut
    dec, _, _ = ctx.decompress_ex(g_stream, data.dtype)
    o_dec, _ = oracle.decompress(o_stream, data.dtype)
    assert np.array_equal(dec.reshape(shape).view(np.uint8), o_dec.reshape(shape).view(np.uint8))
    bound = eb if mode == 0 else eb * (float(data.max()) - float(data.min()))
    assert np.abs(dec.reshape(shape).astype(np.float64) - data.astype(np.float64)).max() <= bound
    return opts


def test_sinusoids_128_abs(ctx, oracle):
    data = synth.sinusoids((128, 128, 128), seed=1)
    _field_case(ctx, oracle, data, 2, 0, 1e-3)


@pytest.mark.parametrize("rel", [1e-3, 1e-4])
def test_k3_field_512_rel(ctx, oracle, rel):
    data = synth.power_law((512, 512, 512), -3.0, seed=2, dtype=np.float32)
    opts = _field_case(ctx, oracle, data, 2, 1, rel)
    assert opts.method == 1  # order-1 tANS, as the reference picks for this field


def test_front_f64_lorenzo(ctx, oracle):
    data = synth.front((256, 384, 384), seed=3)
    _field_case(ctx, oracle, data, 1, 0, 1e-6)
    g_ratios = ctx.tune_report()
    r = np.zeros(16)
    n = oracle.L.fo_tune_report(r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 16)
    assert n == 4 and len(g_ratios) == n  # {linear, cubic} x {direction 0, 5}
    assert list(g_ratios) == list(r[:n])


def _chunk_sample(count=4096, take=64, seed=4):
    sides, rels = synth.chunk_mix(count, seed=seed)
    vol = sides.prod(axis=1)
    largest = list(np.argsort(-vol, kind="stable")[:8])
    spread = [i for i in np.linspace(0, count - 1, take).astype(int) if i not in largest][: take - len(largest)]
    return sides, rels, sorted(int(i) for i in largest + spread)


def test_chunk_mix_sample_batched(ctx, oracle):
    import falcons_amd
    sides, rels, idx = _chunk_sample()
    items, bufs, expect = [], [], []
    for i in idx:
        f = synth.chunk_field(sides, i)
        c, oc = _pair(oracle, f.shape, falcons_amd.falconsAlgorithm.INTERP, 1, float(rels[i]))
        o_stream, o_mut, o_pre = oracle.compress(f, oc, want_pre=True)
        o_dec, _ = oracle.decompress(o_stream, np.float32)
        d, p = ctx.alloc(f.nbytes), ctx.alloc(ctx.prezstd_bound(c, np.float32))
        d.upload(f)
        items.append((c, np.float32, d, p))
        bufs.append((d, p))
        expect.append((o_pre, o_mut, o_dec))
    ctx.set_workers(24)
    lens, status, _ = ctx.compress_batch(items)
    assert status == [0] * len(items)
    dec_items, outs = [], []
    for (c, _, d, p), n, (o_pre, o_mut, o_dec) in zip(items, lens, expect):
        assert n == len(o_pre)
        assert p.download(n).tobytes() == o_pre, f"chunk {c.dims}: pre-zstd differs from the oracle"
        assert np.array_equal(d.download(o_mut.nbytes).view(np.uint8), o_mut.reshape(-1).view(np.uint8))
        o = ctx.alloc(o_mut.nbytes)
        outs.append(o)
        dec_items.append((c, np.float32, np.frombuffer(o_pre, dtype=np.uint8), p, o))
    status, _ = ctx.decompress_batch(dec_items)
    assert status == [0] * len(dec_items)
    for o, (o_pre, o_mut, o_dec) in zip(outs, expect):
        assert np.array_equal(o.download(o_dec.nbytes).view(np.uint8), o_dec.reshape(-1).view(np.uint8))
    for b in [x for pair in bufs for x in pair] + outs:
        b.free()


def test_serving_mix_sample(oracle):
    """The first requests of the serving mix, submitted at once so that they overlap on the device:
    every compress result is the oracle's stream, every decompress (of the oracle's stream of the
    same field) the oracle's field."""
    import falcons_amd
    reqs = synth.serving_mix(40)[:16]
    base = synth.serving_base(sorted({r["cls"] for r in reqs}))
    work = []
    for k, r in enumerate(reqs):
        f = synth.serving_field(base, r)
        c, oc = _pair(oracle, f.shape, falcons_amd.falconsAlgorithm.INTERP, r["mode"], r["eb"])
        o_stream, _ = oracle.compress(f, oc)
        if r["op"] == 0:
            work.append((k, "c", f, c, o_stream))
        else:
            o_dec, _ = oracle.decompress(o_stream, f.dtype)
            work.append((k, "d", f, o_stream, o_dec.reshape(-1)))
    srv = falcons_amd.Server(0, slots=8)
    try:
        for k, kind, f, a, b in work:
            if kind == "c":
                srv.submit_compress(k, f, a)
            else:
                srv.submit_decompress(k, np.frombuffer(a, dtype=np.uint8), f.dtype, f.size)
        got = {}
        while len(got) < len(work):
            res = srv.poll(timeout_ms=120000)
            assert res is not None, "server stalled"
            assert res["status"] == 0, res["msg"]
            got[res["tag"]] = res["out"]
        for k, kind, f, a, b in work:
            if kind == "c":
                assert got[k].tobytes() == b, f"request {k}: stream differs from the oracle"
            else:
                assert np.array_equal(got[k].view(np.uint8), b.view(np.uint8)), f"request {k}: field differs"
        assert srv.stats()["max_in_flight"] > 1
    finally:
        srv.close()
