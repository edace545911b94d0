"""The oracle against golden vectors written by the reference's own sources (tests/golden/make_golden.py
over oracle/_ref) and against the reference's known-answer tests. Bit-exact throughout."""
import hashlib
import json
import os

import numpy as np
import pytest

from golden_inputs import codes_from_recipe
from oracle_lib import Oracle, ref_or_none

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _split(flat, lens):
    offs = np.concatenate([[0], np.cumsum(lens)])
    return [flat[offs[i]:offs[i + 1]] for i in range(len(lens))]


def test_fse_blocks_match_reference(orc):
    z = np.load(os.path.join(GOLDEN, "fse.npz"))
    srcs = _split(z["src"], z["lens"])
    outs = _split(z["out"], z["out_lens"])
    assert len(srcs) >= 40
    for src, ret, out in zip(srcs, z["ret"], outs):
        r, o = orc.fse_compress(src)
        assert r == ret
        assert o == out.tobytes()
        if r > 1:
            assert np.array_equal(orc.fse_decompress(o, src.size), src)


def test_rlefse_small_match_reference(orc):
    z = np.load(os.path.join(GOLDEN, "rlefse.npz"))
    recipes = json.loads(str(z["recipes"]))
    codes = _split(z["codes"], z["code_lens"])
    encs = _split(z["enc"], z["enc_lens"])
    methods = set()
    for r, c, e in zip(recipes, codes, encs):
        assert np.array_equal(codes_from_recipe(r), c), r
        got = orc.rle_encode(c)
        assert got == e.tobytes(), r
        assert np.array_equal(orc.rle_decode(got, c.size), c), r
        methods.add((e[0] & 0x7F, e[0] >> 7))
    # both entropy methods, with and without the delta transform
    assert {(0, 0), (1, 0), (1, 1)} <= methods


@pytest.mark.parametrize("idx", range(5))
def test_rlefse_large_match_reference(orc, idx):
    with open(os.path.join(GOLDEN, "rlefse_big.json")) as f:
        rec = json.load(f)[idx]
    c = codes_from_recipe(rec)
    assert hashlib.sha256(c.tobytes()).hexdigest() == rec["codes_sha256"]
    e = orc.rle_encode(c)
    assert len(e) == rec["enc_len"]
    assert hashlib.sha256(e).hexdigest() == rec["enc_sha256"]
    assert np.array_equal(orc.rle_decode(e, c.size), c)


@pytest.mark.parametrize("tag", ["f32", "f64"])
def test_quantizer_matches_reference(orc, tag):
    import ctypes as C
    z = np.load(os.path.join(GOLDEN, "quantizer.npz"))
    dt = np.float32 if tag == "f32" else np.float64
    fn = orc.L.fo_quantize_point_f32 if tag == "f32" else orc.L.fo_quantize_point_f64
    ct = C.c_float if tag == "f32" else C.c_double
    ori, pred, eb = z[f"{tag}_ori"], z[f"{tag}_pred"], z[f"{tag}_eb"]
    code, dec, esc = z[f"{tag}_code"], z[f"{tag}_dec"], z[f"{tag}_esc"]
    out, e = ct(0), C.c_int(0)
    got_c = np.zeros(ori.size, np.int32)
    got_d = np.zeros(ori.size, dt)
    got_e = np.zeros(ori.size, np.int8)
    for i in range(ori.size):
        got_c[i] = fn(ori[i], pred[i], eb[i], 0, C.byref(out), C.byref(e))  # scalar compare class
        got_d[i], got_e[i] = out.value, e.value
    assert np.array_equal(got_c, code)
    assert np.array_equal(got_e, esc)
    ui = np.uint32 if tag == "f32" else np.uint64
    assert np.array_equal(got_d.view(ui), dec.view(ui))
    # the fixture spans ordinary codes, escapes and the out-of-range/NaN/Inf branch
    assert (code == -32768).sum() > 50 and (np.abs(code) > 1000).sum() > 100


@pytest.mark.parametrize("tag", ["f32", "f64"])
def test_recover_matches_reference(tag):
    z = np.load(os.path.join(GOLDEN, "quantizer.npz"))
    dt = np.float32 if tag == "f32" else np.float64
    pred, eb, code, rec = z[f"{tag}_pred"], z[f"{tag}_eb"], z[f"{tag}_code"], z[f"{tag}_recover"]
    # recover(pred, q) = (T) fma(q, 2 eb, pred): one rounding of the exact q*2eb + pred to double
    # (Fraction arithmetic is exact and its float() rounds to nearest even), then the cast to T
    from fractions import Fraction
    ok = (code != -32768) & np.isfinite(pred)
    got = np.array([dt(float(Fraction(int(q)) * Fraction(2.0 * float(e)) + Fraction(float(p)))) if o else 0
                    for q, e, p, o in zip(code, eb, pred, ok)], dt)
    ui = np.uint32 if tag == "f32" else np.uint64
    assert np.array_equal(got[ok].view(ui), rec[ok].view(ui))


@pytest.mark.parametrize("tag", ["f32", "f64"])
def test_interpolators_match_reference(orc, tag):
    z = np.load(os.path.join(GOLDEN, "interp.npz"))
    args, res = z[f"{tag}_args"], z[f"{tag}_out"]
    fn = orc.L.fo_interp_scalar_f32 if tag == "f32" else orc.L.fo_interp_scalar_f64
    dt = np.float32 if tag == "f32" else np.float64
    ui = np.uint32 if tag == "f32" else np.uint64
    for k in range(5):
        got = np.array([fn(k, *map(float, a)) for a in args], dt)
        assert np.array_equal(got.view(ui), res[k].view(ui)), k


def test_interpolator_kats(orc):
    # tools/test/test_interpolators.cpp:67-70
    assert orc.L.fo_interp_scalar_f64(0, 10.0, 30.0, 0.0, 0.0) == 20.0
    assert orc.L.fo_interp_scalar_f64(4, 0.0, 20.0, 40.0, 60.0) == 30.0
    assert orc.L.fo_interp_scalar_f32(0, 10.0, 30.0, 0.0, 0.0) == 20.0
    assert orc.L.fo_interp_scalar_f32(4, 0.0, 20.0, 40.0, 60.0) == 30.0
    # the one-sided forms reproduce polynomials they are exact for: linear1 extrapolates a line,
    # quad_1 / quad_2 interpolate a parabola at the half-step
    x = lambda t: 3.0 * t + 1.0  # noqa: E731
    assert orc.L.fo_interp_scalar_f64(1, x(0), x(2), 0, 0) == x(3)
    p = lambda t: 0.5 * t * t - 2.0 * t + 7.0  # noqa: E731
    assert orc.L.fo_interp_scalar_f64(2, p(0), p(2), p(4), 0) == p(1)
    assert orc.L.fo_interp_scalar_f64(3, p(0), p(2), p(4), 0) == p(3)
    c = lambda t: t ** 3 - 4 * t * t + t - 2  # noqa: E731
    assert orc.L.fo_interp_scalar_f64(4, c(0), c(2), c(4), c(6)) == c(3)


def test_oracle_matches_reference_build_randomized(orc):
    """Wider randomized sweep when the reference-source build is present in this checkout."""
    ref = ref_or_none()
    if ref is None:
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(99)
    for i in range(25):
        n = int(rng.integers(1, 300_000))
        kind = ["sparse", "laplace", "escapes", "runs"][i % 4]
        c = codes_from_recipe({"kind": kind, "n": n, "seed": 1000 + i, "scale": int(rng.integers(1, 6))})
        assert orc.rle_encode(c) == ref.rle_encode(c)
    for i in range(200):
        n = int(rng.integers(1, 65537))
        src = np.minimum(rng.geometric(rng.uniform(0.01, 0.9), n) - 1, 255).astype(np.uint8)
        assert orc.fse_compress(src) == ref.fse_compress(src)
