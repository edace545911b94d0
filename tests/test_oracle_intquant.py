"""The oracle's integer quantizer and interpolators against the reference's own sources.

LinearQuantizer<T> (quantize_integral / recover_integral, LinearQuantizer.hpp:190-248) and the integer
branches of the interpolators (Interpolators.hpp:33-117) compile from their own headers into
oracle/_ref; tests/golden/intquant.npz holds their outputs for the eight integer element types on
near and far pairs, several bounds and the types' limits (distances that overflow the narrower
arithmetic, reconstructions that leave the range, clamped predictions). The oracle, whose generic
traversal uses exactly these functions, must return the same codes, values, escape decisions,
overflow refusals and predictions.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle_lib import INT_CTYPES, REF_SO, Oracle, Ref

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "intquant.npz"))
TAGS = list(INT_CTYPES)


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _fns(orc, tag):
    ct = INT_CTYPES[tag]
    q = getattr(orc.L, f"fo_quantize_point_{tag}")
    q.restype, q.argtypes = C.c_int, [ct, ct, C.c_double, C.POINTER(ct), C.POINTER(C.c_int)]
    r = getattr(orc.L, f"fo_recover_point_{tag}")
    r.restype, r.argtypes = C.c_int, [ct, C.c_int, C.c_double, C.POINTER(ct)]
    i = getattr(orc.L, f"fo_interp_point_{tag}")
    i.restype, i.argtypes = ct, [C.c_int, ct, ct, ct, ct]
    return ct, q, r, i


@pytest.mark.parametrize("tag", TAGS)
def test_integer_quantizer_matches_reference(orc, tag):
    ct, q, _, _ = _fns(orc, tag)
    ori, pred, eb = G[f"{tag}_ori"], G[f"{tag}_pred"], G[f"{tag}_eb"]
    code, dec, esc = G[f"{tag}_code"], G[f"{tag}_dec"], G[f"{tag}_esc"]
    bad = []
    for i in range(ori.size):
        out, un = ct(0), C.c_int(0)
        c = q(ct(int(ori[i])), ct(int(pred[i])), float(eb[i]), C.byref(out), C.byref(un))
        if c != code[i] or out.value != int(dec[i]) or un.value != esc[i]:
            bad.append(i)
    assert not bad, f"{len(bad)} points differ, first {bad[:5]}"
    assert esc.sum() > 0 and (esc == 0).sum() > 0  # both outcomes are exercised


@pytest.mark.parametrize("tag", TAGS)
def test_integer_recover_matches_reference(orc, tag):
    ct, _, r, _ = _fns(orc, tag)
    pred, code, eb, ok, val = (G[f"{tag}_{k}"] for k in ("rpred", "rcode", "reb", "rok", "rout"))
    bad = []
    for i in range(pred.size):
        out = ct(0)
        got = r(ct(int(pred[i])), int(code[i]), float(eb[i]), C.byref(out))
        if got != ok[i] or (got and out.value != int(val[i])):
            bad.append(i)
    assert not bad, f"{len(bad)} reconstructions differ, first {bad[:5]}"
    assert ok.sum() > 0 and (ok == 0).sum() > 0  # in-range and overflowing reconstructions


@pytest.mark.parametrize("tag", TAGS)
def test_integer_interpolators_match_reference(orc, tag):
    ct, _, _, fn = _fns(orc, tag)
    args, want = G[f"{tag}_iargs"], G[f"{tag}_iout"]
    for k in range(6):
        got = [fn(k, *(ct(int(v)) for v in args[i])) for i in range(args.shape[0])]
        assert got == [int(v) for v in want[k]], f"interpolator {k}"


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("tag", TAGS)
def test_integer_quantizer_live_random(orc, tag):
    """Fresh random points against the reference build (beyond the stored vectors)."""
    ref = Ref()
    ct, q, _, _ = _fns(orc, tag)
    info = np.iinfo(np.dtype(INT_CTYPES[tag]))
    rng = np.random.default_rng(7)
    lo, hi = int(info.min), int(info.max)
    for _ in range(400):
        a = lo + int(rng.integers(0, 2 ** 62)) % (hi - lo + 1)
        b = min(max(a + int(rng.integers(-5000, 5000)), lo), hi)
        e = float(rng.choice([0.5, 1.0, 2.5, 77.0, 1e6]))
        out, un = ct(0), C.c_int(0)
        c = q(ct(a), ct(b), e, C.byref(out), C.byref(un))
        assert (c, out.value, un.value) == ref.quantize_int(tag, a, b, e), (a, b, e)
