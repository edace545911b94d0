"""Writes the golden vectors under tests/golden/ from the reference's own sources.

The reference's entropy coder (include/Falcons/encoder/RleFseEncoder.hpp with the vendored
zfse/*.c), its scalar LinearQuantizer (include/Falcons/quantizer/LinearQuantizer.hpp) and its
interpolators (include/Falcons/utils/Interpolators.hpp) compile from their own files; oracle/Makefile
builds them in place into oracle/_ref/libfalcons_ref.so (`make -C oracle ref`). This script calls
that build on seeded inputs and stores inputs and outputs:

  fse.npz        FSE_compress (zfse/fse_compress.c:632-694) on byte blocks: return code + bytes
  rlefse.npz     RleFseEncoder save+encode on int16 code arrays: encoded bytes
  rlefse_big.json  the same for arrays too large to store: integer-only generator recipe, length,
                 sha256 of the encoded bytes (regenerated bit-exactly by codes_from_recipe)
  quantizer.npz  LinearQuantizer::quantize_and_overwrite and recover (f32 and f64)
  interp.npz     interp_linear / linear1 / quad_1 / quad_2 / cubic (f32 and f64)
  intquant.npz   LinearQuantizer<T>::quantize_and_overwrite / recover_pred and the interpolators for
                 the eight integer element types (quantize_integral / recover_integral, wide-integer
                 predictions; LinearQuantizer.hpp:190-248, Interpolators.hpp:33-117)
  sampler.json   profiling_block + sampleBlocks (utils/Sample.hpp:5-344) of the ALGO_INTERP_LORENZO
                 parameter search on integer-generated fields (recipes in golden_inputs.sampler_field):
                 the profiling branch, the profiled block starts and the sha256 of the sampled
                 blocks, at the sample block side the search picks for each field

Run from the repository root: python tests/golden/make_golden.py [--only sampler|intquant]
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from golden_inputs import codes_from_recipe, sampler_field  # noqa: E402
from oracle_lib import Ref  # noqa: E402

BIG_RECIPES = [
    # > 2^20 codes: sampled order-1 histograms; > 16 Mi codes: two tANS segments
    {"kind": "laplace", "n": 1_300_000, "seed": 41, "scale": 3},
    {"kind": "escapes", "n": 2_500_000, "seed": 42, "scale": 5},
    {"kind": "laplace", "n": 17_000_000, "seed": 43, "scale": 2},
    {"kind": "sparse", "n": 3_000_000, "seed": 44, "scale": 1},
    {"kind": "runs", "n": 2_000_000, "seed": 45, "scale": 3},
]

SMALL_RECIPES = [
    {"kind": "zeros", "n": 100_000, "seed": 1, "scale": 0},
    {"kind": "sparse", "n": 100_000, "seed": 2, "scale": 1},
    {"kind": "sparse", "n": 5_000, "seed": 3, "scale": 2},
    {"kind": "laplace", "n": 200_000, "seed": 4, "scale": 2},
    {"kind": "laplace", "n": 76_799, "seed": 5, "scale": 4},
    {"kind": "laplace", "n": 76_800, "seed": 6, "scale": 4},
    {"kind": "escapes", "n": 150_000, "seed": 7, "scale": 6},
    {"kind": "runs", "n": 120_000, "seed": 8, "scale": 3},
    {"kind": "runs", "n": 9_000, "seed": 9, "scale": 1},
    {"kind": "laplace", "n": 1, "seed": 10, "scale": 2},
    {"kind": "laplace", "n": 2, "seed": 11, "scale": 2},
    {"kind": "laplace", "n": 31, "seed": 12, "scale": 2},
    {"kind": "escapes", "n": 8192, "seed": 13, "scale": 9},
    {"kind": "wide", "n": 90_000, "seed": 14, "scale": 0},
]


# Sampler recipes: sbs is the sample block side Falcons_compress_Interp_lorenzo picks for the shape
# (FalconsAlgoInterp.hpp:297-305), eb the absolute bound profiling_block compares ranges with.
SAMPLER_RECIPES = [
    {"kind": "walk", "shape": [160, 192, 160], "dtype": "f32", "seed": 61, "sbs": 32, "eb": 0.5},
    {"kind": "patch", "shape": [136, 160, 240], "dtype": "f64", "seed": 62, "sbs": 32, "eb": 1e-6,
     "box": [[20, 90], [0, 70], [100, 240]]},
    {"kind": "flat", "shape": [96, 100, 110], "dtype": "f32", "seed": 63, "sbs": 16, "eb": 1e-4},
    {"kind": "patch", "shape": [96, 100, 110], "dtype": "f32", "seed": 64, "sbs": 16, "eb": 1e-3,
     "box": [[40, 50], [40, 50], [40, 50]]},
    {"kind": "walk", "shape": [40, 44, 48], "dtype": "f64", "seed": 65, "sbs": 8, "eb": 0.25},
    {"kind": "walk", "shape": [1000, 1200], "dtype": "f32", "seed": 66, "sbs": 64, "eb": 2.0},
    {"kind": "patch", "shape": [700, 900], "dtype": "f64", "seed": 67, "sbs": 64, "eb": 1e-3,
     "box": [[100, 400], [300, 500]]},
    {"kind": "flat", "shape": [900, 800], "dtype": "f64", "seed": 68, "sbs": 64, "eb": 1e-3},
    {"kind": "walk", "shape": [2000000], "dtype": "f32", "seed": 69, "sbs": 4096, "eb": 0.5},
    {"kind": "patch", "shape": [4000000], "dtype": "f64", "seed": 70, "sbs": 4096, "eb": 1e-3,
     "box": [[500000, 502000]]},
    # the 256 x 384 x 384 workload's shape: 847 candidate blocks of 33^3
    {"kind": "patch", "shape": [256, 384, 384], "dtype": "f32", "seed": 71, "sbs": 32, "eb": 1e-3,
     "box": [[0, 256], [0, 384], [100, 384]]},
]


def sampler_vectors(ref):
    res = []
    for r in SAMPLER_RECIPES:
        f = sampler_field(r)
        prof, starts, blocks = ref.sample(f, r["sbs"], r["eb"])
        res.append(dict(r, profiling=int(prof), starts=starts[:, : f.ndim].tolist(),
                        nblocks=int(blocks.size // (r["sbs"] + 1) ** f.ndim),
                        blocks_sha256=hashlib.sha256(blocks.tobytes()).hexdigest()))
    return res


INT_TYPES = ["u8", "i8", "u16", "i16", "u32", "i32", "u64", "i64"]


def int_points(tag, count=1500):
    """(data, pred, eb) triples for an integer type: near pairs at several bounds, far pairs, and the
    type's limits, where distances overflow the narrower arithmetic and reconstructions leave the range."""
    dt = np.dtype({"u": "uint", "i": "int"}[tag[0]] + tag[1:])
    info = np.iinfo(dt)
    rng = np.random.default_rng(100 + INT_TYPES.index(tag))
    lo, hi = int(info.min), int(info.max)
    def draw(k):
        r = rng.integers(0, 2 ** 62, k, dtype=np.int64)
        return np.array([lo + (int(x) % (hi - lo + 1)) for x in r], dtype=object)
    base = draw(count)
    delta = rng.integers(-20000, 20000, count)
    pred = np.array([min(max(int(b) + int(d), lo), hi) for b, d in zip(base, delta)], dtype=object)
    ebs = rng.choice([0.25, 0.5, 1.0, 1.5, 2.0, 3.7, 17.0, 1000.0, 1e6, 1e-3], count)
    far_d, far_p = draw(300), draw(300)
    lim = [lo, hi, lo + 1, hi - 1, 0, 0]
    lim_d = np.array([lim[i % 6] for i in range(120)], dtype=object)
    lim_p = np.array([lim[(i // 6) % 6] for i in range(120)], dtype=object)
    data = np.concatenate([base, far_d, lim_d]).astype(dt)
    pr = np.concatenate([pred, far_p, lim_p]).astype(dt)
    eb = np.concatenate([ebs, rng.choice([0.5, 3.0, 1e9, 1e18], 300), rng.choice([0.5, 2.0, 1e15], 120)])
    return data, pr, eb.astype(np.float64)


def int_recover_args(tag, count=1200):
    dt = np.dtype({"u": "uint", "i": "int"}[tag[0]] + tag[1:])
    info = np.iinfo(dt)
    rng = np.random.default_rng(200 + INT_TYPES.index(tag))
    lo, hi = int(info.min), int(info.max)
    r = rng.integers(0, 2 ** 62, count, dtype=np.int64)
    pred = np.array([lo + (int(x) % (hi - lo + 1)) for x in r], dtype=object)
    pred[:40] = [lo, hi] * 20
    code = rng.integers(-32767, 32768, count)
    code[:40] = rng.choice([-32767, -1, 1, 32767], 40)
    eb = rng.choice([0.25, 0.5, 1.0, 3.7, 1e3, 1e12, 1e17], count)
    return pred.astype(dt), code.astype(np.int32), eb.astype(np.float64)


def int_interp_args(tag, count=800):
    dt = np.dtype({"u": "uint", "i": "int"}[tag[0]] + tag[1:])
    info = np.iinfo(dt)
    rng = np.random.default_rng(300 + INT_TYPES.index(tag))
    lo, hi = int(info.min), int(info.max)
    r = rng.integers(0, 2 ** 62, (count, 4), dtype=np.int64)
    a = np.array([[lo + (int(x) % (hi - lo + 1)) for x in row] for row in r], dtype=object)
    a[:32] = np.array([[lo, hi, lo, hi], [hi, hi, hi, lo], [lo, lo, lo, hi], [hi, lo, hi, lo]] * 8, dtype=object)
    return a.astype(dt)


def intquant_vectors(ref):
    out = {}
    for tag in INT_TYPES:
        d, p, e = int_points(tag)
        codes = np.zeros(d.size, np.int32)
        dec = np.zeros_like(d)
        esc = np.zeros(d.size, np.int8)
        for i in range(d.size):
            codes[i], dec[i], esc[i] = ref.quantize_int(tag, d[i], p[i], e[i])
        rp, rc, re = int_recover_args(tag)
        rok = np.zeros(rp.size, np.int8)
        rout = np.zeros_like(rp)
        for i in range(rp.size):
            rok[i], rout[i] = ref.recover_int(tag, rp[i], int(rc[i]), re[i])
        ia = int_interp_args(tag)
        iout = np.zeros((6, ia.shape[0]), ia.dtype)
        for k in (0, 1, 2, 3, 4, 5):
            for i in range(ia.shape[0]):
                iout[k, i] = ref.interp_int(tag, k, *ia[i])
        out.update({f"{tag}_ori": d, f"{tag}_pred": p, f"{tag}_eb": e, f"{tag}_code": codes, f"{tag}_dec": dec,
                    f"{tag}_esc": esc, f"{tag}_rpred": rp, f"{tag}_rcode": rc, f"{tag}_reb": re, f"{tag}_rok": rok,
                    f"{tag}_rout": rout, f"{tag}_iargs": ia, f"{tag}_iout": iout})
    return out


def fse_blocks():
    rng = np.random.default_rng(7)
    blocks = []
    for n in (1, 2, 3, 100, 1000, 4097, 65536):
        blocks.append(np.full(n, 5, dtype=np.uint8))                                     # one symbol
        blocks.append(rng.integers(0, 256, n, dtype=np.uint8))                           # incompressible
        blocks.append(np.minimum(rng.geometric(0.3, n) - 1, 255).astype(np.uint8))       # skewed
        blocks.append(np.minimum(rng.geometric(0.05, n) - 1, 255).astype(np.uint8))      # wide
        blocks.append(rng.choice(np.array([0, 1, 2, 17], np.uint8), n, p=[0.7, 0.2, 0.09, 0.01]))
    for n in (65536, 30000):
        b = np.zeros(n, np.uint8)
        b[rng.integers(0, n, 3)] = 1                                                      # nearly RLE
        blocks.append(b)
        blocks.append((np.arange(n) % 3).astype(np.uint8))
        blocks.append(np.minimum(rng.poisson(12, n), 255).astype(np.uint8))
    return blocks


def quantizer_points(dtype, count=1200):
    rng = np.random.default_rng(11 if dtype == np.float32 else 12)
    ebs = [1e-1, 1e-3, 1e-4, 6.1e-5, 1e-7] + ([1e-10, 3e-12] if dtype == np.float64 else [])
    ori, pred, eb = [], [], []
    for e in ebs:
        base = rng.normal(0, 10, count).astype(dtype)
        k = rng.integers(-40000, 40000, count)
        frac = rng.uniform(-0.5, 0.5, count)
        ori += list(base)
        pred += list((base.astype(np.float64) - (k + frac) * 2 * e).astype(dtype))
        eb += [e] * count
        # points landing on the bound: err == +-eb after rounding
        p = rng.normal(0, 1, 200).astype(dtype)
        ori += list(p)
        pred += list((p.astype(np.float64) - np.float64(e) * rng.choice([-1.0, 1.0], 200)).astype(dtype))
        eb += [e] * 200
    special = [np.nan, np.inf, -np.inf, 0.0, -0.0, 1e30, -1e30, np.finfo(dtype).tiny, np.finfo(dtype).max]
    for s in special:
        for pv in (0.0, 1.0, np.nan, np.inf):
            ori.append(s)
            pred.append(pv)
            eb.append(1e-3)
    return np.array(ori, dtype), np.array(pred, dtype), np.array(eb, np.float64)


def interp_args(dtype, count=600):
    rng = np.random.default_rng(21 if dtype == np.float32 else 22)
    a = rng.normal(0, 1, (count, 4)) * rng.choice([1e-3, 1.0, 1e3, 1e7], (count, 1))
    return a.astype(dtype)


def main():
    ref = Ref()
    out = {}
    if "--only" in sys.argv and sys.argv[sys.argv.index("--only") + 1] == "intquant":
        np.savez_compressed(os.path.join(HERE, "intquant.npz"), **intquant_vectors(ref))
        print(json.dumps({"intquant": INT_TYPES}))
        return
    samp = sampler_vectors(ref)
    with open(os.path.join(HERE, "sampler.json"), "w") as f:
        json.dump(samp, f, indent=1)
    out["sampler"] = [(s["nblocks"], s["profiling"], len(s["starts"])) for s in samp]
    if "--only" in sys.argv and sys.argv[sys.argv.index("--only") + 1] == "sampler":
        print(json.dumps(out))
        return
    # FSE
    blocks = fse_blocks()
    src = np.concatenate(blocks)
    lens = np.array([b.size for b in blocks], np.int64)
    rets, outs = [], []
    for b in blocks:
        r, o = ref.fse_compress(b)
        rets.append(r)
        outs.append(np.frombuffer(o, np.uint8))
    np.savez_compressed(os.path.join(HERE, "fse.npz"), src=src, lens=lens, ret=np.array(rets, np.int64),
                        out=np.concatenate(outs) if outs else np.zeros(0, np.uint8),
                        out_lens=np.array([o.size for o in outs], np.int64))
    out["fse"] = len(blocks)
    # RLE/FSE + order-1 tANS, stored arrays
    arrs, encs = [], []
    for r in SMALL_RECIPES:
        c = codes_from_recipe(r)
        arrs.append(c)
        encs.append(np.frombuffer(ref.rle_encode(c), np.uint8))
        assert np.array_equal(ref.rle_decode(encs[-1].tobytes(), c.size), c)
    np.savez_compressed(os.path.join(HERE, "rlefse.npz"),
                        codes=np.concatenate(arrs), code_lens=np.array([a.size for a in arrs], np.int64),
                        enc=np.concatenate(encs), enc_lens=np.array([e.size for e in encs], np.int64),
                        recipes=np.array(json.dumps(SMALL_RECIPES)))
    out["rlefse"] = len(arrs)
    # large arrays: recipe + hash
    big = []
    for r in BIG_RECIPES:
        c = codes_from_recipe(r)
        e = ref.rle_encode(c)
        big.append(dict(r, enc_len=len(e), enc_sha256=hashlib.sha256(e).hexdigest(), method=e[0] & 0x7F,
                        delta=e[0] >> 7, codes_sha256=hashlib.sha256(c.tobytes()).hexdigest()))
    with open(os.path.join(HERE, "rlefse_big.json"), "w") as f:
        json.dump(big, f, indent=1)
    out["rlefse_big"] = len(big)
    # quantizer
    q = {}
    for dt, tag, fn in ((np.float32, "f32", ref.quantize_f32), (np.float64, "f64", ref.quantize_f64)):
        o, p, e = quantizer_points(dt)
        codes = np.zeros(o.size, np.int32)
        dec = np.zeros(o.size, dt)
        esc = np.zeros(o.size, np.int8)
        for i in range(o.size):
            codes[i], dec[i], esc[i] = fn(o[i], p[i], e[i])
        q.update({f"{tag}_ori": o, f"{tag}_pred": p, f"{tag}_eb": e, f"{tag}_code": codes, f"{tag}_dec": dec,
                  f"{tag}_esc": esc})
        rec = np.array([(ref.recover_f32 if tag == "f32" else ref.recover_f64)(p[i], int(codes[i]), e[i])
                        for i in range(o.size)], dt)
        q[f"{tag}_recover"] = rec
    np.savez_compressed(os.path.join(HERE, "quantizer.npz"), **q)
    out["quantizer"] = int(q["f32_ori"].size + q["f64_ori"].size)
    # interpolators
    it = {}
    for dt, tag, fn in ((np.float32, "f32", ref.interp_f32), (np.float64, "f64", ref.interp_f64)):
        a = interp_args(dt)
        res = np.zeros((5, a.shape[0]), dt)
        for k in range(5):
            for i in range(a.shape[0]):
                res[k, i] = fn(k, *a[i])
        it[f"{tag}_args"] = a
        it[f"{tag}_out"] = res
    np.savez_compressed(os.path.join(HERE, "interp.npz"), **it)
    out["interp"] = 5 * 2 * 600
    print(json.dumps(out))


if __name__ == "__main__":
    main()
