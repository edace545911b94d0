"""Seeded synthetic fields for the benchmark workloads and tests (no datasets are available).

  sinusoids(shape, seed)    sum of 16 sinusoids, amplitude U(0.1, 1), integer frequency 1..8 per
                            axis, random phase, plus N(0, 1e-3) noise
  power_law(shape, slope)   Gaussian random field with power spectrum |k|^slope (amplitude
                            |k|^(slope/2)), zero mean and unit variance, synthesised with the FFT
  front(shape)              power-law field plus a smooth radial tanh front of width 8 cells
"""
import numpy as np


def sinusoids(shape, seed=1, dtype=np.float32):
    rng = np.random.default_rng(seed)
    axes = [np.arange(s, dtype=np.float64) / s for s in shape]
    grids = np.meshgrid(*axes, indexing="ij", sparse=True)
    f = np.zeros(shape, dtype=np.float64)
    for _ in range(16):
        amp = rng.uniform(0.1, 1.0)
        freqs = rng.integers(1, 9, size=len(shape))
        phase = rng.uniform(0, 2 * np.pi)
        arg = sum(2 * np.pi * fr * g for fr, g in zip(freqs, grids))
        f += amp * np.sin(arg + phase)
    f += rng.normal(0, 1e-3, size=shape)
    return f.astype(dtype)


def power_law(shape, slope=-3.0, seed=2, dtype=np.float32):
    rng = np.random.default_rng(seed)
    white = rng.standard_normal(shape, dtype=np.float32)
    spec = np.fft.rfftn(white)
    del white
    ks = [np.fft.fftfreq(s) for s in shape[:-1]] + [np.fft.rfftfreq(shape[-1])]
    k2 = None
    for i, k in enumerate(ks):
        sh = [1] * len(shape)
        sh[i] = k.size
        kk = (k.astype(np.float32) ** 2).reshape(sh)
        k2 = kk if k2 is None else k2 + kk
    amp = np.zeros_like(k2)
    nz = k2 > 0
    amp[nz] = k2[nz] ** (slope / 4.0)  # |k|^(slope/2)
    spec *= amp
    del amp, k2
    f = np.fft.irfftn(spec, s=shape, axes=tuple(range(len(shape)))).astype(np.float64)
    del spec
    f -= f.mean()
    f /= f.std()
    return f.astype(dtype)


def front(shape, seed=3, dtype=np.float64):
    f = power_law(shape, slope=-5.0 / 3.0, seed=seed, dtype=np.float64)
    r0 = 0.3 * min(shape)
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) - (s - 1) / 2 for s in shape], indexing="ij", sparse=True)
    r = np.sqrt(sum(g * g for g in grids))
    f += np.tanh((r - r0) / 8.0)
    return f.astype(dtype)


def chunk_sides(count, seed=4, lo=32, hi=128):
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 3)))).astype(int)


def chunk_mix(count, seed=4):
    """The batched chunk mix: per-axis sides floor(exp(U(ln 32, ln 128))) (seed), chunk i a k^-3
    float32 field of seed seed + i, REL bound drawn uniformly from {1e-2, 1e-3, 1e-4} (seed + 7919).
    Returns (sides [count, 3], rels [count]); chunk_field(sides, i, seed) makes chunk i."""
    sides = chunk_sides(count, seed=seed)
    rels = np.random.default_rng(seed + 7919).choice([1e-2, 1e-3, 1e-4], size=count)
    return sides, rels


def chunk_field(sides, i, seed=4):
    return power_law(tuple(int(v) for v in sides[i]), -3.0, seed=seed + i)


# request classes of the serving mix: (side lo, side hi, ndim, dtype, base field shape, base seed)
SERVE_CLASSES = {
    0: (64, 512, 3, np.float32, (512, 512, 512), 5),
    1: (1024, 8192, 2, np.float32, (8192, 8192), 6),
    2: (128, 256, 3, np.float64, (256, 256, 256), 7),
}


def serving_mix(count, seed=5):
    """Request mix: 60% compress / 40% decompress; 70% 3D float32 side in [64, 512], 20% 2D float32 side
    in [1024, 8192], 10% 3D float64 side in [128, 256] (sides log-uniform); ABS or REL bound
    log-uniform in [1e-5, 1e-2]. Fields are crops of one k^-3 field per class (serving_base) at
    random offsets (serving_field)."""
    rng = np.random.default_rng(seed)
    return [{
        "op": 0 if rng.random() < 0.6 else 1,
        "cls": int(rng.choice(3, p=[0.7, 0.2, 0.1])),
        "u": rng.random(3),
        "mode": int(rng.integers(0, 2)),
        "eb": float(10 ** rng.uniform(-5, -2)),
    } for _ in range(count)]


def serving_base(classes=(0, 1, 2)):
    return {c: power_law(SERVE_CLASSES[c][4], -3.0, seed=SERVE_CLASSES[c][5], dtype=SERVE_CLASSES[c][3])
            for c in classes}


def serving_field(base, r):
    lo, hi, nd = SERVE_CLASSES[r["cls"]][:3]
    side = int(np.floor(np.exp(np.log(lo) + r["u"][0] * (np.log(hi) - np.log(lo)))))
    b = base[r["cls"]]
    off = [int(r["u"][1 + (k % 2)] * (b.shape[k] - side)) for k in range(nd)]
    return np.ascontiguousarray(b[tuple(slice(o, o + side) for o in off)])
