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
