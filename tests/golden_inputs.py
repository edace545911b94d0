"""Deterministic int16 code arrays for the entropy-coder fixtures, from integer arithmetic only
(splitmix64 over the element index), so that an array is regenerated bit for bit from its recipe on
any machine and numpy version; the large fixtures are stored as recipe + sha256 of the output."""
import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, n, stream=0):
    with np.errstate(over="ignore"):
        z = (np.arange(n, dtype=np.uint64) + np.uint64(seed * 1_000_003 + stream * 7_919)) * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def _geometric(h, bits=24):
    """Integer geometric(1/2) magnitudes: leading zero count of a `bits`-bit field."""
    u = (h >> np.uint64(64 - bits)).astype(np.int64) + 1
    # frexp of an integer below 2^53 is exact: exponent = bit length
    return bits - np.frexp(u.astype(np.float64))[1]


def codes_from_recipe(r):
    kind, n, seed, scale = r["kind"], int(r["n"]), int(r["seed"]), int(r["scale"])
    h = splitmix64(seed, n)
    sign = np.where((h >> np.uint64(3)) & np.uint64(1), -1, 1).astype(np.int64)
    if kind == "zeros":
        v = np.zeros(n, np.int64)
    elif kind == "sparse":  # about 90% zeros, small values otherwise
        nz = (h % np.uint64(100)) >= np.uint64(90)
        v = np.where(nz, sign * (1 + _geometric(splitmix64(seed, n, 1)) * max(scale, 1)), 0)
    elif kind == "laplace":  # two-sided geometric, spread by scale
        g = _geometric(h)
        low = (splitmix64(seed, n, 2) % np.uint64(max(scale, 1))).astype(np.int64)
        v = sign * (g * max(scale, 1) + low)
        v = np.where(g == 0, low * sign * (low > 0), v)
    elif kind == "escapes":  # laplace body plus ~3% large values and some -32768 escape codes
        g = _geometric(h)
        v = sign * g * max(scale, 1)
        h2 = splitmix64(seed, n, 3)
        big = (h2 % np.uint64(100)) < np.uint64(3)
        mag = 128 + (splitmix64(seed, n, 4) % np.uint64(32640)).astype(np.int64)
        v = np.where(big, sign * mag, v)
        v = np.where((h2 % np.uint64(1000)) == np.uint64(999), -32768, v)
    elif kind == "runs":  # piecewise-constant nonzero runs: the delta transform turns them into zeros
        starts = (h % np.uint64(40)) == np.uint64(0)
        level = ((splitmix64(seed, n, 5) % np.uint64(2 * scale + 1)).astype(np.int64) - scale)
        idx = np.maximum.accumulate(np.where(starts, np.arange(n), 0))
        v = level[idx]
    elif kind == "wide":  # full int16 range, uniform
        v = (h >> np.uint64(48)).astype(np.int64) - 32768
    else:
        raise ValueError(kind)
    return np.clip(v, -32768, 32767).astype(np.int16)


def sampler_field(r):
    """Fields for the parameter search's block sampler, from integer arithmetic scaled by powers of
    two (exact in float32 and float64):
      walk   an integer random walk along the flattened field: every block varies
      patch  zero except a box where a walk runs: profiling keeps only the blocks that touch it
      flat   a constant plus steps of 2^-20 (below any bound used): no block is kept
    """
    kind, shape, seed = r["kind"], tuple(r["shape"]), int(r["seed"])
    dt = np.float32 if r["dtype"] == "f32" else np.float64
    n = int(np.prod(shape))
    h = splitmix64(seed, n)
    if kind == "flat":
        v = 3.0 + (h % np.uint64(3)).astype(np.float64) * 2.0 ** -20
        return v.astype(dt).reshape(shape)
    step = (h % np.uint64(9)).astype(np.int64) - 4
    walk = (np.cumsum(step) % 4096).astype(np.float64) * 2.0 ** -6
    if kind == "walk":
        return walk.astype(dt).reshape(shape)
    if kind == "patch":
        f = np.zeros(shape, np.float64)
        box = tuple(slice(int(a), int(b)) for a, b in r["box"])
        f[box] = walk.reshape(shape)[box]
        return f.astype(dt)
    raise ValueError(kind)
