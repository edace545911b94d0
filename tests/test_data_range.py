"""data_range (Statistic.hpp:160-204, openmp off) on fields that hold NaNs, in the oracle.

The expected ranges are worked out by hand from the instruction rule, not by running a model:
MAXPS / MINPS (Intel SDM, MAXPS: "If only one value is a NaN ... the second operand (source operand)
... is written to the result") return the element whenever the lane or the element is NaN, so in
avx2_minmax (:97-157) a NaN element enters its lane and the lane's next element replaces it together
with everything the lane held; std::max / std::min (the lane fold and the tail) keep a NaN they hold
and pass over a NaN they are given. scalar_minmax (:38-44; the scalar build, and every field of 16
values or fewer) starts from d[0] with std::max / std::min.
"""
import ctypes

import numpy as np
import pytest

from oracle_lib import Oracle


@pytest.fixture(scope="module")
def oracle():
    o = Oracle()
    yield o
    o.L.fo_set_simd(1)


def rng_field(n, dtype, seed=3):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n).astype(dtype)


def data_range(oracle, d, simd):
    oracle.L.fo_set_simd(simd)
    return oracle.L.fo_data_range(0 if d.dtype == np.float32 else 1, d.ctypes.data_as(ctypes.c_void_p), d.size)


def span(d, keep):
    v = d[keep].astype(np.float64)
    return float(d.dtype.type(v.max()) - d.dtype.type(v.min()))


LANES = {np.float32: 8, np.float64: 4}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_no_nan_both_builds(oracle, dtype):
    d = rng_field(203, dtype)
    d[77] = 5.0
    d[-1] = -7.0
    for simd in (0, 1):
        assert data_range(oracle, d, simd) == 12.0
    d[0] = np.inf
    assert data_range(oracle, d, 1) == np.inf


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nan_first(oracle, dtype):
    n = 203
    d = rng_field(n, dtype)
    d[0] = np.nan
    # scalar: d[0] stays NaN through std::max / std::min
    assert np.isnan(data_range(oracle, d, 0))
    # AVX2: every lane starts at NaN and takes its first element; lane 0's first element is d[0]
    # itself (NaN), and its second, d[L], replaces it: the range of everything but d[0]
    keep = np.ones(n, bool)
    keep[0] = False
    assert data_range(oracle, d, 1) == span(d, keep)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nan_last_full_vector_lane0(oracle, dtype):
    n, L = 203, LANES[dtype]
    nfull = n // L * L
    d = rng_field(n, dtype)
    d[nfull - L] = np.nan
    assert np.isnan(data_range(oracle, d, 1))  # lane 0 ends NaN and the fold keeps it
    keep = ~np.isnan(d)
    assert data_range(oracle, d, 0) == span(d, keep)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nan_drops_lane_history(oracle, dtype):
    n, L = 203, LANES[dtype]
    d = rng_field(n, dtype)
    d[3] = 40.0    # the maximum, lane 3
    d[7 * L + 3] = -30.0  # the minimum, lane 3
    d[9 * L + 3] = np.nan  # lane 3: everything before is replaced by d[10 L + 3]
    keep = np.ones(n, bool)
    keep[[3, 7 * L + 3, 9 * L + 3]] = False
    assert data_range(oracle, d, 1) == span(d, keep)
    assert data_range(oracle, d, 0) == 70.0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nan_last_full_vector_other_lane(oracle, dtype):
    n, L = 203, LANES[dtype]
    nfull = n // L * L
    d = rng_field(n, dtype)
    d[L + 2] = 40.0                # lane 2 holds the maximum ...
    d[nfull - L + 2] = np.nan      # ... and ends NaN: the fold passes over the whole lane
    keep = np.ones(n, bool)
    keep[np.arange(2, nfull, L)] = False
    assert data_range(oracle, d, 1) == span(d, keep)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nan_in_tail(oracle, dtype):
    n, L = 203, LANES[dtype]
    assert n % L
    d = rng_field(n, dtype)
    d[-1] = np.nan
    keep = ~np.isnan(d)
    for simd in (0, 1):
        assert data_range(oracle, d, simd) == span(d, keep)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sixteen_values_scan_scalar(oracle, dtype):
    d = rng_field(16, dtype)
    d[15] = 9.0
    d[8] = np.nan  # would be lane 0 of the last full vector in the AVX2 scan
    keep = ~np.isnan(d)
    for simd in (0, 1):
        assert data_range(oracle, d, simd) == span(d, keep)
    d[0] = np.nan
    for simd in (0, 1):
        assert np.isnan(data_range(oracle, d, simd))
