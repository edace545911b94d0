"""The parameter search's block sampler (ALGO_INTERP_LORENZO) against the reference's own source.

profiling_block and sampleBlocks (utils/Sample.hpp:5-344) compile from their own header and
Falcons/def.hpp, so oracle/_ref holds them as built from the reference tree; tests/golden/sampler.json
holds what they returned on integer-generated fields (recipes in golden_inputs.sampler_field).
Checked for every recipe: the sample block side the oracle picks for the shape, the profiling
branch, the profiled block starts in order, and the sampled blocks (gathered by the oracle from its
selected starts) byte for byte. The GPU tuner is checked against the oracle by the tuner's GPU tests.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from golden_inputs import sampler_field
from oracle_lib import REF_SO, Oracle, Ref

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "sampler.json")) as _f:
    CASES = json.load(_f)


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _gather(f, sel, side):
    out = []
    for s in sel:
        sl = tuple(slice(int(s[a]), int(s[a]) + side) for a in range(f.ndim))
        out.append(np.ascontiguousarray(f[sl]).reshape(-1))
    return np.concatenate(out) if out else np.zeros(0, f.dtype)


def _oracle_sample(orc, f, eb):
    conf = orc.config(f.shape, cmpr_algo=1, eb_mode=0, abs_eb=eb)
    return orc.tune_sample(f, conf)


@pytest.mark.parametrize("case", CASES, ids=[f"{c['kind']}-{'x'.join(map(str, c['shape']))}-{c['dtype']}"
                                             for c in CASES])
def test_sampler_matches_reference_vectors(orc, case):
    f = sampler_field(case)
    res = _oracle_sample(orc, f, case["eb"])
    assert res is not None, "the field is tuned by the reference"
    sbs, prof, starts, sel = res
    assert sbs == case["sbs"]
    assert prof == case["profiling"]
    assert starts[:, : f.ndim].tolist() == case["starts"]
    assert len(sel) == case["nblocks"]
    blocks = _gather(f, sel, sbs + 1)
    assert hashlib.sha256(blocks.tobytes()).hexdigest() == case["blocks_sha256"]


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("case", CASES[:9], ids=[f"{c['kind']}-{'x'.join(map(str, c['shape']))}-{c['dtype']}"
                                                 for c in CASES[:9]])
@pytest.mark.parametrize("scale", [1.0, 0.25])
def test_sampler_matches_reference_build(orc, case, scale):
    """The same, live against the reference build, also at a quarter of the bound (more blocks kept)."""
    ref = Ref()
    f = sampler_field(case)
    eb = case["eb"] * scale
    sbs, prof, starts, sel = _oracle_sample(orc, f, eb)
    r_prof, r_starts, r_blocks = ref.sample(f, sbs, eb)
    assert prof == r_prof
    assert np.array_equal(starts, r_starts)
    assert np.array_equal(_gather(f, sel, sbs + 1).view(np.uint8), r_blocks.view(np.uint8))


def test_sampler_skips_small_fields(orc):
    # a block of side+1 on every axis must stay under 5% of the field, as the reference requires
    f = np.zeros((20, 20, 20), np.float32)
    assert _oracle_sample(orc, f, 1e-3) is None
