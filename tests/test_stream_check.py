"""fz_check_prezstd on the CPU: the host-side parse that every decode runs before its first kernel.

Valid pre-zstd buffers come from the oracle (RLE+FSE and order-1 tANS streams); each corruption below
must be refused with ValueError (FZ_ERR_INVALID_ARGUMENT) without touching a GPU, so that a damaged
buffer can never send a decoder kernel outside the buffer."""
import struct

import numpy as np
import pytest

import falcons_amd as fa
from falcons_amd import synth
from oracle_lib import Oracle


def _oracle_pre(shape, data, **kw):
    o = Oracle()
    oc = o.config(shape, cmpr_algo=2, **kw)
    stream, _, pre = o.compress(data, oc, want_pre=True)
    cfg = fa.read_config(stream)
    return bytearray(pre), cfg, data.dtype


@pytest.fixture(scope="module")
def rle_case():
    d = synth.sinusoids((32, 32, 32), seed=3)  # 32768 codes < 76,800: RLE + FSE (RleFse:850-895)
    return _oracle_pre(d.shape, d, eb_mode=0, abs_eb=1e-2)


@pytest.fixture(scope="module")
def tans_case():
    d = synth.power_law((48, 48, 48), -3.0, seed=5)
    return _oracle_pre(d.shape, d, eb_mode=1, rel_eb=1e-4)


def enc_offset(pre, cfg, dtype):
    """Offset of the entropy stream: decomposition header, quantizer header, unpredictable values and
    the code count (InterpolationDecomposition.hpp:241-267, LinearQuantizer.hpp:142-180)."""
    n = cfg.conf.N
    q = 8 * n + 4 + 4 + 4 + 8 + 8 + 8
    nun = struct.unpack_from("<i", pre, q + 1 + 8 + 4)[0]
    return q + 17 + nun * np.dtype(dtype).itemsize + 8, q


def refused(pre, cfg, dtype):
    with pytest.raises(ValueError):
        fa.check_prezstd(bytes(pre), cfg, dtype)


def test_valid_buffers_pass(rle_case, tans_case):
    for pre, cfg, dt in (rle_case, tans_case):
        fa.check_prezstd(bytes(pre), cfg, dt)
    assert rle_case[0][enc_offset(*rle_case)[0]] & 0x7F == 0
    assert tans_case[0][enc_offset(*tans_case)[0]] & 0x7F == 1


@pytest.mark.parametrize("cut", [1, 7, 100])
def test_truncation(rle_case, tans_case, cut):
    for pre, cfg, dt in (rle_case, tans_case):
        refused(pre[:-cut], cfg, dt)


def test_decomposition_headers(rle_case):
    pre, cfg, dt = rle_case
    e, q = enc_offset(pre, cfg, dt)
    bad = bytearray(pre)
    bad[q] = 3  # LinearQuantizer uid
    refused(bad, cfg, dt)
    bad = bytearray(pre)
    struct.pack_into("<i", bad, q + 13, 1 << 30)  # more unpredictable values than elements
    refused(bad, cfg, dt)
    bad = bytearray(pre)
    struct.pack_into("<Q", bad, e - 8, struct.unpack_from("<Q", pre, e - 8)[0] + 1)  # code count
    refused(bad, cfg, dt)


def test_rle_headers(rle_case):
    pre, cfg, dt = rle_case
    e, _ = enc_offset(pre, cfg, dt)
    h = e + 5  # [u64 nr, nv, s_rcl, s_vcl, rbp, vbp] (RleFse:972-979)
    bad = bytearray(pre)
    bad[e] = 2  # no method 2
    refused(bad, cfg, dt)
    for field, value in ((0, 1 << 40), (1, 1 << 40), (2, 1 << 50), (3, 1 << 50), (4, 1 << 50), (5, 1 << 50)):
        bad = bytearray(pre)
        struct.pack_into("<Q", bad, h + 8 * field, value)
        refused(bad, cfg, dt)
    nr, nv, s_rcl = struct.unpack_from("<QQQ", pre, h)
    assert nr > 0 and s_rcl > 0
    first = h + 48  # the run-class stream's first FSE block header
    bad = bytearray(pre)
    bad[first] = 7  # no such block flag
    refused(bad, cfg, dt)
    if pre[first] == 1:
        bad = bytearray(pre)
        struct.pack_into("<I", bad, first + 1, 1 << 30)  # block longer than the stream
        refused(bad, cfg, dt)


def test_tans_headers(tans_case):
    pre, cfg, dt = tans_case
    e, _ = enc_offset(pre, cfg, dt)
    h = e + 5  # [u64 S] then per segment [u64 N][u64 K][u64 blen x K][u64 nvk x K] (RleFse:728-739)
    S = struct.unpack_from("<Q", pre, h)[0]
    assert 1 <= S <= 16
    bad = bytearray(pre)
    struct.pack_into("<Q", bad, h, 17)
    refused(bad, cfg, dt)
    K = struct.unpack_from("<Q", pre, h + 16)[0]
    for off, value in ((h + 8, 12345), (h + 16, 0), (h + 16, 17), (h + 24, 1 << 50), (h + 24 + 8 * K, 1 << 50)):
        bad = bytearray(pre)
        struct.pack_into("<Q", bad, off, value)
        refused(bad, cfg, dt)
