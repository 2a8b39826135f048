"""The 1-4-11 16-bit deviation format of the oracle (CPU).

Pinned by known codes, a round trip over all 65,536 codes, strict
monotonicity of the decoded values, and brute-force nearest-value rounding
(ties to even) against the table of all codes.  A planted truncating
encoder must fail the rounding check.
"""

import numpy as np
import pytest


def test_known_codes(oracle):
    enc = oracle.f16c_encode
    vals = np.array([0.0, -0.0, 2.0 ** -16, 2.0 ** -15, 0.5, 0.75, -0.75, 2.0 ** -26,
                     1.0, -1.0, np.inf, np.nan, 3 * 2.0 ** -27], dtype=np.float32)
    codes = [0x0000, 0x8000, 0x0400, 0x0800, 0x7800, 0x7C00, 0xFC00, 0x0001,
             0x7FFF, 0xFFFF, 0x7FFF, 0x7FFF, 0x0002]   # 1.5 * 2^-26 ties to the even 2
    assert list(enc(vals)) == codes
    dec = oracle.f16c_decode(np.array([0x0800, 0x7800, 0x7FFE, 0x0001, 0x8001], np.uint16))
    assert list(dec) == [2.0 ** -15, 0.5, np.float32((1 + 2046 / 2048) * 0.5), 2.0 ** -26,
                         -(2.0 ** -26)]


def _table(oracle):
    codes = np.arange(0x7FFF, dtype=np.uint16)          # every non-negative finite code
    return codes, oracle.f16c_decode(codes)


def test_round_trip_and_monotone(oracle):
    codes, vals = _table(oracle)
    assert np.all(np.diff(vals.astype(np.float64)) > 0)
    np.testing.assert_array_equal(oracle.f16c_encode(vals), codes)
    np.testing.assert_array_equal(oracle.f16c_encode(-vals[1:]), codes[1:] | 0x8000)


def _check_nearest(encode, oracle, x):
    codes, vals = _table(oracle)
    v = vals.astype(np.float64)
    a = np.abs(x).astype(np.float64)
    got = oracle.f16c_decode(encode(x)).astype(np.float64)
    i = np.clip(np.searchsorted(v, a), 1, len(v) - 1)
    lo, hi = v[i - 1], v[i]
    dlo, dhi = a - lo, hi - a
    want = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(codes[i - 1] % 2 == 0, lo, hi)))
    return np.array_equal(np.abs(got), want)


def test_rounding_is_nearest_ties_even(oracle):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-0.99, 0.99, 200000),
                        rng.uniform(-2e-4, 2e-4, 50000)]).astype(np.float32)
    # exact midpoints between neighbouring codes
    _, vals = _table(oracle)
    mid = ((vals[1:4000].astype(np.float64) + vals[:3999]) / 2).astype(np.float32)
    x = np.concatenate([x, mid[mid.astype(np.float64) == (vals[1:4000].astype(np.float64) + vals[:3999]) / 2]])
    assert _check_nearest(oracle.f16c_encode, oracle, x)
    rel = np.abs(oracle.f16c_decode(oracle.f16c_encode(x)) - x)
    big = np.abs(x) >= 2.0 ** -15
    assert np.all(rel[big] <= 2.0 ** -12 * np.abs(x[big]))
    assert np.all(rel[~big] <= 2.0 ** -27)


def test_planted_truncation_is_caught(oracle):
    def trunc(v):
        v = np.asarray(v, np.float32)
        t = (v.view(np.uint32) & np.uint32(0xFFFFF000)).view(np.float32)
        return oracle.f16c_encode(t)
    rng = np.random.default_rng(1)
    x = rng.uniform(-0.9, 0.9, 20000).astype(np.float32)
    assert not _check_nearest(trunc, oracle, x)
