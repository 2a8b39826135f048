"""The 16-bit deviation code of the oracle (CPU).

The code of |v| is F = fl32(|v| + 2^-15) rounded to 12 significant bits,
stored as a 4-bit exponent and 11 mantissa bits; decoding gives F - 2^-15.
Exponent code 15 is not finite (0x7800 infinity, above it NaN), so the
finite range is |v| < 1.  Pinned by known codes, a round trip over all
30,720 finite magnitudes and their negatives, strict monotonicity,
brute-force nearest rounding of F (ties to even) against the table of all
finite codes, the resulting error bounds, and non-finite codes for every
magnitude that rounds to F >= 1.  A planted truncating encoder must fail the
rounding check.
"""

import numpy as np

OFF = np.float32(2.0 ** -15)


def test_known_codes(oracle):
    vals = np.array([0.0, -0.0, 2.0 ** -26, 2.0 ** -15, 0.5, -0.5, 1.0, 3 * 2.0 ** -27,
                     1.9999, 2.5, np.inf, np.nan], dtype=np.float32)
    codes = [0x0000, 0x8000, 0x0001, 0x0800, 0x7000, 0xF000, 0x7800, 0x0002,
             0x7FFF, 0x7FFF, 0x7FFF, 0x7FFF]     # 1.5 * 2^-26 ties to the even code 2
    assert list(oracle.f16c_encode(vals)) == codes
    dec = oracle.f16c_decode(np.array([0x0800, 0x7000, 0x0001, 0x8001, 0x77FF], np.uint16))
    assert list(dec) == [2.0 ** -15, np.float32(0.5 - 2.0 ** -15), 2.0 ** -26, -(2.0 ** -26),
                         np.float32((1 - 2.0 ** -12) - 2.0 ** -15)]
    assert list(oracle.f16c_decode(np.array([0x7800, 0xF800], np.uint16))) == [np.inf, -np.inf]
    zero = oracle.f16c_decode(np.array([0x0000, 0x8000], np.uint16))
    assert list(zero) == [0.0, 0.0] and not np.signbit(zero).any()
    assert np.isnan(oracle.f16c_decode(np.array([0x7801, 0x7FFF, 0xFFFF], np.uint16))).all()


def _table(oracle):
    codes = np.arange(0x7800, dtype=np.uint16)
    return codes, oracle.f16c_decode(codes)


def test_magnitudes_from_one_up_are_not_finite(oracle):
    top = oracle.f16c_decode(np.array([0x77FF], np.uint16))[0].astype(np.float64)
    F_top = top + 2.0 ** -15
    # the first magnitude whose F rounds up to 1: halfway between F_top and 1
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(1.0, 1.9, 1000), [1.0, 0.99999, (F_top + 1.0) / 2],
                        rng.uniform(2.0, 1e6, 1000)]).astype(np.float32)
    x = x[(np.abs(x).astype(np.float32) + np.float32(2.0 ** -15)).astype(np.float64) >
          (F_top + 1.0) / 2]
    codes = oracle.f16c_encode(np.concatenate([x, -x]))
    assert np.all((codes & 0x7FFF) >= 0x7800)
    assert not np.isfinite(oracle.f16c_decode(codes)).any()


def test_round_trip_and_monotone(oracle):
    codes, vals = _table(oracle)
    assert np.all(np.diff(vals.astype(np.float64)) > 0)
    np.testing.assert_array_equal(oracle.f16c_encode(vals), codes)
    np.testing.assert_array_equal(oracle.f16c_encode(-vals[1:]), codes[1:] | 0x8000)


def _check_nearest(encode, oracle, x):
    """The decoded F is the nearest code value to fl32(|x| + 2^-15), ties to
    the even code."""
    codes, vals = _table(oracle)
    F = vals.astype(np.float64) + 2.0 ** -15
    t = (np.abs(x).astype(np.float32) + OFF).astype(np.float64)
    got = oracle.f16c_decode(encode(x)).astype(np.float64)
    i = np.clip(np.searchsorted(F, t), 1, len(F) - 1)
    lo, hi = F[i - 1], F[i]
    dlo, dhi = t - lo, hi - t
    want = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(codes[i - 1] % 2 == 0, lo, hi)))
    return np.array_equal(np.abs(got) + 2.0 ** -15, want)


def test_rounding_is_nearest_ties_even(oracle):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-0.9997, 0.9997, 200000),
                        rng.uniform(-2e-4, 2e-4, 50000)]).astype(np.float32)
    _, vals = _table(oracle)
    F = vals[:4000].astype(np.float64) + 2.0 ** -15
    mid = (F[1:] + F[:-1]) / 2 - 2.0 ** -15          # magnitudes whose F is an exact tie
    mid32 = mid.astype(np.float32)
    x = np.concatenate([x, mid32[(mid32.astype(np.float64) + 2.0 ** -15) == (F[1:] + F[:-1]) / 2]])
    assert _check_nearest(oracle.f16c_encode, oracle, x)
    q = oracle.f16c_decode(oracle.f16c_encode(x)).astype(np.float64)
    a = np.abs(x).astype(np.float64)
    Fx = a + 2.0 ** -15
    ulp12 = 2.0 ** (np.floor(np.log2(Fx)) - 11)
    assert np.all(np.abs(np.abs(q) - a) <= 0.5 * ulp12 + 2.0 ** -24 * Fx)
    small = a < 2.0 ** -15
    assert np.all(np.abs(np.abs(q[small]) - a[small]) <= 2.0 ** -27 + 2.0 ** -39)


def test_planted_truncation_is_caught(oracle):
    def trunc(v):
        v = np.asarray(v, np.float32)
        t = (np.abs(v) + OFF).astype(np.float32)
        t = (t.view(np.uint32) & np.uint32(0xFFFFF000)).view(np.float32)
        return oracle.f16c_encode(np.copysign((t - OFF).astype(np.float32), v))
    rng = np.random.default_rng(1)
    x = rng.uniform(-0.9, 0.9, 20000).astype(np.float32)
    assert not _check_nearest(trunc, oracle, x)
