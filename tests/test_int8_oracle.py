"""CPU pinning of the int8 storage oracle (oracle/int8.py).

Checked against what the format's definition fixes, not against a retyping
of the oracle: every represented value is an integer multiple q s of its
segment's scale with |q| <= 127, the nearest such multiple to the entry
(brute force over all 255 candidates), ties to even q; the segment's largest
magnitude maps to q = +-127; a segment of zeros stores zeros; segments are
64 columns wide.
"""
import numpy as np

from oracle import int8 as oi


def test_nearest_multiple_by_brute_force():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((5, 300)) * np.exp(rng.uniform(-8, 2, 300))[None, :]
    A[2] *= 1e-12
    codes, scales = oi.cast(A)
    assert scales.shape == (5, 5)
    qs = np.arange(-127, 128, dtype=np.float64)
    for r in range(5):
        for c in range(300):
            s = float(scales[r, c // 64])
            cand = qs * s
            d = np.abs(cand - A[r, c])
            best = qs[d == d.min()]
            q = best[0] if best.size == 1 else best[best % 2 == 0][0]
            assert codes[r, c] == int(q) + 128, (r, c)
    D = oi.dequantize(codes, scales)
    step = scales[:, np.arange(300) // 64].astype(np.float64)
    assert np.all(np.abs(D - A) <= 0.5 * step * (1 + 1e-12))


def test_segment_extremes_ties_and_zeros():
    A = np.zeros((3, 130))
    A[0, :64] = np.linspace(-1.0, 1.0, 64)
    A[0, 5] = 63.5                     # s = 0.5 exactly
    A[0, 6:9] = [0.25, 0.75, -1.25]    # q = 0.5, 1.5, -2.5: ties to 0, 2, -2
    A[1, 64:128] = -3.0                # constant segment: all -127
    codes, scales = oi.cast(A)
    assert scales.shape == (3, 3)
    assert scales[0, 0] == np.float32(0.5)
    assert codes[0, 5] == 127 + 128
    assert list(codes[0, 6:9].astype(int) - 128) == [0, 2, -2]
    assert np.all(codes[1, 64:128] == 1) and scales[1, 1] == np.float32(3.0 / 127.0)
    # all-zero segments: scale 0, every code 128 (value 0)
    assert scales[2, 0] == 0 and np.all(codes[2] == 128)
    assert scales[0, 2] == 0 and np.all(codes[0, 128:] == 128)
    assert np.all(oi.dequantize(codes, scales)[2] == 0)


def test_split_remainder_is_the_nearest_integer_exactly():
    """The remainder e of the split form against exact rational arithmetic:
    t = (A - s q) / s * 2^24 computed with fractions, |e - t| <= 1/2 (plus
    float64's rounding of the division, below 2^-28), so the represented
    value s (q + e 2^-24) is A to within s 2^-25 (and float64's rounding)
    on every entry of a matrix spanning 30 decades."""
    from fractions import Fraction
    rng = np.random.default_rng(5)
    A = rng.standard_normal((6, 400)) * np.exp(rng.uniform(-35, 5, 400))[None, :]
    A[3] *= 1e-9
    codes, scales, resid = oi.cast_split(A)
    assert resid.dtype == np.int32 and resid.shape == A.shape
    for r in range(6):
        for c in range(0, 400, 7):
            s = Fraction(float(scales[r, c // 64]))
            q = int(codes[r, c]) - 128
            t = (Fraction(float(A[r, c])) - s * q) / s * 2 ** 24
            assert abs(int(resid[r, c]) - t) <= Fraction(1, 2) + Fraction(1, 2 ** 28), (r, c)
    D = oi.dequantize_split(codes, scales, resid)
    step = scales[:, np.arange(400) // 64].astype(np.float64)
    assert np.all(np.abs(D - A) <= step * 2.0 ** -25 * (1 + 1e-6) + np.abs(A) * 2.0 ** -52)
    # the codes are the int8 form's, unchanged
    c8, s8 = oi.cast(A)
    assert np.array_equal(codes, c8) and np.array_equal(scales, s8)


def test_split_remainder_known_answers():
    """Entries built as s (q + k 2^-24) exactly (s = 1/8: the segment holds
    127 / 8) give back q and k; a constant segment and a zero segment have
    zero remainders; the remainder's sign follows A - s q."""
    rng = np.random.default_rng(6)
    q = rng.integers(-126, 127, 63)
    k = rng.integers(-(2 ** 23) + 1, 2 ** 23 - 1, 63)
    A = np.zeros((3, 128))
    A[0, 0] = 127 / 8
    A[0, 1:64] = q / 8 + k * 2.0 ** -27
    A[1, :64] = -2.5
    A[2, 64:] = 0.0
    A[2, :64] = 1.0
    A[2, 5] = 1.0 + 2.0 ** -22          # 2^-22 / s * 2^24 = 508 above A[2, 0]'s remainder
    codes, scales, resid = oi.cast_split(A)
    assert scales[0, 0] == np.float32(0.125)
    assert list(codes[0, 1:64].astype(int) - 128) == list(q)
    assert list(resid[0, 1:64].astype(np.int64)) == list(k)
    assert resid[0, 0] == 0
    assert np.all(resid[1, :64] == resid[1, 0]) and np.all(resid[:, 64:] == 0)
    s = np.float64(scales[2, 0])
    for c in (0, 5):
        want = np.rint((A[2, c] - s * 127) / s * 2.0 ** 24)
        assert resid[2, c] == want
    assert resid[2, 5] - resid[2, 0] == np.rint(4.0 / s)
    D = oi.dequantize_split(codes, scales, resid)
    assert np.array_equal(D[0], A[0])


def test_split_remainder_packing():
    """Four remainders to three little-endian words of 24-bit two's
    complement, a row padded to a multiple of four columns with zeros."""
    r = np.array([[1, -1, 2 ** 23 - 1, -(2 ** 23) + 1, 0x123456]], dtype=np.int32)
    b = oi.pack24(r)
    assert b.shape == (1, 24)
    assert list(b[0, :12]) == [1, 0, 0, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0x7F, 1, 0, 0x80]
    assert list(b[0, 12:15]) == [0x56, 0x34, 0x12] and not b[0, 15:].any()
    words = b[0, :12].view("<u4")
    vals = [words[0] & 0xFFFFFF, (words[0] >> 24) | ((words[1] & 0xFFFF) << 8),
            (words[1] >> 16) | ((words[2] & 0xFF) << 16), words[2] >> 8]
    signed = [int(v) - (1 << 24) if v >= 1 << 23 else int(v) for v in vals]
    assert signed == [1, -1, 2 ** 23 - 1, -(2 ** 23) + 1]
