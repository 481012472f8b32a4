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
