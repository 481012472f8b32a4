"""CPU oracle of the int8 far-field storage (test infrastructure only:
imported by tests, never by the product path).

The storage has no counterpart in the reference (whose FGMRES operator is the
FP64 matrix, la/solver.py:178); this restates what mbem_amd/csrc/mixed.hip
does, so the GPU bytes, scales and products can be checked bit for bit:

  for row r and 64-column segment k, s = float32(max |A[r, seg]| / 127);
  q = rint(A / s) (ties to even, the division in float64) clamped to
  [-127, 127], q = 0 where s = 0; the byte is q + 128; the represented value
  is s (byte - 128).

The split form (mbem_cast_i8e, the mixed method's outer operator) keeps each
code's rounding remainder as well: e = rint((A - s q) / s * 2^24) (the
difference, the division and the product by 2^24 in float64), clamped to
+-(2^23 - 1), e = 0 where s = 0; the represented value is
s (byte - 128 + e 2^-24).  On the device the remainders are packed four to
three little-endian 32-bit words of 24-bit two's complement (``pack24``).
"""
from __future__ import annotations

import numpy as np

SEG = 64


def segments(n: int) -> int:
    return -(-n // SEG) if n > 0 else 0


def cast(A) -> tuple:
    """(codes (m, n) uint8, scales (m, nseg) float32) of a float64 matrix."""
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    nseg = segments(n)
    codes = np.empty((m, n), dtype=np.uint8)
    scales = np.empty((m, nseg), dtype=np.float32)
    for k in range(nseg):
        blk = A[:, k * SEG:(k + 1) * SEG]
        s = (np.abs(blk).max(axis=1) / 127.0).astype(np.float32)
        scales[:, k] = s
        sd = s.astype(np.float64)[:, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            q = np.where(sd > 0, np.rint(blk / np.where(sd > 0, sd, 1.0)), 0.0)
        codes[:, k * SEG:(k + 1) * SEG] = (np.clip(q, -127, 127) + 128).astype(np.uint8)
    return codes, scales


def dequantize(codes, scales) -> np.ndarray:
    """float64 matrix the int8 storage represents."""
    n = codes.shape[1]
    seg = np.arange(n) // SEG
    return (codes.astype(np.float64) - 128.0) * scales[:, seg].astype(np.float64)


def cast_split(A) -> tuple:
    """(codes, scales, remainders (m, n) int32) of a float64 matrix."""
    A = np.asarray(A, dtype=np.float64)
    codes, scales = cast(A)
    n = A.shape[1]
    sd = scales[:, np.arange(n) // SEG].astype(np.float64)
    q = codes.astype(np.float64) - 128.0
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.where(sd > 0, np.rint((A - sd * q) / np.where(sd > 0, sd, 1.0) * 16777216.0), 0.0)
    e = np.clip(e, -8388607.0, 8388607.0).astype(np.int32)
    return codes, scales, e


def dequantize_split(codes, scales, resid) -> np.ndarray:
    """float64 matrix the split storage represents."""
    n = codes.shape[1]
    seg = np.arange(n) // SEG
    v = (codes.astype(np.float64) - 128.0) + resid.astype(np.float64) * 2.0 ** -24
    return v * scales[:, seg].astype(np.float64)


def pack24(resid) -> np.ndarray:
    """(m, 3 ceil(n/4) 4) bytes: each row's remainders, zero-padded to a
    multiple of 4 columns, as 24-bit little-endian two's complement."""
    resid = np.asarray(resid, dtype=np.int64)
    m, n = resid.shape
    n4 = -(-n // 4) * 4
    r = np.zeros((m, n4), dtype=np.int64)
    r[:, :n] = resid
    u = (r & 0xFFFFFF).astype(np.uint32)
    b = np.stack([u & 0xFF, (u >> 8) & 0xFF, (u >> 16) & 0xFF], axis=-1).astype(np.uint8)
    return b.reshape(m, 3 * n4)
