"""CPU oracle for the FP8-e4m3 block-scaled GEMM (test infrastructure only:
imported by tests and the benchmark's CPU baseline, never by the library).

The reference has no low-precision arithmetic (everything is DMatrix<f64>;
src/lora_drivers.rs:30-69 runs its GEMMs in FP64).  This module states the
number format and the block scaling the GPU path uses, so that the GPU GEMM
can be checked element for element:

* OCP FP8 E4M3 ("e4m3fn"): 1 sign, 4 exponent (bias 7), 3 mantissa bits, no
  infinities, S.1111.111 is NaN, largest finite 448, smallest subnormal 2^-9;
  conversion from FP32 rounds to nearest, ties to even.
* Scales are E8M0 bytes e (value 2^(e - 127)), one per 128 x 128 block of A
  and one per (column, 128-long K block) of B, chosen as the smallest power
  of two with absmax / scale <= 448 (a zero block gets 2^0).
* C = sum_k sA(i, k) q(A[i, k] / sA) * sB(k, j) q(B[k, j] / sB), here summed
  exactly (FP64).  The GPU's scaled matrix-core instruction does not sum its
  64-term dot products exactly in FP32: measured, its error stays below
  2^-12 sum_k |qa| |qb| (and shrinks about like 1/sqrt(K) with more K
  blocks); integer operands whose sums fit its internal precision come out
  exactly.  The tests use 2^-10 sum_k |qa| |qb|.

Pinned by tests/test_oracle_fp8.py: the 256-code table is derived from the
bit layout and checked against torch.float8_e4m3fn (an independent library
implementation of the same OCP format), rounding against a brute-force
nearest search over the table, and the scaled GEMM on blocks whose values
are exactly representable against the plain FP64 product.
"""
import numpy as np

FP8_MAX = 448.0
BLOCK = 128


def e4m3_table():
    """Value of each of the 256 codes (NaN for S.1111.111)."""
    v = np.empty(256)
    for c in range(256):
        s, e, m = c >> 7, (c >> 3) & 15, c & 7
        if e == 15 and m == 7:
            v[c] = np.nan
        elif e == 0:
            v[c] = (m / 8.0) * 2.0 ** -6
        else:
            v[c] = (1.0 + m / 8.0) * 2.0 ** (e - 7)
        if s:
            v[c] = -v[c]
    return v


def e4m3_round(x, truncate=False):
    """Round (FP32 values, given as any float array) to the nearest E4M3
    value, ties to even, saturating at +-448.  ``truncate`` rounds toward
    zero instead (a planted mistake for the tests, not the format's rule)."""
    x = np.asarray(x, dtype=np.float64)
    a = np.abs(x)
    with np.errstate(divide="ignore"):
        ex = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    # spacing of the grid around a: 2^(e - 3) for normals, 2^-9 below 2^-6
    ulp = 2.0 ** (np.maximum(ex, -6.0) - 3.0)
    r = (np.floor(a / ulp) if truncate else np.rint(a / ulp)) * ulp   # np.rint: half to even
    r = np.minimum(r, FP8_MAX)
    return np.copysign(r, x)


def e8m0_for(amax):
    """Smallest e with amax * 2^-e <= 448, as the E8M0 byte e + 127
    (127 for a zero block), clamped to the byte's range."""
    amax = float(amax)
    if not amax > 0.0:
        return 127
    e = int(np.ceil(np.log2(amax / FP8_MAX)))
    while np.ldexp(amax, -e) > FP8_MAX:
        e += 1
    while e > -127 and np.ldexp(amax, -(e - 1)) <= FP8_MAX:
        e -= 1
    return max(-127, min(127, e)) + 127


def quantize_blocks(A, rb, cb, truncate=False):
    """Block-scaled E4M3 copy of A (FP32 values): returns (Q, E) with Q the
    dequantised values (FP64, exactly sA * q) and E the E8M0 bytes, one per
    rb x cb block (E[bi, bj])."""
    A = np.asarray(A, dtype=np.float32).astype(np.float64)
    m, n = A.shape
    mb, nb = -(-m // rb), -(-n // cb)
    E = np.zeros((mb, nb), dtype=np.int64)
    Q = np.zeros_like(A)
    for bi in range(mb):
        for bj in range(nb):
            blk = A[bi * rb:(bi + 1) * rb, bj * cb:(bj + 1) * cb]
            e = e8m0_for(np.abs(blk).max() if blk.size else 0.0)
            E[bi, bj] = e
            s = 2.0 ** (e - 127)
            # the GPU scales in FP32 by the exact power of two, then rounds
            v = (blk.astype(np.float32) * np.float32(1.0 / s)).astype(np.float64)
            Q[bi * rb:(bi + 1) * rb, bj * cb:(bj + 1) * cb] = e4m3_round(v, truncate) * s
    return Q, E


def fp8_gemm(A, B, trans_a=False):
    """op(A) @ B with A quantised per 128 x 128 block (of A as stored) and B
    per column and 128-row block, products summed exactly (FP64)."""
    qa, _ = quantize_blocks(A, BLOCK, BLOCK)
    qb, _ = quantize_blocks(B, BLOCK, 1)
    return (qa.T if trans_a else qa) @ qb
