"""GPU parity: the FP8-e4m3 block-scaled GEMM (128 x 128 block scales on A,
per-column 128-K scales on B, E8M0 powers of two, FP32 accumulation) against
the FP8 oracle (oracle/fp8.py).

Tolerance: the scaled matrix-core instruction's 64-term dot products are not
summed exactly in FP32 (measured error below 2^-12 sum_k |qa||qb|), so
results are compared to the exact sum of the quantised products within
TOL = 2^-10 sum_k |qa||qb| elementwise.  That is far below the FP8
quantisation error itself (2^-4 per operand); the tests also check that the
GPU result lies several times closer to the right oracle than to oracles
with a planted mistake (truncating rounding) and outside the tolerance of
one that takes A's scale from the neighbouring K block.  Integer blocks (entries in [-7, 7], block maximum 7) quantise and
sum exactly: there the GPU result must equal the FP64 product bit for bit,
which pins the operand and scale layout of the instruction."""
import numpy as np
import pytest

import randnla_amd as L
from oracle import fp8 as F

pytestmark = pytest.mark.gpu

TOL = 2.0 ** -10
rng = np.random.default_rng(12)
VALS = F.e4m3_table()
VALS = VALS[~np.isnan(VALS)]


def int_blocks(r, c, rb, cb):
    X = rng.integers(-7, 8, (r, c)).astype(np.float64)
    for bi in range(0, r, rb):
        for bj in range(0, c, cb):
            X[bi, bj] = 7.0 * rng.choice([-1.0, 1.0])
    return X


def representable(r, c, rb, cb, spread=(-20, 10)):
    """E4M3 values times one power of two per rb x cb block, block maximum
    448 * 2^k (scale exactly 2^k)."""
    X = rng.choice(VALS, size=(r, c))
    for bi in range(0, r, rb):
        for bj in range(0, c, cb):
            X[bi, bj] = 448.0 * rng.choice([-1.0, 1.0])
            X[bi:bi + rb, bj:bj + cb] *= 2.0 ** int(rng.integers(*spread))
    return X


def oracle(A, B, trans, truncate=False):
    qa, _ = F.quantize_blocks(A, 128, 128, truncate)
    qb, _ = F.quantize_blocks(B, 128, 1, truncate)
    qa = qa.T if trans else qa
    return qa @ qb, TOL * (np.abs(qa) @ np.abs(qb))


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("M,N,K", [(300, 70, 260), (128, 320, 128), (33, 5, 17), (257, 641, 1000)])
def test_gemm_fp8_integer_blocks_bit_exact(M, N, K, trans):
    A = int_blocks(K, M, 128, 128) if trans else int_blocks(M, K, 128, 128)
    B = int_blocks(K, N, 128, 1)
    got = L.gemm(A, B, trans_a=trans, precision=L.FP8)
    assert np.array_equal(got, (A.T if trans else A) @ B)


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("M,N,K", [(300, 70, 260), (33, 5, 17)])
def test_gemm_fp8_representable_wide_scales(M, N, K, trans):
    A = representable(K, M, 128, 128) if trans else representable(M, K, 128, 128)
    B = representable(K, N, 128, 1)
    got = L.gemm(A, B, trans_a=trans, precision=L.FP8)
    want, tol = oracle(A, B, trans)
    # the operands are exact E4M3 multiples: the oracle is the FP64 product
    assert np.array_equal(want, (A.T if trans else A) @ B)
    assert np.all(np.abs(got - want) <= tol)


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("M,N,K", [(300, 70, 260), (1000, 400, 3000), (513, 321, 129)])
def test_gemm_fp8_matches_oracle(M, N, K, trans):
    shape = (K, M) if trans else (M, K)
    # rows and columns scaled over e^6 each: block scales differ from
    # per-row or per-column ones
    A = (rng.standard_normal(shape) * np.exp(rng.uniform(-3, 3, (1, shape[1])))
         * np.exp(rng.uniform(-3, 3, (shape[0], 1))))
    B = rng.standard_normal((K, N))
    A = A.astype(np.float32).astype(np.float64)
    B = B.astype(np.float32).astype(np.float64)
    got = L.gemm(A, B, trans_a=trans, precision=L.FP8)
    want, tol = oracle(A, B, trans)
    assert np.all(np.abs(got - want) <= tol), np.max(np.abs(got - want) / tol)
    # planted mistakes in the oracle: the GPU result lies several times
    # closer (Frobenius norm) to the right oracle than to each mistaken one
    d = np.linalg.norm(got - want)
    wt, _ = oracle(A, B, trans, truncate=True)          # truncation instead of RNE
    assert np.linalg.norm(got - wt) > 4 * d
    # scale of A taken from the next K block (an off-by-one in the scale
    # index): elementwise far outside the tolerance
    qa, E = F.quantize_blocks(A, 128, 128)
    qb, _ = F.quantize_blocks(B, 128, 1)
    S = 2.0 ** (E - 127.0)
    grow = lambda X: np.kron(X, np.ones((128, 128)))[:A.shape[0], :A.shape[1]]
    qw = qa / grow(S) * grow(np.roll(S, -1, axis=0 if trans else 1))
    ww = (qw.T if trans else qw) @ qb
    assert np.any(np.abs(got - ww) > tol)
    # and the FP8 product approximates the FP64 one at the FP8 level
    exact = (A.T if trans else A) @ B
    assert np.linalg.norm(got - exact) / np.linalg.norm(exact) < 0.05


def test_gemm_fp8_zero_and_ragged_blocks():
    A = np.zeros((200, 300))
    A[150:, 250:] = rng.standard_normal((50, 50))
    B = rng.standard_normal((300, 3))
    got = L.gemm(A, B, precision=L.FP8)
    want, tol = oracle(A, B, False)
    assert np.all(np.abs(got - want) <= tol)
    assert not np.any(got[:128])
