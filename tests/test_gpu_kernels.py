"""GPU parity: counter streams, sketch operators, MFMA GEMMs and the small
factorizations, through the C ABI, against the oracle."""
import numpy as np
import pytest

import randnla_amd as L
from oracle import la, rng as R

pytestmark = pytest.mark.gpu


def _stream(gen, key, ctr=(0, 0)):
    return R.Stream(gen, key, ctr), L.rng(gen, key, ctr)


@pytest.mark.parametrize("gen", [R.THREEFRY2X64, R.PHILOX4X32])
@pytest.mark.parametrize("ctr", [(0, 0), (2 ** 64 - 3, 1), (12345, 2 ** 63)])
def test_device_stream_words_bit_exact(gen, ctr):
    o, r = _stream(gen, (0x0123456789ABCDEF, 0x0FEDCBA987654321 if gen == 0 else 0x1234), ctr)
    got = L.stream_words(r, 5, 1000)
    np.testing.assert_array_equal(got, o.words64(5, 1000))


def test_device_kat(golden):
    k = golden["rng_kat"]
    _, r = _stream(R.THREEFRY2X64, k["threefry_key"])
    assert [int(w) for w in L.stream_words(r, 0, 20)] == k["threefry2x64"]
    _, p = _stream(R.PHILOX4X32, (k["philox_seed1"], k["philox_seed2"]))
    w = L.stream_words(p, 0, 20)
    v = k["philox4x32"]
    assert [int(x) for x in w] == [v[2 * i] | (v[2 * i + 1] << 32) for i in range(20)]
    s42 = L.rng_from_seed(L.THREEFRY2X64, 42)
    assert int(L.stream_words(s42, 0, 1)[0]) == k["threefry_seed42_first_u64"]


@pytest.mark.parametrize("dist", [0, 1, 2, 3])
def test_sketching_operator_matches_oracle(dist):
    rows, cols = 37, 129
    got = L.sketching_operator(dist, rows, cols)            # reference stream
    want = R.sketching_operator(dist, rows, cols, la.reference_stream())
    if dist == 0:
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
    else:
        np.testing.assert_array_equal(got, want)


def test_sketching_operator_errors():
    with pytest.raises(L.RandNLAError) as e:
        L.sketching_operator(0, 0, 5)
    assert e.value.kind == "InvalidDimensions"


def _bf16(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("M,N,K,trans", [(130, 70, 40, False), (64, 3, 1000, True),
                                         (257, 320, 129, False), (33, 65, 4100, True)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_gemm_f64(M, N, K, trans, dt):
    r = np.random.default_rng(M + N + K)
    A = r.standard_normal((K, M) if trans else (M, K)).astype(dt)
    B = r.standard_normal((K, N)).astype(dt)
    got = L.gemm(A, B, trans_a=trans, precision=L.FP64)
    want = (A.T if trans else A).astype(np.float64) @ B.astype(np.float64)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.sqrt(K) * 8)


@pytest.mark.parametrize("M,N,K,trans", [(256, 320, 512, False), (512, 320, 4096, True),
                                         (1000, 64, 264, False), (136, 200, 8000, True),
                                         (24, 40, 16, False)])
def test_gemm_bf16(M, N, K, trans):
    r = np.random.default_rng(7 * M + N)
    A = _bf16(r.standard_normal((K, M) if trans else (M, K)))
    B = _bf16(r.standard_normal((K, N)))
    got = L.gemm(A, B, trans_a=trans, precision=L.BF16)
    Aop = (A.T if trans else A).astype(np.float64)
    want = Aop @ B.astype(np.float64)
    # BF16 x BF16 products are exact in FP32; the error is FP32 accumulation,
    # bounded per entry by a small multiple of sqrt(K) u32 (|A| |B|)_ij.
    bound = 8 * np.sqrt(K) * 2.0 ** -24 * (np.abs(Aop) @ np.abs(B.astype(np.float64)))
    assert np.all(np.abs(got - want) <= bound)


@pytest.mark.parametrize("M,N,K,trans", [(512, 320, 2048, True), (256, 96, 512, False),
                                         (100, 30, 72, True)])
def test_gemm_f16x3(M, N, K, trans):
    r = np.random.default_rng(11 * M + N)
    A = (r.standard_normal((K, M) if trans else (M, K)) * 1e-3).astype(np.float32)
    B = (r.standard_normal((K, N)) * 3.0).astype(np.float32)
    got = L.gemm(A, B, trans_a=trans, precision=L.F16X3)
    want = (A.T if trans else A).astype(np.float64) @ B.astype(np.float64)
    scale = np.abs(A).max() * np.abs(B).max() * np.sqrt(K)
    # ~22-bit products, FP32 accumulation
    assert np.abs(got - want).max() <= 3e-6 * scale


@pytest.mark.parametrize("M,N,K,trans", [(512, 320, 2048, True), (256, 96, 512, False),
                                         (100, 30, 72, True)])
def test_gemm_f16x2(M, N, K, trans):
    """op(A) split into FP16 hi/lo, B rounded once to FP16 after the power-of-
    two scale that brings max|B| into [2^13, 2^14]: the product of ~22-bit A
    with the rounded B, FP32 accumulation."""
    r = np.random.default_rng(13 * M + N)
    A = (r.standard_normal((K, M) if trans else (M, K)) * 1e-3).astype(np.float32)
    B = (r.standard_normal((K, N)) * 3.0).astype(np.float32)
    got = L.gemm(A, B, trans_a=trans, precision=L.F16X2)
    f, e = np.frexp(float(np.abs(B).max()))
    e = e - 1 if f == 0.5 else e
    sb = np.float32(np.ldexp(1.0, 14 - e))
    B16 = (B * sb).astype(np.float16).astype(np.float64) / float(sb)
    Aop = (A.T if trans else A).astype(np.float64)
    scale = np.abs(A).max() * np.abs(B).max() * np.sqrt(K)
    assert np.abs(got - Aop @ B16).max() <= 3e-6 * scale
    # planted check: B's rounding is visible against the exact product
    assert np.abs(got - Aop @ B.astype(np.float64)).max() > 10 * np.abs(got - Aop @ B16).max()


@pytest.mark.parametrize("M,N,K,trans", [(512, 320, 2048, True), (256, 96, 512, False),
                                         (100, 30, 72, True)])
def test_gemm_f16(M, N, K, trans):
    """op(A) and B each rounded once to FP16 after the power-of-two scales that
    bring their maxima into [2^13, 2^14]: the product of the rounded operands,
    FP32 accumulation."""
    r = np.random.default_rng(17 * M + N)
    A = (r.standard_normal((K, M) if trans else (M, K)) * 1e-3).astype(np.float32)
    B = (r.standard_normal((K, N)) * 3.0).astype(np.float32)
    got = L.gemm(A, B, trans_a=trans, precision=L.F16)

    def rounded(X):
        f, e = np.frexp(float(np.abs(X).max()))
        e = e - 1 if f == 0.5 else e
        s = np.float32(np.ldexp(1.0, 14 - e))
        return (X * s).astype(np.float16).astype(np.float64) / float(s)

    A16, B16 = rounded(A), rounded(B)
    Aop16 = A16.T if trans else A16
    scale = np.abs(A).max() * np.abs(B).max() * np.sqrt(K)
    assert np.abs(got - Aop16 @ B16).max() <= 3e-6 * scale
    # planted check: A's rounding is visible against the product with A exact
    Aop = (A.T if trans else A).astype(np.float64)
    assert np.abs(got - Aop @ B16).max() > 10 * np.abs(got - Aop16 @ B16).max()


@pytest.mark.parametrize("M,N,K,trans", [(512, 320, 2048, True), (256, 96, 512, False),
                                         (100, 30, 72, True)])
def test_gemm_f16_bsplit(M, N, K, trans):
    """op(A) rounded once to FP16, B split into FP16 hi/lo (both after their
    power-of-two scales): the product of the rounded A with ~22-bit B, FP32
    accumulation."""
    r = np.random.default_rng(19 * M + N)
    A = (r.standard_normal((K, M) if trans else (M, K)) * 1e-3).astype(np.float32)
    B = (r.standard_normal((K, N)) * 3.0).astype(np.float32)
    got = L.gemm(A, B, trans_a=trans, precision=L.F16_BSPLIT)
    f, e = np.frexp(float(np.abs(A).max()))
    e = e - 1 if f == 0.5 else e
    sa = np.float32(np.ldexp(1.0, 14 - e))
    A16 = (A * sa).astype(np.float16).astype(np.float64) / float(sa)
    Aop16 = A16.T if trans else A16
    Bd = B.astype(np.float64)
    scale = np.abs(A).max() * np.abs(B).max() * np.sqrt(K)
    assert np.abs(got - Aop16 @ Bd).max() <= 3e-6 * scale
    # planted check: A's rounding is visible against the product with A exact
    Aop = (A.T if trans else A).astype(np.float64)
    assert np.abs(got - Aop @ Bd).max() > 10 * np.abs(got - Aop16 @ Bd).max()


def test_orth_special_cases():
    # src/lora_helpers.rs:325-340 and :359-365
    np.testing.assert_allclose(L.Orth(np.zeros((5, 5))), np.eye(5), atol=1e-6)
    np.testing.assert_allclose(L.Orth(np.eye(5)), np.eye(5), atol=1e-6)
    np.testing.assert_allclose(L.Stabilizer(np.zeros((5, 5))), np.eye(5), atol=1e-6)
    x = np.random.default_rng(3).standard_normal((300, 40)) @ np.diag(np.logspace(0, 6, 40))
    q, rr = L.Orth(x, return_r=True)
    np.testing.assert_allclose(q.T @ q, np.eye(40), atol=1e-13)
    np.testing.assert_allclose(q @ rr, x, rtol=0, atol=1e-9 * np.abs(x).max())
    q2, r2 = np.linalg.qr(x)
    np.testing.assert_allclose(q, q2 * np.sign(np.diag(r2)), atol=1e-8)
    # nalgebra's convention: R has a non-negative diagonal
    assert np.all(np.diag(rr) >= 0) and np.all(np.tril(rr, -1) == 0)


def test_orth_ill_conditioned_and_wide():
    """Householder Orth, as the reference's thin QR: any conditioning, and a
    wide X gives an m x m orthogonal Q with R = Q^T X (m x n)."""
    r = np.random.default_rng(31)
    x = r.standard_normal((500, 30)) @ np.diag(np.logspace(0, -14, 30))
    q, rr = L.Orth(x, return_r=True)
    np.testing.assert_allclose(q.T @ q, np.eye(30), atol=1e-13)
    np.testing.assert_allclose(q @ rr, x, rtol=0, atol=1e-13)
    q2, r2 = np.linalg.qr(x)
    np.testing.assert_allclose(np.abs(np.diag(rr)), np.abs(np.diag(r2)), rtol=1e-6)
    for m, n in ((5, 10), (7, 64), (1, 3)):
        w = r.standard_normal((m, n))
        q, rr = L.Orth(w, return_r=True)
        assert q.shape == (m, m) and rr.shape == (m, n)
        np.testing.assert_allclose(q.T @ q, np.eye(m), atol=1e-13)
        np.testing.assert_allclose(q @ rr, w, atol=1e-13)
        q2, r2 = np.linalg.qr(w)
        sg = np.where(np.diag(r2) < 0, -1.0, 1.0)
        np.testing.assert_allclose(q, q2 * sg, atol=1e-12)
        np.testing.assert_allclose(rr, r2 * sg[:, None], atol=1e-12)


@pytest.mark.parametrize("n", [1, 17, 32, 33, 64, 100, 160, 161, 320, 321, 416, 512, 600, 1100,
                               1600])
def test_orth_cholesky_block_sizes(n):
    """Cholesky-QR across the single-launch Cholesky + inverse (n <= 512:
    32-row blocks, partial last block, up to 16 block rows; the inverse's
    K-step count changes between 160 and 161 and between 320 and 321) and the
    blocked fallback (n > 512: 64-row steps in 512-row panels, a partial last
    panel at 1100): Q orthonormal, Q R = X, R upper triangular with a positive
    diagonal (Householder QR's R up to column signs)."""
    rng = np.random.default_rng(n)
    m = max(2 * n, 64)
    x = rng.standard_normal((m, n)) @ np.diag(np.logspace(0, 3, n))
    q, rr = L.cholqr2(x)
    np.testing.assert_allclose(q.T @ q, np.eye(n), atol=1e-13)
    np.testing.assert_allclose(q @ rr, x, rtol=0, atol=1e-11 * np.abs(x).max())
    assert np.all(np.tril(rr, -1) == 0) and np.all(np.diag(rr) > 0)
    r2 = np.linalg.qr(x, mode="r")
    np.testing.assert_allclose(rr, r2 * np.sign(np.diag(r2))[:, None], rtol=0,
                               atol=1e-10 * np.abs(r2).max())


@pytest.mark.parametrize("n", [40, 100, 320, 1100])
def test_orth_dependent_columns_across_blocks(n):
    """Dependent columns in several 32-column blocks get a zero row in the
    first pass's R1 and a unit-vector completion in Q; the second pass
    re-orthogonalises the completion, so R = R2 R1 keeps an exactly zero
    diagonal entry there (R2 is upper triangular)."""
    rng = np.random.default_rng(100 + n)
    m = 3 * n
    x = rng.standard_normal((m, n))
    dep = [3, n // 2, n - 1]
    for j in dep:
        x[:, j] = x[:, :j] @ rng.standard_normal(j)
    q, rr = L.cholqr2(x)
    np.testing.assert_allclose(q.T @ q, np.eye(n), atol=1e-12)
    np.testing.assert_allclose(q @ rr, x, rtol=0, atol=1e-10 * np.abs(x).max())
    for j in dep:
        assert rr[j, j] == 0.0, j
    keep = [j for j in range(n) if j not in dep]
    assert np.all(np.diag(rr)[keep] > 0)


def test_orth_rank_deficient_completes_with_unit_vectors():
    x = np.zeros((8, 4))
    x[:, 0] = np.arange(1.0, 9.0)
    x[:, 2] = 2 * x[:, 0]
    for q in (L.Orth(x), L.cholqr2(x)[0]):
        np.testing.assert_allclose(q.T @ q, np.eye(4), atol=1e-12)
        np.testing.assert_allclose(np.abs(q[:, 0]), x[:, 0] / np.linalg.norm(x[:, 0]), atol=1e-12)


@pytest.mark.parametrize("rows,cols,attr", [(60, 200, L.ROW), (200, 60, L.COLUMN),
                                            (1000, 600, L.COLUMN)])
def test_haar_sample(rows, cols, attr):
    q = L.haar_sample(rows, cols, attr)
    if attr == L.ROW:
        np.testing.assert_allclose(q @ q.T, np.eye(rows), atol=1e-6)
        g = R.gaussian(la.reference_stream(), rows * cols).reshape(rows, cols).T
    else:
        np.testing.assert_allclose(q.T @ q, np.eye(cols), atol=1e-6)
        g = R.gaussian(la.reference_stream(), rows * cols).reshape(cols, rows).T
    qq, rr = np.linalg.qr(g)
    want = qq * np.where(np.diag(rr) < 0, -1.0, 1.0)
    np.testing.assert_allclose(q if attr == L.COLUMN else q.T, want, atol=1e-10)
    with pytest.raises(L.RandNLAError):
        L.haar_sample(cols, rows, attr)


@pytest.mark.parametrize("dist", [0, 1, 2])
def test_fused_sketch_apply(dist):
    r = np.random.default_rng(5)
    A = r.standard_normal((700, 40))
    rs, s = _stream(R.PHILOX4X32, (9, 10))
    got = L.sketch_apply(L.LEFT, dist, A, 96, rng=s)
    S = R.sketching_operator(dist, 96, 700, rs)
    np.testing.assert_allclose(got, S @ A, rtol=0, atol=1e-12 * np.sqrt(700) * 8)
    got = L.sketch_apply(L.RIGHT, dist, A, 24, rng=s)
    Om = R.sketching_operator(dist, 40, 24, rs)
    np.testing.assert_allclose(got, A @ Om, rtol=0, atol=1e-12 * 40)


@pytest.mark.parametrize("zeta,d", [(1, 128), (5, 128), (8, 128), (12, 128), (16, 128),
                                    (17, 128), (64, 128), (12, 10)])
def test_sparse_sign_apply(zeta, d):
    # zeta <= 8 and <= 16 take the register kernels, larger zeta the generic
    # one; zeta > d clips to d nonzeros per column
    r = np.random.default_rng(6)
    A = r.standard_normal((2000, 33))
    rs, s = _stream(R.PHILOX4X32, (1, 2))
    got = L.sketch_apply(L.LEFT, L.SPARSE_SIGN, A, d, rng=s, zeta=zeta)
    S = R.sketching_operator(3, d, 2000, rs, zeta=zeta)
    np.testing.assert_allclose(got, S @ A, rtol=0, atol=1e-12)
    # f32 input path
    got32 = L.sketch_apply(L.LEFT, L.SPARSE_SIGN, A.astype(np.float32), d, rng=s, zeta=zeta)
    np.testing.assert_allclose(got32, S @ A.astype(np.float32).astype(np.float64), atol=1e-12)


def test_test_matrix_matches_oracle():
    sig = np.logspace(0, -4, 64)
    rs, s = _stream(R.PHILOX4X32, (77, 78))
    got = L.test_matrix(256, 64, sig, noise=1e-3, rng=s, dtype=L.F64, device=False)
    want = la.hadamard_test_matrix(256, 64, sig, 1e-3, rs)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)
    clean = L.test_matrix(256, 64, sig, noise=0.0, rng=s, dtype=L.F64, device=False)
    np.testing.assert_allclose(np.linalg.svd(clean, compute_uv=False), sig, rtol=1e-12)
