"""Pins the FP8-e4m3 oracle (oracle/fp8.py) against an independent
implementation of the OCP format (torch.float8_e4m3fn on the CPU), brute
force over the code table, and exact products."""
import numpy as np
import torch

from oracle import fp8 as F


def test_table_matches_torch_e4m3fn():
    codes = torch.arange(256, dtype=torch.int32).to(torch.uint8)
    ref = codes.view(torch.float8_e4m3fn).to(torch.float64).numpy()
    t = F.e4m3_table()
    assert np.array_equal(np.isnan(t), np.isnan(ref))
    ok = ~np.isnan(t)
    assert np.array_equal(t[ok], ref[ok])
    assert np.nanmax(t) == 448.0 and np.nanmin(np.abs(t[t != 0])) == 2.0 ** -9


def test_rounding_matches_brute_force_and_torch():
    t = F.e4m3_table()
    vals = np.unique(t[~np.isnan(t)])
    rng = np.random.default_rng(0)
    # random magnitudes over the whole range, every representable value,
    # and every midpoint between neighbours (ties)
    x = np.concatenate([np.exp(rng.uniform(np.log(2.0 ** -12), np.log(448.0), 20000)),
                        vals, 0.5 * (vals[1:] + vals[:-1])])
    x = np.concatenate([x, -x]).astype(np.float32)
    got = F.e4m3_round(x)
    # brute force: nearest table value, ties to the even code (mantissa lsb 0)
    pos = np.array(sorted(set(np.abs(vals))))
    codes = {v: c for c, v in enumerate(t) if not np.isnan(v) and c < 128}
    a = np.abs(x.astype(np.float64))
    d = np.abs(pos[None, :] - a[:, None])
    best = d.min(axis=1)
    want = np.empty_like(a)
    for i in range(len(a)):
        cand = pos[d[i] == best[i]]
        want[i] = cand[0] if len(cand) == 1 else [c for c in cand if codes[c] % 2 == 0][0]
    want = np.copysign(want, x)
    assert np.array_equal(got, want)
    tq = torch.from_numpy(x).to(torch.float8_e4m3fn).to(torch.float64).numpy()
    assert np.array_equal(got, tq)


def test_rounding_planted_mistake_is_caught():
    """Truncation instead of round-to-nearest differs from the brute force
    on a value just below a grid point."""
    x = np.float32(1.0 + 0.124)          # grid spacing 1/8 near 1
    assert F.e4m3_round(x) == 1.125
    assert np.floor(x * 8) / 8 == 1.0


def test_e8m0_scale():
    assert F.e8m0_for(0.0) == 127
    assert F.e8m0_for(448.0) == 127
    assert F.e8m0_for(448.0 * 1.0001) == 128
    assert F.e8m0_for(1.0) == 127 - 8            # 1 * 2^8 = 256 <= 448, 512 > 448
    assert F.e8m0_for(448.0 * 2.0 ** -20) == 107
    rng = np.random.default_rng(1)
    for a in np.exp(rng.uniform(-60, 60, 500)):
        e = F.e8m0_for(a) - 127
        assert a * 2.0 ** -e <= 448.0 < a * 2.0 ** -(e - 1)


def test_scaled_gemm_exact_on_representable_blocks():
    """Blocks whose entries are e4m3 values times one power of two (with the
    block maximum 448 * 2^k) quantise without error: the scaled GEMM equals
    the FP64 product exactly, for both op(A)."""
    rng = np.random.default_rng(2)
    t = F.e4m3_table()
    vals = t[(~np.isnan(t))]

    def mat(r, c, rb, cb):
        X = rng.choice(vals, size=(r, c))
        for bi in range(0, r, rb):
            for bj in range(0, c, cb):
                X[bi, bj] = 448.0
                X[bi:bi + rb, bj:bj + cb] *= 2.0 ** int(rng.integers(-20, 10))
        return X

    A = mat(300, 260, 128, 128)
    B = mat(260, 70, 128, 1)
    assert np.array_equal(F.fp8_gemm(A, B), A @ B)
    At = mat(260, 300, 128, 128)
    assert np.array_equal(F.fp8_gemm(At, B, trans_a=True), At.T @ B)
    # a planted mistake: one scale per 128 x 128 block of op(A) taken from the
    # wrong block (transposed block index) breaks the exactness
    qa, E = F.quantize_blocks(At, 128, 128)
    assert E.shape == (3, 3) and not np.array_equal(E, E.T)


def test_scaled_gemm_error_bound():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((256, 384)).astype(np.float32)
    B = rng.standard_normal((384, 16)).astype(np.float32)
    C = F.fp8_gemm(A, B)
    exact = A.astype(np.float64) @ B.astype(np.float64)
    # each operand carries relative error <= 2^-4 per entry
    bound = (2 * 2.0 ** -4 + 2.0 ** -8) * (np.abs(A) @ np.abs(B))
    assert np.all(np.abs(C - exact) <= bound)
    assert np.abs(C - exact).max() > 1e-3 * np.abs(exact).max()   # it is FP8, not FP32
