"""CPU: the RNG oracle and the library's host RNG exports against the
reference's known-answer tests (rust-random123/src/philox.rs:242-363,
threefry.rs:97-142)."""
import numpy as np
import pytest
from scipy import stats

from oracle import rng as R


def test_philox4x32_kat(golden):
    k = golden["rng_kat"]
    key = (k["philox_seed1"], k["philox_seed2"])
    ctr = np.array([[i, 0, 0, 0] for i in range(10)], dtype=np.uint64)
    out = R.philox4x32(ctr, key).reshape(-1)
    assert [int(x) for x in out] == k["philox4x32"]


def test_philox_other_widths_kat(golden):
    k = golden["rng_kat"]
    s1, s2 = k["philox_seed1"], k["philox_seed2"]
    got2x32 = sum((R.philox2x32([i, 0], s1) for i in range(10)), [])
    got2x64 = sum((R.philox2x64([i, 0], s1) for i in range(10)), [])
    got4x64 = sum((R.philox4x64([i, 0, 0, 0], (s1, s2)) for i in range(10)), [])
    assert got2x32 == k["philox2x32"]
    assert got2x64 == k["philox2x64"]
    assert got4x64 == k["philox4x64"]


def test_threefry_kat(golden):
    k = golden["rng_kat"]
    ctr = np.array([[i, 0] for i in range(10)], dtype=np.uint64)
    out = R.threefry2x64(ctr, k["threefry_key"]).reshape(-1)
    assert [int(x) for x in out] == k["threefry2x64"]


def test_seed_from_u64_vector(golden):
    # threefry.rs:140-141: ThreeFryRng::seed_from_u64(42).next_u64()
    s = R.Stream.from_seed(R.THREEFRY2X64, 42)
    assert int(s.words64(0, 1)[0]) == golden["rng_kat"]["threefry_seed42_first_u64"]
    d = golden["rng_kat"]["derived"]
    assert list(R.Stream.from_seed(R.THREEFRY2X64, 0).key) == d["threefry_seed0_key"]
    assert list(R.Stream.from_seed(R.PHILOX4X32, 42).key) == d["philox_seed42_key"]


def test_stream_counter_carry():
    # next() carries ctr[0] overflow into ctr[1] (threefry.rs:16-19)
    s = R.Stream(R.THREEFRY2X64, (1, 2), (2 ** 64 - 1, 5))
    b = s.blocks([0, 1])
    want0 = R.threefry2x64(np.array([[2 ** 64 - 1, 5]], dtype=np.uint64), (1, 2))[0]
    want1 = R.threefry2x64(np.array([[0, 6]], dtype=np.uint64), (1, 2))[0]
    assert (b[0] == want0).all() and (b[1] == want1).all()
    p = R.Stream(R.PHILOX4X32, (7, 9), (2 ** 64 - 1, 0))
    pb = p.blocks([1])[0]
    o = R.philox4x32(np.array([[0, 0, 1, 0]], dtype=np.uint64), (7, 9))[0]
    assert int(pb[0]) == int(o[0]) | (int(o[1]) << 32)


def test_gaussian_box_muller_invariants():
    # z0^2 + z1^2 = -2 ln u1 and atan2(z1, z0) = 2 pi u2 (mod 2 pi), u from the words.
    s = R.Stream.from_seed(R.PHILOX4X32, 42)
    z = R.gaussian(s, 2000)
    blk = s.blocks(np.arange(1000))
    u1 = ((blk[:, 0] >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (blk[:, 1] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    np.testing.assert_allclose(z[0::2] ** 2 + z[1::2] ** 2, -2 * np.log(u1), rtol=1e-12)
    ang = np.mod(np.arctan2(z[1::2], z[0::2]), 2 * np.pi)
    d = np.abs(ang - 2 * np.pi * u2)
    assert np.all(np.minimum(d, 2 * np.pi - d) < 1e-9)


def test_gaussian_distribution():
    z = R.gaussian(R.Stream.from_seed(R.PHILOX4X32, 42), 200000)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-3


def test_uniform_and_rademacher():
    s = R.Stream.from_seed(R.THREEFRY2X64, 0)
    u = R.uniform_pm1(s, 50000)
    w = s.words64(0, 50000)
    np.testing.assert_array_equal(u, 2 * (w >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 - 1)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.02
    r = R.rademacher(s, 640)
    w = s.words64(0, 10)
    bits = np.array([(int(w[e // 64]) >> (e % 64)) & 1 for e in range(640)])
    np.testing.assert_array_equal(r, np.where(bits == 1, 1.0, -1.0))


def test_sparse_sign_invariants():
    s = R.Stream.from_seed(R.PHILOX4X32, 3)
    d, m, z = 64, 300, 8
    S = R.sketching_operator(3, d, m, s, zeta=z)
    nnz = (S != 0).sum(axis=0)
    assert (nnz == z).all()
    np.testing.assert_allclose(np.abs(S[S != 0]), 1 / np.sqrt(z))
    np.testing.assert_allclose(np.linalg.norm(S, axis=0), 1.0)
    # E[S^T S] = I: off-diagonal mean near 0
    G = S.T @ S
    assert abs(G[~np.eye(m, dtype=bool)].mean()) < 0.02


def test_sketching_operator_zero_dims():
    with pytest.raises(ValueError):
        R.sketching_operator(0, 0, 5, R.Stream.from_seed(0, 0))


def test_library_host_rng_kat(golden):
    import randnla_amd as L
    k = golden["rng_kat"]
    key = (k["philox_seed1"], k["philox_seed2"])
    assert sum((L.philox4x32([i, 0, 0, 0], key) for i in range(10)), []) == k["philox4x32"]
    assert sum((L.philox2x32([i, 0], key[0]) for i in range(10)), []) == k["philox2x32"]
    assert sum((L.philox2x64([i, 0], key[0]) for i in range(10)), []) == k["philox2x64"]
    assert sum((L.philox4x64([i, 0, 0, 0], key) for i in range(10)), []) == k["philox4x64"]
    assert sum((L.threefry2x64([i, 0], k["threefry_key"]) for i in range(10)), []) == k["threefry2x64"]
    r = L.rng_from_seed(L.PHILOX4X32, 42)
    assert [r.key[0], r.key[1]] == k["derived"]["philox_seed42_key"]
    r0 = L.rng_from_seed(L.THREEFRY2X64, 0)
    assert [r0.key[0], r0.key[1]] == k["derived"]["threefry_seed0_key"]


def test_batch_sketch_stream_layout_and_distribution():
    """The batched path's sketch: entry (r, k) of substream id comes from
    Philox4x32 block r*ceil(m/4) + k//4 at counter + ((id+1) << 96) (words
    pinned by the Philox KATs above), Box-Muller on 24-bit uniforms."""
    s = R.Stream.from_seed(R.PHILOX4X32, 42)
    d, m = 40, 1003                     # m not a multiple of 4
    S = R.batch_gauss_sketch(s, 5, d, m)
    assert S.shape == (d, m)
    # Block layout: row r, columns 4q..4q+3 come from one block, independent
    # of how many rows are drawn.
    np.testing.assert_array_equal(R.batch_gauss_sketch(s, 5, 3, m), S[:3])
    # The substream offset is a counter offset of (id+1) << 96: the same
    # entries from stream id 5 equal stream id 0 of a stream whose counter
    # high word is 5 << 32 larger.
    s2 = R.Stream(R.PHILOX4X32, s.key, (s.ctr[0], s.ctr[1] + (5 << 32)))
    np.testing.assert_array_equal(R.batch_gauss_sketch(s2, 0, d, m), S)
    # Distinct substreams are uncorrelated; entries are standard normal.
    T = R.batch_gauss_sketch(s, 6, d, m)
    assert abs(np.corrcoef(S.ravel(), T.ravel())[0, 1]) < 0.02
    z = R.batch_gauss_sketch(s, 0, 200, 4000).ravel()
    assert stats.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.01
    # 24-bit uniforms: u in (0, 1) strictly, so every entry is finite.
    assert np.isfinite(z).all() and np.abs(z).max() < 6.0
