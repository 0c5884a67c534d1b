"""GPU parity: rand_svd and its helpers (src/lora_drivers.rs,
src/lora_helpers.rs) against the oracle, the configurations' closed forms and
the reference's own special cases."""
import numpy as np
import pytest

import randnla_amd as L
from oracle import la, rng as R

pytestmark = pytest.mark.gpu


def c1_matrix(m=2000, n=1000):
    """Plain rand_svd configuration: sigma_i = 10^(-i/10) for i < 50, then 1e-8;
    U, V from QR of seeded Gaussians."""
    sig = np.r_[10.0 ** (-np.arange(50) / 10.0), np.full(n - 50, 1e-8)]
    a = la.lowrank_matrix(m, n, sig, R.Stream.from_seed(R.PHILOX4X32, 1),
                          R.Stream.from_seed(R.PHILOX4X32, 2))
    return a, sig


def omega(n, l):
    z = R.gaussian(R.Stream.from_seed(R.PHILOX4X32, 42), n * l)
    return z.reshape(l, n).T


def test_rand_svd_c1_fp64_matches_oracle():
    a, sig = c1_matrix()
    k, p, q = 50, 10, 2
    U, sv, Vt = L.rand_svd_factors(a, k, 1e-6, p, precision=L.FP64, power_iters=q)
    Uo, So, Vto = la.rand_svd_power(a, k, p, q, omega(1000, k + p))
    assert U.shape == (2000, 50) and sv.shape == (50,) and Vt.shape == (50, 1000)
    # same Omega stream, same method: singular values agree to rounding
    np.testing.assert_allclose(sv, np.diag(So), rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(sv, sig[:50], rtol=1e-7)
    # closed form: ||A - A_50||_2 = sigma_51 = 1e-8; target <= 1e-6 relative
    err = np.linalg.norm(a - (U * sv) @ Vt, 2)
    assert err <= 1e-6 * sig[0] and err < 3e-8
    np.testing.assert_allclose((U * sv) @ Vt, Uo @ So @ Vto, rtol=0, atol=1e-12)
    np.testing.assert_allclose(U.T @ U, np.eye(50), atol=1e-12)
    np.testing.assert_allclose(Vt @ Vt.T, np.eye(50), atol=1e-12)
    assert np.all(np.diff(sv) <= 0)


def test_rand_svd_c1_against_reference_faithful_path():
    # The reference's own rand_svd (deterministic in A, src/lora_helpers.rs:70-72)
    a, sig = c1_matrix(600, 300)
    Uf, Sf, Vtf = la.rand_svd(a, 50, 1e-6, 10)
    U, sv, Vt = L.rand_svd_factors(a, 50, 1e-6, 10)
    # both methods resolve sigma to well below 1e-6 * sigma_1
    np.testing.assert_allclose(sv, np.diag(Sf), rtol=0, atol=1e-9)
    ef = np.linalg.norm(a - Uf @ Sf @ Vtf, 2)
    eg = np.linalg.norm(a - (U * sv) @ Vt, 2)
    assert eg <= 1e-6 and ef <= 1e-6


@pytest.mark.parametrize("m,n,k,s", [(20, 10, 3, 2), (5, 10, 3, 2), (100, 80, 45, 2), (64, 64, 8, 4)])
def test_rand_svd_reference_shapes(m, n, k, s):
    a = np.random.default_rng(m * n).standard_normal((m, n))
    U, S, Vt = L.rand_svd(a, k, 1e-6, s)
    assert U.shape == (m, k) and S.shape == (k, k) and Vt.shape == (k, n)
    d = np.diag(S)
    assert np.all(d[:-1] >= d[1:])
    np.testing.assert_allclose(U.T @ U, np.eye(k), atol=1e-10)
    # Ritz values bound the true singular values from below
    s_true = np.linalg.svd(a, compute_uv=False)
    assert np.all(d <= s_true[:k] * (1 + 1e-12)) and d[0] >= 0.8 * s_true[0]


def test_rand_svd_zero_matrix_kat():
    # src/lora_drivers.rs:341-357
    U, S, Vt = L.rand_svd(np.zeros((10, 5)), 3, 1e-6, 2)
    np.testing.assert_allclose(U, np.eye(10, 3), atol=1e-6)
    np.testing.assert_allclose(S, np.zeros((3, 3)), atol=1e-6)
    np.testing.assert_allclose(Vt.T, np.eye(5, 3), atol=1e-6)


def test_rand_svd_identity():
    U, S, Vt = L.rand_svd(np.eye(10, 5), 3, 1e-6, 2)
    np.testing.assert_allclose(np.diag(S), np.ones(3), atol=1e-12)
    np.testing.assert_allclose((U * np.diag(S)) @ Vt @ Vt.T, U * np.diag(S), atol=1e-12)


@pytest.mark.parametrize("k,eps,s", [(0, 1e-6, 2), (3, 0.0, 2), (3, -1.0, 2), (3, 1e-6, 0)])
def test_rand_svd_invalid_parameters(k, eps, s):
    with pytest.raises(L.RandNLAError) as e:
        L.rand_svd(np.ones((10, 5)), k, eps, s)
    assert e.value.kind == "InvalidParameters"


def test_rand_svd_mixed_precision_tracks_fp64():
    # Polynomial-decay spectrum + noise floor, FP32 storage (scaled-down
    # mixed-precision configuration).
    import torch
    m, n, k, p = 4096, 1024, 64, 16
    sig = (1.0 + np.arange(n)) ** -2.0
    A = L.test_matrix(m, n, sig, noise=1e-7, rng=L.rng(L.PHILOX4X32, (5, 6)), dtype=L.F32)
    res = {}
    for prec in (L.FP64, L.BF16, L.FP8):
        U, sv, Vt = L.rand_svd_factors(A, k, 1e-6, p, precision=prec, power_iters=3)
        torch.cuda.synchronize()
        e, an = L.spectral_error(A, U, sv, Vt, iters=30)
        res[prec] = (e / an, sv.cpu().numpy())
    e64, s64 = res[L.FP64]
    for prec in (L.BF16, L.FP8):
        el, sl = res[prec]
        assert abs(el - e64) <= 1e-6, (prec, el, e64)
        assert np.max(np.abs(sl - s64)) / s64[0] <= 1e-5, prec


def test_qb1_rf1_tsog1():
    r = np.random.default_rng(21)
    a = r.standard_normal((120, 6)) @ r.standard_normal((6, 50)) + 1e-3 * r.standard_normal((120, 50))
    Q, B = L.QB1(a, 10, 1e-6)
    assert Q.shape == (120, 10) and B.shape == (10, 50)
    np.testing.assert_allclose(Q.T @ Q, np.eye(10), atol=1e-12)
    np.testing.assert_allclose(B, Q.T @ a, atol=1e-12)
    assert np.linalg.norm(a - Q @ B) < np.linalg.norm(a)
    # same Omega and method as the oracle: identical subspace
    Qo = la.orth(a @ la.orth(a.T @ la.orth(a @ omega(50, 10))))
    np.testing.assert_allclose(Q @ Q.T, Qo @ Qo.T, atol=1e-10)
    Q2 = L.RF1(a, 10)
    np.testing.assert_allclose(Q2 @ Q2.T, Q @ Q.T, atol=1e-12)
    for passes in (2, 3, 4):
        S = L.tsog1(a, 3, passes, 1)
        assert S.shape == (50, 3)
        np.testing.assert_allclose(S.T @ S, np.eye(3), atol=1e-10)
