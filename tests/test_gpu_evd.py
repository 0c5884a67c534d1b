"""GPU parity: rand_evd1 / rand_evd2 (src/lora_drivers.rs:87-224) against the
oracle on the GPU path's own range basis / test matrix, closed-form spectra
and the reference's own tests (src/lora_drivers.rs:480-870)."""
import numpy as np
import pytest

import randnla_amd as L
from oracle import la

pytestmark = pytest.mark.gpu


def _rs(seed):
    return np.random.default_rng(seed)


def hermitian(n, seed):
    """generate_random_hermitian_matrix (src/test_assist.rs:132-140)."""
    x = _rs(seed).standard_normal((n, n))
    return 0.5 * (x + x.T)


def psd(n, seed):
    """generate_random_psd_matrix (src/test_assist.rs:143-151)."""
    x = _rs(seed).standard_normal((n, n))
    return x @ x.T


def known_spectrum(d, seed):
    q, _ = np.linalg.qr(_rs(seed).standard_normal((len(d), len(d))))
    a = (q * np.asarray(d, float)) @ q.T
    return (a + a.T) / 2, q


def assert_same_columns(v, w, tol):
    """Columns equal up to sign."""
    s = np.sign(np.sum(v * w, axis=0))
    np.testing.assert_allclose(v * s, w, atol=tol)


def _kind(e):
    return e.value.kind


# ------------------------------------------------------------------ rand_evd1
@pytest.mark.parametrize("n,k,s", [(100, 3, 2), (100, 10, 5), (300, 20, 10)])
def test_rand_evd1_matches_oracle(n, k, s):
    a = hermitian(n, n + k)
    V, lam = L.rand_evd1(a, k, 1e-6, s)
    q = L.RF1(a, min(k + s, n))            # the same Q (same stream and passes)
    Vo, lamo = la.rand_evd1(a, k, 1e-6, s, q=q)
    assert V.shape == (n, k) and lam.shape == (k,)
    np.testing.assert_allclose(lam, lamo, rtol=0, atol=1e-11 * np.abs(lamo).max())
    assert_same_columns(V, Vo, 1e-9)
    np.testing.assert_allclose(V.T @ V, np.eye(k), atol=1e-12)
    assert np.all(np.diff(np.abs(lam)) <= 0)
    # the reference's bound (:659-684) with epsilon = 2
    err = np.linalg.norm(a - (V * lam) @ V.T) / np.linalg.norm(a)
    assert err <= 2.0


def test_rand_evd1_full_basis_signed_pairs():
    # l = n: exact eigenpairs.  +-3 and +-1 pairs have equal |lambda|; an SVD
    # of Q^T A Q could not separate them, the shifted Jacobi must.
    d = [3.0, -3.0, 1.0, -1.0, 0.5, -0.2, 2.0, 0.0]
    a, _ = known_spectrum(d, 11)
    V, lam = L.rand_evd1(a, 8, 1e-6, 1)
    np.testing.assert_allclose(np.sort(lam), np.sort(d), atol=1e-13)
    assert np.all(np.diff(np.abs(lam)) <= 0)
    np.testing.assert_allclose(a @ V, V * lam, atol=1e-13)
    np.testing.assert_allclose(V.T @ V, np.eye(8), atol=1e-13)


def test_rand_evd1_reference_cases():
    # zero matrix (:532-549): V = I[:, :3] exactly, lambda all exactly 0
    V, lam = L.rand_evd1(np.zeros((5, 5)), 3, 1e-6, 2)
    np.testing.assert_array_equal(V, np.eye(5)[:, :3])
    assert lam.shape == (3,) and np.all(lam == 0.0)
    # identity (:551-563)
    V, lam = L.rand_evd1(np.eye(5), 3, 1e-6, 2)
    assert V.shape == (5, 3) and np.allclose(lam, 1.0, atol=1e-14)
    # PSD 5 x 5, k = 3, s = 1 (:565-615): the top-3 eigenvectors of a
    # 4-dimensional Rayleigh-Ritz space; eigenvalues bounded by A's
    a = psd(5, 3)
    V, lam = L.rand_evd1(a, 3, 0.01, 1)
    w = np.sort(np.linalg.eigvalsh(a))[::-1]
    assert np.all(lam <= w[:3] * (1 + 1e-12)) and np.all(lam >= w[-1] - 1e-12)
    # not Hermitian (:505-514), NaN, non-square
    bad = np.array([[1.0, 2.0, 3.0], [100.0, 200.0, 4.0], [3.0, 38.0, 1.0]])
    for m in (bad, np.full((4, 4), np.nan), np.ones((4, 3))):
        with pytest.raises(L.RandNLAError) as e:
            L.rand_evd1(m, 3, 1e-6, 2)
        assert _kind(e) == "NotHermitian"
    # one asymmetric entry far from the diagonal of a larger matrix
    a = hermitian(200, 4)
    a[170, 3] += 1e-14
    with pytest.raises(L.RandNLAError) as e:
        L.rand_evd1(a, 3, 1e-6, 2)
    assert _kind(e) == "NotHermitian"
    for args in ((0, 1e-6, 2), (3, 0.0, 2), (3, 1e-6, 0)):
        with pytest.raises(L.RandNLAError) as e:
            L.rand_evd1(np.eye(5), *args)
        assert _kind(e) == "InvalidParameters"


# ------------------------------------------------------------------ rand_evd2
@pytest.mark.parametrize("n,k,s", [(100, 3, 2), (100, 10, 5), (100, 3, 0), (400, 16, 8)])
def test_rand_evd2_matches_oracle(n, k, s):
    a = psd(n, n + 7 * k)
    V, lam = L.rand_evd2(a, k, s)
    sm = L.tsog1(a, k + s, 3, 1)           # the same S (same stream and passes)
    Vo, lamo = la.rand_evd2(a, k, s, s_mat=sm)
    assert V.shape == Vo.shape == (n, k) and lam.shape == (k,)
    np.testing.assert_allclose(lam, lamo, rtol=1e-10)
    assert_same_columns(V, Vo, 1e-8)
    np.testing.assert_allclose(V.T @ V, np.eye(k), atol=1e-12)
    assert np.all(np.diff(lam) <= 0)
    err = np.linalg.norm(a - (V * lam) @ V.T) / np.linalg.norm(a)   # :817-829
    assert err <= 2.0


def test_rand_evd2_comparison_with_deterministic_evd():
    # :748-776: k + s = n, so the Nystrom approximation is exact
    for seed in range(3):
        a = psd(5, 20 + seed)
        V, lam = L.rand_evd2(a, 3, 2)
        w, u = np.linalg.eigh(a)
        idx = np.argsort(-np.abs(w))[:3]
        np.testing.assert_allclose(lam, w[idx], rtol=0, atol=1e-6)
        np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-6)
        ud = u[:, idx]
        np.testing.assert_allclose(V @ V.T @ a, ud @ ud.T @ a, atol=1e-6)


def test_rand_evd2_known_spectrum_large():
    # decaying known spectrum floored at 1e-8 (so that the reference's
    # eigenvalue sign test accepts it too), k + s < n: the top eigenvalues to
    # 1e-9 (three passes of tsog1 resolve them to ~1e-11)
    n = 600
    d = np.maximum(10.0 ** (-np.arange(n) / 20.0), 1e-8)
    a, q = known_spectrum(d, 5)
    V, lam = L.rand_evd2(a, 20, 20)
    np.testing.assert_allclose(lam[:10], d[:10], rtol=1e-9)
    np.testing.assert_allclose(np.abs(np.sum(V[:, :5] * q[:, :5], axis=0)), 1.0, atol=1e-8)


def test_rand_evd2_reference_cases():
    bad = np.array([[1.0, 2.0, 3.0], [100.0, 200.0, 4.0], [3.0, 38.0, 1.0]])   # :698-706
    with pytest.raises(L.RandNLAError) as e:
        L.rand_evd2(bad, 3, 0)
    assert _kind(e) == "NotPositiveSemiDefinite"
    # an indefinite matrix whose lower triangle is the only one read
    a = -psd(50, 9)
    with pytest.raises(L.RandNLAError) as e:
        L.rand_evd2(a, 3, 2)
    assert _kind(e) == "NotPositiveSemiDefinite"
    with pytest.raises(L.RandNLAError) as e:                                   # :725-733
        L.rand_evd2(np.zeros((5, 5)), 3, 2)
    assert _kind(e) == "MatrixDecompositionError"
    V, lam = L.rand_evd2(np.eye(5), 3, 2)                                      # :735-746
    assert V.shape == (5, 3)
    np.testing.assert_allclose(lam, 1.0, rtol=1e-13)
    np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-13)
    with pytest.raises(L.RandNLAError) as e:
        L.rand_evd2(np.eye(5), 0, 2)
    assert _kind(e) == "InvalidParameters"


def test_rand_evd_device_inputs():
    torch = pytest.importorskip("torch")
    a = psd(256, 31)
    at = torch.from_numpy(a).cuda()
    V, lam = L.rand_evd2(at, 8, 8)
    Vh, lamh = L.rand_evd2(a, 8, 8)
    assert V.is_cuda and lam.is_cuda
    np.testing.assert_array_equal(lam.cpu().numpy(), lamh)
    np.testing.assert_array_equal(V.cpu().numpy(), Vh)
    V1, lam1 = L.rand_evd1(at, 8, 1e-6, 8)
    V1h, lam1h = L.rand_evd1(a, 8, 1e-6, 8)
    np.testing.assert_array_equal(lam1.cpu().numpy(), lam1h)
    np.testing.assert_array_equal(V1.cpu().numpy(), V1h)


def test_rand_evd_rank_above_n():
    # rand_evd1: l = min(k + s, n), r = min(k, l) (src/lora_drivers.rs:117, :140)
    a = hermitian(6, 41)
    V, lam = L.rand_evd1(a, 9, 1e-6, 2)
    w = np.linalg.eigvalsh(a)
    assert V.shape == (6, 6) and lam.shape == (6,)
    np.testing.assert_allclose(np.sort(lam), np.sort(w), atol=1e-12)
    # rand_evd2 needs k + s <= n here (the reference's S^T (A S) is then singular)
    with pytest.raises(L.RandNLAError) as e:
        L.rand_evd2(psd(6, 42), 5, 2)
    assert _kind(e) == "InvalidParameters"
    V, lam = L.rand_evd2(psd(6, 42), 4, 2)
    np.testing.assert_allclose(lam, np.sort(np.linalg.eigvalsh(psd(6, 42)))[::-1][:4], rtol=1e-9)
