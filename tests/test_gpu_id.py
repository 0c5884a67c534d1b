"""GPU parity: interpolative decompositions and CUR (src/id.rs) and
economic_qrcp (src/pivot_decompositions.rs:194-269) against the oracle on the
same inputs and sketches, closed forms on rank-k matrices, and the
reference's own tests (src/id.rs:319-560)."""
import numpy as np
import pytest

import randnla_amd as L
from oracle import la

pytestmark = pytest.mark.gpu


def _rs(seed):
    return np.random.default_rng(seed)


def rank_k(m, n, k, seed):
    """rank_k_matrix (src/test_assist.rs:7-30): a sum of k Gaussian rank-1 terms."""
    r = _rs(seed)
    return r.standard_normal((m, k)) @ r.standard_normal((k, n))


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("m,n,k", [(40, 30, 10), (30, 40, 30), (300, 200, 64)])
def test_economic_qrcp_matches_oracle(m, n, k):
    a = _rs(m + n).standard_normal((m, n)) * np.logspace(0, 1, n)
    Q, R, p = L.economic_qrcp(a, k)
    Qo, Ro, po = la.economic_qrcp(a, k)
    np.testing.assert_array_equal(p[:k], po[:k])
    np.testing.assert_allclose(R, Ro, rtol=0, atol=1e-11 * np.abs(Ro).max())
    np.testing.assert_allclose(Q, Qo, rtol=0, atol=1e-11)


@pytest.mark.parametrize("attr", [L.COLUMN, L.ROW])
def test_osid_qrcp_matches_oracle(attr):
    a = _rs(3).standard_normal((90, 70))
    X, J = L.osid_qrcp(a, 25, attr)
    Xo, Jo = la.osid_qrcp(a, 25, attr)
    np.testing.assert_array_equal(J, Jo)
    np.testing.assert_allclose(X, Xo, rtol=0, atol=1e-10 * np.abs(Xo).max())


def test_osid_randomised_matches_oracle():
    a = _rs(4).standard_normal((120, 80))
    k = 20
    X, J = L.osid_randomised(a, k, L.COLUMN)
    S = L.sketching_operator(0, k, 120)                 # the same Gaussian stream
    Xo, Jo = la.osid_randomised(a, k, la.COLUMN, s_mat=S)
    np.testing.assert_array_equal(J, Jo)
    np.testing.assert_allclose(X, Xo, rtol=0, atol=1e-9 * np.abs(Xo).max())
    # Row branch: defined for n == k (the reference multiplies A by S^T)
    b = _rs(5).standard_normal((150, k))
    X, I = L.osid_randomised(b, k, L.ROW)
    S = L.tsog1(b, k, 2, 1)                             # the same tsog1 stream
    Xo, Io = la.osid_randomised(b, k, la.ROW, s_mat=S)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_allclose(X, Xo, rtol=0, atol=1e-9 * np.abs(Xo).max())
    with pytest.raises(L.RandNLAPanic, match="dimensions mismatch"):
        L.osid_randomised(a, k, L.ROW)


@pytest.mark.parametrize("shape", [(110, 90), (90, 110)])
def test_cur_and_two_sided_match_oracle(shape):
    a = _rs(shape[0]).standard_normal(shape)
    k = 30
    J, U, I = L.cur(a, k)
    Jo, Uo, Io = la.cur(a, k)
    np.testing.assert_array_equal(J, Jo)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_allclose(U, Uo, rtol=0, atol=1e-8 * np.abs(Uo).max())
    Z, I, J, X = L.two_sided_id(a, k)
    Zo, Io, Jo, Xo = la.two_sided_id(a, k)
    np.testing.assert_array_equal(J, Jo)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_allclose(Z, Zo, rtol=0, atol=1e-9 * np.abs(Zo).max())
    np.testing.assert_allclose(X, Xo, rtol=0, atol=1e-9 * np.abs(Xo).max())


@pytest.mark.parametrize("seed", range(3))
def test_reference_rank_k_reconstructions(seed):
    # src/id.rs:450-559: m, n in [100, 110), k in [min/2, min), rank-k input
    r = _rs(100 + seed)
    m, n = int(r.integers(100, 110)), int(r.integers(100, 110))
    k = int(r.integers(min(m, n) // 2, min(m, n)))
    a = rank_k(m, n, k, seed)
    X, J = L.osid_qrcp(a, k, L.COLUMN)
    assert rel(a[:, J] @ X, a) < 1e-10
    X, J = L.osid_randomised(a, k, L.COLUMN)
    assert rel(a[:, J] @ X, a) < 1e-9
    for f in (L.two_sided_id, L.two_sided_id_randomised):
        Z, I, J, X = f(a, k)
        assert rel(Z @ a[np.ix_(I, J)] @ X, a) < 1e-8
    for f in (L.cur, L.cur_randomised):
        J, U, I = f(a, k)
        assert rel(a[:, J] @ U @ a[I, :], a) < 1e-8
    at = a.T.copy()
    for f in (L.cur, L.cur_randomised):
        J, U, I = f(at, k)
        assert rel(at[:, J] @ U @ at[I, :], at) < 1e-8


def test_reference_panics():
    # src/id.rs:327-446: k = 0 and k in (n, m) on a 100 x 10 matrix
    data = L.sketching_operator(1, 100, 10)
    for f in (lambda d: L.osid_qrcp(data, d, L.COLUMN),
              lambda d: L.osid_randomised(data, d, L.COLUMN),
              lambda d: L.cur(data, d), lambda d: L.cur_randomised(data, d),
              lambda d: L.two_sided_id(data, d), lambda d: L.two_sided_id_randomised(data, d)):
        with pytest.raises(L.RandNLAPanic, match="k must be positive"):
            f(0)
        with pytest.raises(L.RandNLAPanic, match=r"k must be <= min\(l,w\)"):
            f(int(_rs(7).integers(11, 100)))
    with pytest.raises(L.RandNLAPanic, match=r"k must be <= min\(m,n\)"):
        L.economic_qrcp(data, 11)
    with pytest.raises(L.RandNLAPanic, match="k must be positive"):
        L.economic_qrcp(data, 0)


def test_cur_randomised_large_device():
    torch = pytest.importorskip("torch")
    a = rank_k(3000, 2000, 96, 21)
    at = torch.from_numpy(a).cuda()
    J, U, I = L.cur_randomised(at, 96)
    assert J.is_cuda and U.is_cuda
    Jh, Uh, Ih = J.cpu().numpy(), U.cpu().numpy(), I.cpu().numpy()
    assert rel(a[:, Jh] @ Uh @ a[Ih, :], a) < 1e-8
    Z, I, J, X = L.two_sided_id_randomised(at, 96)
    Zh, Ih, Jh, Xh = (t.cpu().numpy() for t in (Z, I, J, X))
    assert rel(Zh @ a[np.ix_(Ih, Jh)] @ Xh, a) < 1e-8


def test_economic_qrcp_ties_and_wide():
    # equal largest norms (a column and its negation): the first one wins
    a = _rs(8).standard_normal((50, 40))
    a[:, 3] *= 10.0
    a[:, 7] = -a[:, 3]
    _, _, p = L.economic_qrcp(a, 20)
    _, _, po = la.economic_qrcp(a, 20)
    assert p[0] == 3
    np.testing.assert_array_equal(p[:20], po[:20])
    # far more columns than threads in the pivot search
    b = _rs(9).standard_normal((64, 20000)) * np.exp(_rs(10).uniform(0, 2, 20000))
    Q, R, p = L.economic_qrcp(b, 32)
    Qo, Ro, po = la.economic_qrcp(b, 32)
    np.testing.assert_array_equal(p[:32], po[:32])
    np.testing.assert_allclose(R[:, :32], Ro[:, :32], rtol=0, atol=1e-11 * np.abs(Ro).max())


def test_cur_ill_conditioned_core():
    # sigma_i = 10^(-i/8): A[I, :] and A[:, J] are ill-conditioned (kappa ~ 1e10);
    # the backward-stable pseudo-inverse keeps the reconstruction at the
    # oracle's (SVD-based) level
    m, n, k = 300, 200, 80
    r = _rs(31)
    u, _ = np.linalg.qr(r.standard_normal((m, n)))
    v, _ = np.linalg.qr(r.standard_normal((n, n)))
    a = (u * 10.0 ** (-np.arange(n) / 8.0)) @ v.T
    for mat in (a, a.T.copy()):
        J, U, I = L.cur(mat, k)
        Jo, Uo, Io = la.cur(mat, k)
        e = rel(mat[:, J] @ U @ mat[I, :], mat)
        eo = rel(mat[:, Jo] @ Uo @ mat[Io, :], mat)
        assert e <= 10 * eo + 1e-12, (e, eo)


@pytest.mark.parametrize("shape", [(140, 100), (100, 140)])
def test_randomised_cur_and_two_sided_match_oracle(shape):
    # the same Gaussian sketch (the dense-sketch stream) and the same tsog1
    # basis as the GPU path, fed to the oracle's restatement
    a = _rs(shape[1]).standard_normal(shape)
    m, n = shape
    k = 24
    J, U, I = L.cur_randomised(a, k)
    if m >= n:
        S = L.sketching_operator(0, k, m)
        osid = lambda y, kk, attr: la.osid_randomised(y, kk, attr, s_mat=S)
    else:
        S = L.sketching_operator(0, k, n)                  # the sketch of A^T
        osid = lambda y, kk, attr: la.osid_randomised(y, kk, attr, s_mat=S)
    Jo, Uo, Io = la.cur(a, k, osid=osid)
    np.testing.assert_array_equal(J, Jo)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_allclose(U, Uo, rtol=0, atol=1e-8 * np.abs(Uo).max())
    Z, I, J, X = L.two_sided_id_randomised(a, k)
    S = L.sketching_operator(0, k, m)
    Xo, Jo = la.osid_randomised(a, k, la.COLUMN, s_mat=S)
    np.testing.assert_array_equal(J, Jo)
    np.testing.assert_allclose(X, Xo, rtol=0, atol=1e-9 * np.abs(Xo).max())
    St = L.tsog1(a[:, J], k, 2, 1)
    Zo, Io = la.osid_randomised(a[:, J], k, la.ROW, s_mat=St)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_allclose(Z, Zo, rtol=0, atol=1e-9 * np.abs(Zo).max())
