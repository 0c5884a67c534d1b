"""CPU: pins for the linear-algebra oracle (closed forms, LAPACK / scipy
cross-checks, the reference's special cases).  Each check is chosen so that a
plausible mistake in the restatement (dropped term, wrong sign or index,
transposed operand) fails it."""
import numpy as np
import pytest
import scipy.linalg as sla
from scipy.sparse.linalg import lsqr as scipy_lsqr

from oracle import la, rng as R


def _rs(seed):
    return np.random.default_rng(seed)


# ---- solvers.rs ----------------------------------------------------------
def test_upper_triangular_kats(golden):
    k = golden["solvers_kat"]
    for name in ("upper_basic", "upper_identity"):
        c = k[name]
        x = la.solve_upper_triangular_system(np.array(c["u"], float), np.array(c["y"], float))
        np.testing.assert_allclose(x.ravel(), c["x"], atol=1e-10)
    c = k["upper_zero_diag"]
    u = np.array(c["u"], float)
    x = la.solve_upper_triangular_system(u, np.array(c["y"], float))
    assert abs((u @ x)[2, 0] - c["x_row2"]) < 1e-10
    assert x[1, 0] == 0.0          # zero pivot leaves x_i = 0
    c = k["upper_f32"]
    x = la.solve_upper_triangular_system(np.array(c["u"], np.float32), np.array(c["y"], np.float32))
    assert x.dtype == np.float32
    np.testing.assert_allclose(x.ravel(), c["x"], atol=1e-6)


def test_lsqr_kat(golden):
    c = golden["solvers_kat"]["lsqr_simple"]
    x = la.lsqr_dense(np.array(c["a"], float), np.array(c["b"], float))[0]
    np.testing.assert_allclose(x, c["x"], atol=c["tol"])


@pytest.mark.parametrize("damp,seed", [(0.0, 0), (0.0, 1), (0.5, 2)])
def test_lsqr_matches_scipy(damp, seed):
    # Same algorithm as scipy's (src/solvers.rs:111): identical iterates.
    r = _rs(seed)
    a = r.standard_normal((200, 100)) * np.logspace(0, 3, 100)
    b = r.standard_normal(200)
    mine = la.lsqr_dense(a, b, damp=damp, atol=1e-10, btol=1e-10, iter_lim=300, calc_var=True)
    ref = scipy_lsqr(a, b, damp=damp, atol=1e-10, btol=1e-10, iter_lim=300, calc_var=True)
    assert mine[1] == ref[1] and mine[2] == ref[2]
    np.testing.assert_allclose(mine[0], ref[0], rtol=1e-9, atol=1e-12)
    for i in (3, 4, 5, 6, 8):
        np.testing.assert_allclose(mine[i], ref[i], rtol=1e-9)
    np.testing.assert_allclose(mine[9], ref[9], rtol=1e-9)


def test_lsqr_x0_and_zero_rhs():
    r = _rs(3)
    a = r.standard_normal((50, 20))
    x, istop, itn, r1, *_ = la.lsqr_dense(a, np.zeros(50))
    assert istop == 0 and itn == 0 and not x.any()
    b = r.standard_normal(50)
    x0 = r.standard_normal(20)
    mine = la.lsqr_dense(a, b, x0=x0, atol=1e-12, btol=1e-12)
    ref = scipy_lsqr(a, b, x0=x0, atol=1e-12, btol=1e-12)
    np.testing.assert_allclose(mine[0], ref[0], rtol=1e-9)


def test_cgls_converges_to_lstsq():
    r = _rs(4)
    a = r.standard_normal((120, 30))
    b = r.standard_normal(120)
    x, it, conv = la.cgls(lambda v: a @ v, lambda u: a.T @ u, 30, b, 1e-10, 200)
    assert conv and it <= 60
    np.testing.assert_allclose(x, np.linalg.lstsq(a, b, rcond=None)[0], rtol=1e-8)
    # stops on the absolute ||A^T r|| test (src/cg.rs:46)
    assert np.linalg.norm(a.T @ (b - a @ x)) < 1e-10


# ---- lora_helpers.rs -----------------------------------------------------
def test_orth_special_cases():
    np.testing.assert_allclose(la.orth(np.zeros((5, 5))), np.eye(5), atol=1e-12)
    np.testing.assert_allclose(np.abs(la.orth(np.eye(5))), np.eye(5), atol=1e-12)
    q = la.orth(_rs(5).standard_normal((10, 5)))
    np.testing.assert_allclose(q.T @ q, np.eye(5), atol=1e-12)


def test_full_pivot_lu():
    x = _rs(6).standard_normal((9, 4))
    lu, p, q = la.full_piv_lu(x)
    L = np.tril(lu[:, :4], -1) + np.eye(9, 4)
    U = np.triu(lu[:4, :])
    np.testing.assert_allclose(x[p][:, q], L @ U, atol=1e-12)
    # the first pivot is the entry of largest magnitude
    i, j = np.unravel_index(np.argmax(np.abs(x)), x.shape)
    assert p[0] == i and q[0] == j
    assert np.all(np.abs(L) <= 1 + 1e-12)
    np.testing.assert_array_equal(la.stabilizer(np.zeros((5, 5))), np.eye(5))


def test_tsog1_even_branch_discards_gaussian():
    a = _rs(7).standard_normal((30, 12))
    s = la.tsog1(a, 4, 2, 1)
    # A*0 -> Stabilizer(0) = [I;0] -> A^T [I;0] = A[:4].T -> Stabilizer
    np.testing.assert_allclose(s, la.stabilizer(a[:4, :].T), atol=1e-14)
    assert la.tsog1(a, 3, 3, 1).shape == (12, 3)


def test_rand_svd_faithful_exact_rank():
    r = _rs(8)
    a = r.standard_normal((40, 6)) @ r.standard_normal((6, 30))
    U, S, Vt = la.rand_svd(a, 6, 1e-6, 2)
    np.testing.assert_allclose(np.diag(S), np.linalg.svd(a, compute_uv=False)[:6], rtol=1e-10)
    np.testing.assert_allclose(U @ S @ Vt, a, atol=1e-10)


def test_rand_svd_zero_matrix_kat():
    # src/lora_drivers.rs:341-357
    U, S, Vt = la.rand_svd(np.zeros((10, 5)), 3, 1e-6, 2)
    np.testing.assert_allclose(np.abs(U), np.eye(10, 3), atol=1e-12)
    np.testing.assert_allclose(S, np.zeros((3, 3)))
    np.testing.assert_allclose(np.abs(Vt.T), np.eye(5, 3), atol=1e-12)


def test_rand_svd_invalid():
    for k, e, s in ((0, 1e-6, 2), (3, 0.0, 2), (3, 1e-6, 0)):
        with pytest.raises(ValueError):
            la.rand_svd(np.ones((6, 5)), k, e, s)


def test_rand_svd_power_closed_form():
    # C1 construction at reduced size: ||A - A_k||_2 = sigma_{k+1}
    m, n, k, p = 300, 200, 20, 10
    sig = np.r_[10.0 ** (-np.arange(k) / 10), np.full(n - k, 1e-8)]
    a = la.lowrank_matrix(m, n, sig, R.Stream.from_seed(1, 1), R.Stream.from_seed(1, 2))
    np.testing.assert_allclose(np.linalg.svd(a, compute_uv=False), np.sort(sig)[::-1], rtol=1e-9,
                               atol=1e-14)
    om = R.gaussian(R.Stream.from_seed(1, 42), n * (k + p)).reshape(k + p, n).T
    U, S, Vt = la.rand_svd_power(a, k, p, 2, om)
    err = np.linalg.norm(a - U @ S @ Vt, 2)
    assert err < 2e-8


def test_cholqr2_matches_householder():
    y = _rs(9).standard_normal((500, 40)) @ np.diag(np.logspace(0, 4, 40))
    q, r = la.cholqr2(y)
    q2, r2 = np.linalg.qr(y)
    sg = np.sign(np.diag(r2))
    np.testing.assert_allclose(q, q2 * sg, atol=1e-10)
    np.testing.assert_allclose(q.T @ q, np.eye(40), atol=1e-13)
    np.testing.assert_allclose(q @ r, y, rtol=1e-12, atol=1e-10)


# ---- sketch_and_precondition.rs --------------------------------------------
def _ls_problem(m, n, seed, cond=1e3):
    r = _rs(seed)
    q, _ = np.linalg.qr(r.standard_normal((m, n)))
    v, _ = np.linalg.qr(r.standard_normal((n, n)))
    a = (q * np.logspace(0, -np.log10(cond), n)) @ v.T
    xs = r.standard_normal(n)
    b = a @ xs + 1e-6 * r.standard_normal(m)
    return a, b


@pytest.mark.parametrize("fn", [la.blendenpik_overdetermined, la.lsrn_overdetermined])
def test_preconditioned_ls_matches_lstsq(fn):
    a, b = _ls_problem(400, 20, 10)
    x, it = fn(a, b, 1e-12, 200, 2.0)
    np.testing.assert_allclose(x, np.linalg.lstsq(a, b, rcond=None)[0], rtol=1e-8, atol=1e-10)
    assert it < 60


def test_saddle_point_closed_forms():
    a, b = _ls_problem(300, 15, 11, cond=10)
    c = _rs(12).standard_normal(15)
    # mu = 0: stationarity of ||Ax-b||^2 + 2<c,x>: A^T A x = A^T b - c.
    x, y = la.sketch_saddle_point_precondition(a, b, c, 0.0, 1e-13, 500, 3.0)
    np.testing.assert_allclose(x, np.linalg.solve(a.T @ a, a.T @ b - c), rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(y, b - a @ x, atol=1e-12)
    # mu > 0: src/sketch_and_precondition.rs:185-212 uses mu only in the
    # preconditioner and in the shift of b; the CGLS objective has no mu term,
    # so x = A^+ (b - S^T U diag(1/sqrt(s^2+mu)) V^T c).
    mu = 0.3
    x, y = la.sketch_saddle_point_precondition(a, b, c, mu, 1e-13, 500, 3.0)
    d = int(np.floor(3.0 * 15))
    s = la.sketch(0, d, 300)
    u, sg, vt = np.linalg.svd(s @ a, full_matrices=False)
    w = s.T @ (u @ ((vt @ c) / np.sqrt(sg ** 2 + mu)))
    np.testing.assert_allclose(x, np.linalg.lstsq(a, b - w, rcond=None)[0], rtol=1e-7, atol=1e-9)


def test_ls_error_paths():
    a = np.ones((3, 5))
    with pytest.raises(ValueError):
        la.blendenpik_overdetermined(a, np.ones(3), 1e-6, 10, 2.0)
    with pytest.raises(ValueError):
        la.lsrn_overdetermined(np.ones((10, 3)), np.ones(10), 1e-6, 10, 0.5)


# ---- pivot_decompositions.rs, cqrrpt.rs -------------------------------------
def test_qrcp_invariants():
    a = _rs(13).standard_normal((60, 12)) * np.linspace(1, 5, 12)
    q, r, p = la.qrcp(a)
    np.testing.assert_allclose(q.T @ q, np.eye(60), atol=1e-12)
    np.testing.assert_allclose(q @ r, a[:, p], atol=1e-12)
    assert np.allclose(np.tril(r, -1), 0, atol=1e-12)
    d = np.abs(np.diag(r))
    assert np.all(d[:-1] >= d[1:] - 1e-12)        # pivoting orders |R_ii|
    # first pivot = the column of largest norm
    assert p[0] == np.argmax(np.linalg.norm(a, axis=0))


def test_cqrrpt_invariants():
    a = R.sketching_operator(1, 100, 10, la.reference_stream())
    q, r, j, k = la.sap_chol_qrcp(a, 30)
    assert k == 10
    np.testing.assert_allclose(q.T @ q, np.eye(10), atol=1e-10)
    assert np.allclose(np.tril(r, -1), 0, atol=1e-10)
    np.testing.assert_allclose(q @ r, a[:, j], atol=1e-10)
    # R equals the (positive-diagonal) R of QR(A[:, J])
    r2 = np.linalg.qr(a[:, j])[1]
    np.testing.assert_allclose(np.abs(r), np.abs(r2), atol=1e-10)
    with pytest.raises(AssertionError, match="d must satisfy n ≤ d ≪ m"):
        la.sap_chol_qrcp(a, 5)


def test_hadamard_test_matrix_singular_values():
    sig = np.logspace(0, -3, 16)
    a = la.hadamard_test_matrix(64, 16, sig, 0.0, R.Stream.from_seed(1, 7))
    np.testing.assert_allclose(np.linalg.svd(a, compute_uv=False), sig, rtol=1e-12)


def test_blendenpik_lsqr_oracle_closed_form():
    """The batched path's per-problem algorithm: sketch-QR preconditioned
    LSQR from z0 = Q^T S b reaches the unique least-squares solution."""
    r = np.random.default_rng(12)
    for m, n in [(500, 20), (64, 16), (40, 30)]:
        a = r.standard_normal((m, n)) * 10.0 ** (4 * r.random(n))
        b = r.standard_normal(m)
        d = m if 2 * n > m else 2 * n
        S = R.batch_gauss_sketch(R.Stream.from_seed(R.PHILOX4X32, 42), 0, d, m)
        x, istop, itn = la.blendenpik_lsqr(a, b, 1e-12, 400, 2.0, S)
        xs = np.linalg.lstsq(a, b, rcond=None)[0]
        assert istop in (1, 2) and 0 < itn < 100
        assert np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)
    # Preconditioning works: kappa(A) = 1e4 but kappa(A R^-1) is O(1), so a
    # nearly consistent system converges in a few tens of iterations, and an
    # identity "preconditioner" (a planted mistake: R dropped) does not.
    a = r.standard_normal((800, 30)) * 10.0 ** (4 * r.random(30))
    b = a @ r.standard_normal(30) + 1e-3 * r.standard_normal(800)
    S = R.batch_gauss_sketch(R.Stream.from_seed(R.PHILOX4X32, 42), 0, 60, 800)
    _, _, itn_good = la.blendenpik_lsqr(a, b, 1e-12, 400, 2.0, S)
    assert itn_good < 40
    itn_plain = la.lsqr_dense(a, b, atol=1e-12, btol=1e-12, iter_lim=400)[2]
    assert itn_plain > 3 * itn_good


def test_solve_diagonal_system():
    x = la.solve_diagonal_system([2.0, 0.0, -4.0], [1.0, 5.0, 2.0])
    assert np.array_equal(x, [0.5, 0.0, -0.5])          # zero diagonal -> 0, no inf


def test_sketched_least_squares_match_lstsq():
    r = np.random.default_rng(50)
    a = r.standard_normal((400, 12)) @ np.diag(np.logspace(0, -3, 12))
    b = r.standard_normal(400)
    s = r.standard_normal((100, 400))                  # m // 4 = 100 rows
    want = np.linalg.lstsq(s @ a, s @ b, rcond=None)[0]
    for fn in (la.sketched_least_squares_qr, la.sketched_least_squares_svd):
        np.testing.assert_allclose(fn(a, b, s_op=s), want, rtol=1e-10, atol=1e-12)
    # a planted mistake (Q^T left out) is far from the solution
    q, rr = np.linalg.qr(s @ a)
    bad = np.asarray(la.solve_upper_triangular_system(rr, (s @ b)[:12])).reshape(-1)
    assert np.linalg.norm(bad - want) > 1e-3 * np.linalg.norm(want)


def test_sketched_least_squares_wide_sketch_and_zero_column():
    r = np.random.default_rng(51)
    # m // 4 = 5 < n = 8: QR solves the leading 5 x 5 block (rest 0), SVD gives
    # the minimum-norm solution of the sketched system
    a = r.standard_normal((20, 8)); b = r.standard_normal(20); s = r.standard_normal((5, 20))
    xq = la.sketched_least_squares_qr(a, b, s_op=s)
    np.testing.assert_allclose(xq[:5], np.linalg.solve((s @ a)[:, :5], s @ b), rtol=1e-10)
    assert np.array_equal(xq[5:], np.zeros(3))
    np.testing.assert_allclose(la.sketched_least_squares_svd(a, b, s_op=s),
                               np.linalg.lstsq(s @ a, s @ b, rcond=None)[0], rtol=1e-9, atol=1e-12)
    # a zero column gives R_jj == 0 exactly: x_j = 0 and the triangular
    # system's other rows hold (row j's equation is dropped, as in
    # solve_upper_triangular_system; this is not the lstsq of the reduced
    # problem)
    a = r.standard_normal((200, 6)); a[:, 3] = 0.0; b = r.standard_normal(200)
    s = r.standard_normal((50, 200))
    x = la.sketched_least_squares_qr(a, b, s_op=s)
    q, rr = np.linalg.qr(s @ a)
    z = q.T @ (s @ b)
    keep = [0, 1, 2, 4, 5]
    assert rr[3, 3] == 0.0 and x[3] == 0.0
    np.testing.assert_allclose(rr[np.ix_(keep, keep)] @ x[keep], z[keep], rtol=1e-10, atol=1e-12)
    with pytest.raises(ValueError):
        la.sketched_least_squares_qr(np.ones((3, 1)), np.ones(3))


# ---- lora_drivers.rs: rand_evd1 / rand_evd2 ---------------------------------
def _sym_with_spectrum(d, seed):
    q, _ = np.linalg.qr(_rs(seed).standard_normal((len(d), len(d))))
    return (q * np.asarray(d, float)) @ q.T, q


def test_rand_evd1_closed_form_full_basis():
    # With Q spanning everything (Q = I), Rayleigh-Ritz is exact: lambda is
    # the known spectrum ordered by |lambda| (signs kept), V its eigenvectors.
    d = [1.0, -5.0, 3.0, 0.5, -0.25, 2.0]
    a, q = _sym_with_spectrum(d, 1)
    a = (a + a.T) / 2
    v, lam = la.rand_evd1(a, 3, 1e-6, 3, q=np.eye(6))
    np.testing.assert_allclose(lam, [-5.0, 3.0, 2.0], rtol=1e-12)
    for j, col in enumerate((1, 2, 5)):
        assert abs(abs(v[:, j] @ q[:, col]) - 1.0) < 1e-12
    np.testing.assert_allclose(a @ v, v * lam, atol=1e-12)


def test_rand_evd1_zero_matrix_kat():
    # src/lora_drivers.rs:532-549: V = I[:, :3], lambda all exactly 0
    v, lam = la.rand_evd1(np.zeros((5, 5)), 3, 1e-6, 2)
    np.testing.assert_array_equal(np.abs(v), np.eye(5)[:, :3])
    assert np.all(lam == 0.0) and lam.shape == (3,)


def test_rand_evd1_rayleigh_ritz_and_errors():
    r = _rs(2)
    x = r.standard_normal((40, 40))
    a = (x + x.T) / 2
    q, _ = np.linalg.qr(r.standard_normal((40, 8)))
    v, lam = la.rand_evd1(a, 5, 1e-6, 3, q=q)
    # residual orthogonal to range(Q); V inside range(Q), orthonormal
    np.testing.assert_allclose(q.T @ (a @ v - v * lam), 0.0, atol=1e-12)
    np.testing.assert_allclose(v.T @ v, np.eye(5), atol=1e-12)
    assert np.all(np.diff(np.abs(lam)) <= 0)
    # planted mistake: the eigenpairs of Q^T Q instead of Q^T A Q fail the residual
    w, u = np.linalg.eigh(q.T @ q)
    assert np.abs(q.T @ (a @ (q @ u[:, :5]) - (q @ u[:, :5]) * w[:5])).max() > 1e-3
    bad = np.array([[1.0, 2.0, 3.0], [100.0, 200.0, 4.0], [3.0, 38.0, 1.0]])   # :505-513
    with pytest.raises(ValueError, match="NotHermitian"):
        la.rand_evd1(bad, 3, 1e-6, 2)
    for args in ((0, 1e-6, 2), (3, 0.0, 2), (3, 1e-6, 0)):
        with pytest.raises(ValueError, match="InvalidParameters"):
            la.rand_evd1(a, *args)


def test_rand_evd2_closed_form_and_nystrom():
    d = [9.0, 4.0, 2.5, 1.0, 0.5, 0.1]
    a, q = _sym_with_spectrum(d, 3)
    a = (a + a.T) / 2
    # square invertible S: the Nystrom approximation of A + nu I is exact
    s_sq = _rs(4).standard_normal((6, 6))
    v, lam = la.rand_evd2(a, 3, 3, s_mat=s_sq)
    np.testing.assert_allclose(lam, [9.0, 4.0, 2.5], rtol=1e-10)
    for j in range(3):
        assert abs(abs(v[:, j] @ q[:, j]) - 1.0) < 1e-10
    # planted mistake: sigma instead of sigma^2 misses the spectrum
    assert abs(np.sqrt(lam[0]) - 9.0) > 1.0
    # l < n: eigenvalues of (AS)(S^T A S)^{-1}(AS)^T, invariant under S -> S M
    r = _rs(5)
    x = r.standard_normal((30, 30))
    a = x @ x.T
    sm = r.standard_normal((30, 6))
    v, lam = la.rand_evd2(a, 4, 2, s_mat=sm)
    ys = a @ sm
    nys = ys @ np.linalg.solve(sm.T @ ys, ys.T)
    w = np.sort(np.linalg.eigvalsh((nys + nys.T) / 2))[::-1]
    np.testing.assert_allclose(lam, w[:4], rtol=1e-8)
    v2, lam2 = la.rand_evd2(a, 4, 2, s_mat=sm @ r.standard_normal((6, 6)))
    np.testing.assert_allclose(lam2, lam, rtol=1e-7)
    np.testing.assert_allclose(np.abs(np.sum(v * v2, axis=0)), 1.0, atol=1e-7)


def test_rand_evd2_kats():
    bad = np.array([[1.0, 2.0, 3.0], [100.0, 200.0, 4.0], [3.0, 38.0, 1.0]])   # :698-706
    with pytest.raises(ValueError, match="NotPositiveSemiDefinite"):
        la.rand_evd2(bad, 3, 2)
    with pytest.raises(ValueError, match="MatrixDecompositionError"):          # :725-733
        la.rand_evd2(np.zeros((5, 5)), 3, 2, s_mat=_rs(6).standard_normal((5, 5)))
    v, lam = la.rand_evd2(np.eye(5), 3, 2, s_mat=_rs(7).standard_normal((5, 5)))   # :735-746
    np.testing.assert_allclose(lam, 1.0, rtol=1e-12)
    with pytest.raises(ValueError, match="InvalidParameters"):
        la.rand_evd2(np.eye(5), 0, 2)


# ---- pivot_decompositions.rs economic_qrcp, id.rs ----------------------------
def _rank_k(m, n, k, seed):
    r = _rs(seed)
    return r.standard_normal((m, k)) @ r.standard_normal((k, n))


def test_economic_qrcp_pivot_order_is_greedy():
    # Each step picks the column with the largest residual after projecting
    # out the columns chosen so far (least-squares residuals, an independent
    # statement of the recomputed-norm pivot rule).
    a = _rs(11).standard_normal((30, 20)) * np.logspace(0, 2, 20)
    k = 12
    q, r, p = la.economic_qrcp(a, k)
    chosen = []
    for _ in range(k):
        best, arg = -1.0, None
        for j in range(20):
            if j in chosen:
                continue
            res = a[:, j] if not chosen else a[:, j] - a[:, chosen] @ np.linalg.lstsq(
                a[:, chosen], a[:, j], rcond=None)[0]
            if np.linalg.norm(res) > best:
                best, arg = np.linalg.norm(res), j
        chosen.append(arg)
    assert list(p[:k]) == chosen
    np.testing.assert_allclose(q @ r[:, :k], a[:, p[:k]], atol=1e-10 * np.abs(a).max())
    np.testing.assert_allclose(q.T @ q, np.eye(k), atol=1e-13)
    assert np.all(np.diff(np.abs(np.diag(r[:, :k]))) <= 1e-12)
    assert np.all(np.tril(r[:, :k], -1) == 0.0) or np.abs(np.tril(r[:, :k], -1)).max() < 1e-12
    with pytest.raises(AssertionError, match=r"k must be <= min\(m,n\)"):
        la.economic_qrcp(a, 21)
    with pytest.raises(AssertionError, match="k must be positive"):
        la.economic_qrcp(a, 0)


def test_osid_closed_forms():
    a = _rank_k(70, 50, 9, 12)
    x, j = la.osid_qrcp(a, 9, la.COLUMN)
    np.testing.assert_allclose(x[:, j], np.eye(9), atol=0)
    np.testing.assert_allclose(x, np.linalg.pinv(a[:, j]) @ a, atol=1e-9)
    np.testing.assert_allclose(a[:, j] @ x, a, atol=1e-11 * np.abs(a).max())
    x, i = la.osid_qrcp(a, 9, la.ROW)
    np.testing.assert_allclose(x[i, :], np.eye(9), atol=0)
    np.testing.assert_allclose(x, a @ np.linalg.pinv(a[i, :]), atol=1e-9)
    # planted mistake: X without the pivot scatter ([I T] in pivot order)
    xc, jc = la.osid_qrcp(a, 9, la.COLUMN)
    _, rr, p = la.economic_qrcp(a, 9)
    unscattered = np.hstack([np.eye(9), np.linalg.solve(np.triu(rr[:, :9]), rr[:, 9:])])
    assert np.abs(a[:, jc] @ unscattered - a).max() > 1e-3
    # randomised column ID through a Gaussian S: an ID of A itself
    s = _rs(13).standard_normal((9, 70))
    x, j = la.osid_randomised(a, 9, la.COLUMN, s_mat=s)
    np.testing.assert_allclose(a[:, j] @ x, a, atol=1e-10 * np.abs(a).max())
    for f, args in ((la.osid_qrcp, (a, 0, la.COLUMN)), (la.osid_randomised, (a, 0, la.COLUMN))):
        with pytest.raises(AssertionError, match="k must be positive"):
            f(*args)
    with pytest.raises(AssertionError, match=r"k must be <= min\(l,w\)"):
        la.osid_qrcp(a, 51, la.COLUMN)


@pytest.mark.parametrize("shape", [(80, 60), (60, 80)])
def test_cur_and_two_sided_closed_forms(shape):
    a = _rank_k(*shape, 10, 14)
    j, u, i = la.cur(a, 10)
    c, r = a[:, j], a[i, :]
    np.testing.assert_allclose(u, np.linalg.pinv(c) @ a @ np.linalg.pinv(r), atol=1e-8)
    np.testing.assert_allclose(c @ u @ r, a, atol=1e-10 * np.abs(a).max())
    z, ii, jj, x = la.two_sided_id(a, 10)
    np.testing.assert_allclose(z @ a[np.ix_(ii, jj)] @ x, a, atol=1e-10 * np.abs(a).max())
    # planted mistake: the core from the transposed factors
    assert np.abs(c @ u.T @ r - a).max() > 1e-3
    with pytest.raises(AssertionError, match="k must be positive"):
        la.cur(a, 0)
