"""FP64 linear-algebra oracle (test infrastructure only, see ``oracle/__init__``).

numpy restatement of the reference's sketch-and-factor path.  Dense kernels
that the reference takes from nalgebra 0.33.0 (third-party, not vendored:
Householder QR, SVD, Cholesky) are taken from LAPACK through numpy; nalgebra's
full-pivot LU is restated by hand because its pivot order decides the
reference's `Stabilizer` output.

Function                      follows                                   pinned by
----------------------------  ----------------------------------------  ------------------------------------------
solve_upper_triangular_system src/solvers.rs:22-41                      closed-form KATs src/solvers.rs:290-338
sym_ortho, lsqr               src/solvers.rs:77-102, :115-278           scipy.sparse.linalg.lsqr (same algorithm,
                                                                        scipy 1.14.1 origin, src/solvers.rs:111),
                                                                        KAT src/solvers.rs:391-410
cgls                          src/cg.rs:18-61                           lstsq on small systems (closed form)
orth                          src/lora_helpers.rs:131-133               Orth(0)=I, Orth(I)=I :325-340; QtQ=I
full_piv_lu_l / stabilizer    src/lora_helpers.rs:144-146 (nalgebra     Stabilizer(0)=I :359-365; P A Q = L U
                              FullPivLU, column-major first-max pivot)  reconstruction
tsog1 (faithful)              src/lora_helpers.rs:58-105                shape tests :160-233; even-branch discard
rf1, qb1                      src/lora_helpers.rs:17-23, :37-44         |A - QB| vs closed form
rand_svd (faithful)           src/lora_drivers.rs:30-69                 zero-matrix KAT :341-357; sigma vs LAPACK
rand_evd1                     src/lora_drivers.rs:87-151                zero-matrix KAT :532-549; closed-form
                                                                        spectrum with Q = I (|lambda| order,
                                                                        signs); Rayleigh-Ritz residual
                                                                        Q^T (A V - V L) = 0
rand_evd2                     src/lora_drivers.rs:167-224               KATs :698-706 (not PSD), :725-733 (zero
                                                                        matrix fails); closed-form spectrum
                                                                        with square S; Nystrom formula
                                                                        (AS)(S^T A S)^{-1}(AS)^T for l < n
rand_svd_power (intended)     the library's method (Gaussian Omega,     |A - U S Vt|_2 vs sigma_{k+1} (closed form
                              q power iterations)                       for matrices of known spectrum)
blendenpik / lsrn / saddle    src/sketch_and_precondition.rs:26-216     unique LS / ridge solution (lstsq)
solve_diagonal_system         src/solvers.rs:56-70                      closed form (zero diagonal -> 0)
sketched_least_squares_qr/svd src/sketch_and_solve.rs:24-33, :53-65     lstsq of the sketched problem (min-norm
                                                                        when m / 4 < n); planted zero column
qrcp                          src/pivot_decompositions.rs:106-175       reconstruction, orthogonality :320-348
economic_qrcp                 src/pivot_decompositions.rs:194-269       pivot order = greedy max-residual column
                                                                        (least-squares residuals); A[:, p[:k]] =
                                                                        Q R[:, :k]; |R_ii| non-increasing
osid_qrcp / osid_randomised / src/id.rs:33-305                          rank-k matrices: X = pinv(A[:, J]) A,
two_sided_id / cur                                                      A = A[:, J] X = Z A[I, J] X =
                                                                        A[:, J] U A[I, :] with U = pinv(C) A
                                                                        pinv(R); panic messages :327-446
sap_chol_qrcp                 src/cqrrpt.rs:27-58                       invariants src/cqrrpt.rs:95-109
cholqr2                       no reference counterpart                  QtQ = I, Y = QR, matches Householder up to
                                                                        column signs
"""
import numpy as np

from . import rng as R


# --------------------------------------------------------------- solvers
def solve_upper_triangular_system(u, y):
    """Back substitution that leaves x_i = 0 where u_ii == 0
    (src/solvers.rs:22-41)."""
    u = np.asarray(u)
    y = np.asarray(y).reshape(-1)
    n = u.shape[0]
    x = np.zeros(n, dtype=u.dtype)
    for i in range(n - 1, -1, -1):
        if u[i, i] != 0:
            s = u.dtype.type(0)
            for j in range(i + 1, n):
                s += u[i, j] * x[j]
            x[i] = (y[i] - s) / u[i, i]
    return x.reshape(-1, 1)


def solve_diagonal_system(dg, y):
    """x_i = y_i / d_i, and 0 where d_i == 0 (src/solvers.rs:56-70)."""
    dg = np.asarray(dg, dtype=np.float64).reshape(-1)
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    x = np.zeros_like(dg)
    nz = dg != 0.0
    x[nz] = y[nz] / dg[nz]
    return x


def _sign(x):
    """src/solvers.rs:77-81 (0 maps to 0)."""
    return 1.0 if x > 0 else (-1.0 if x < 0 else 0.0)


def sym_ortho(a, b):
    """Stable Givens rotation (src/solvers.rs:84-102)."""
    if b == 0.0:
        return _sign(a), 0.0, abs(a)
    if a == 0.0:
        return 0.0, _sign(b), abs(b)
    if abs(b) > abs(a):
        tau = a / b
        s = _sign(b) / np.sqrt(1.0 + tau * tau)
        c = s * tau
        return c, s, b / s
    tau = b / a
    c = _sign(a) / np.sqrt(1.0 + tau * tau)
    s = c * tau
    return c, s, a / c


def lsqr(a_mul, at_mul, m, n, b, damp=0.0, atol=1e-8, btol=1e-8, conlim=1e8,
         iter_lim=None, calc_var=False, x0=None):
    """LSQR (src/solvers.rs:115-278) on operator callbacks.

    Returns (x, istop, itn, r1norm, r2norm, anorm, acond, arnorms, xnorm, var)
    exactly as the reference's 10-tuple."""
    eps = np.finfo(np.float64).eps
    if iter_lim is None:
        iter_lim = 2 * n
    var = np.zeros(n)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64).reshape(-1)
    u = np.array(b, dtype=np.float64).reshape(-1)
    bnorm = np.linalg.norm(u)
    if x0 is not None:
        u = u - a_mul(x)
        beta = np.linalg.norm(u)
    else:
        beta = bnorm
    u = u * (1.0 / beta if beta > 0 else 0.0)
    v = at_mul(u)
    alfa = np.linalg.norm(v)
    v = v * (1.0 / alfa if alfa > 0 else 0.0)
    w = v.copy()
    rhobar, phibar = alfa, beta
    rnorm = r1norm = r2norm = beta
    anorm = acond = 0.0
    dampsq = damp * damp
    ddnorm = res2 = xnorm = xxnorm = z = 0.0
    cs2, sn2 = -1.0, 0.0
    arnorm = alfa * beta
    if arnorm == 0.0:
        return x, 0, 0, beta, beta, 0.0, 0.0, [0.0], 0.0, var
    ctol = 1.0 / conlim if conlim > 0 else 0.0
    arnorms = []
    itn = 0
    istop = 0
    while itn < iter_lim:
        arnorms.append(arnorm)
        itn += 1
        u = a_mul(v) - alfa * u
        beta = np.linalg.norm(u)
        if beta > 0.0:
            u = u * (1.0 / beta)
            anorm = np.sqrt(anorm ** 2 + alfa ** 2 + beta ** 2 + dampsq)
            v = at_mul(u) - beta * v
            alfa = np.linalg.norm(v)
            v = v * (1.0 / alfa if alfa > 0 else 0.0)
        rhobar1 = np.sqrt(rhobar ** 2 + dampsq)
        cs1 = rhobar / rhobar1
        sn1 = damp / rhobar1
        psi = sn1 * phibar
        phibar *= cs1
        cs, sn, rho = sym_ortho(rhobar1, beta)
        theta = sn * alfa
        rhobar = -cs * alfa
        phi = cs * phibar
        phibar *= sn
        tau = sn * phi
        t1 = phi / rho
        t2 = -theta / rho
        dk = w * (1.0 / rho)
        x = x + t1 * w
        w = v + t2 * w
        ddnorm += dk @ dk
        if calc_var:
            var += dk * dk
        delta = sn2 * rho
        gambar = -cs2 * rho
        rhs = phi - delta * z
        zbar = rhs / gambar
        xnorm = np.sqrt(xxnorm + zbar ** 2)
        gamma = np.sqrt(gambar ** 2 + theta ** 2)
        cs2 = gambar / gamma
        sn2 = theta / gamma
        z = rhs / gamma
        xxnorm += z ** 2
        acond = anorm * np.sqrt(ddnorm)
        res1 = phibar ** 2
        res2 += psi ** 2
        rnorm = np.sqrt(res1 + res2)
        arnorm = alfa * abs(tau)
        r1sq = rnorm ** 2 - dampsq * xxnorm
        r1norm = np.sqrt(abs(r1sq))
        r2norm = rnorm
        test1 = rnorm / bnorm
        test2 = arnorm / (anorm * rnorm + eps)
        test3 = 1.0 / (acond + eps)
        t1 = test1 / (1.0 + anorm * xnorm / bnorm)
        rtol = atol + btol * (anorm * xnorm / bnorm)
        if itn >= iter_lim:
            istop = 7
        if 1.0 + test3 <= 1.0:
            istop = 6
        if 1.0 + test2 <= 1.0:
            istop = 5
        if 1.0 + t1 <= 1.0:
            istop = 4
        if test3 <= ctol:
            istop = 3
        if test2 <= atol:
            istop = 2
        if test1 <= rtol:
            istop = 1
        if istop != 0:
            break
    return x, istop, itn, r1norm, r2norm, anorm, acond, arnorms, xnorm, var


def lsqr_dense(a, b, **kw):
    a = np.asarray(a, dtype=np.float64)
    return lsqr(lambda v: a @ v, lambda u: a.T @ u, a.shape[0], a.shape[1], b, **kw)


def cgls(a_mul, at_mul, n, b, tol, iters, x0=None):
    """CGLS with the absolute stop ||A^T r|| < tol (src/cg.rs:18-61).
    Returns (x, iterations_run, converged)."""
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64).reshape(-1)
    r = b - a_mul(x)
    s = at_mul(r)
    p = s.copy()
    norm_s = s @ s
    for i in range(iters):
        ap = a_mul(p)
        alpha = norm_s / (ap @ ap)
        x = x + alpha * p
        r = r - alpha * ap
        s_new = at_mul(r)
        norm_s_new = s_new @ s_new
        if np.sqrt(norm_s_new) < tol:
            return x, i + 1, True
        beta = norm_s_new / norm_s
        norm_s = norm_s_new
        p = s_new + beta * p
    return x, iters, False


# ------------------------------------------------------- factorizations
def orth(x):
    """Q of a thin Householder QR, m x min(m, n) (src/lora_helpers.rs:131-133).
    Columns match nalgebra's up to sign."""
    q, _ = np.linalg.qr(np.asarray(x, dtype=np.float64), mode="reduced")
    return q


def full_piv_lu(x):
    """Full-pivot LU as nalgebra 0.33.0 FullPivLU::new (third-party, absent
    here; its published algorithm): at step i the pivot is the first maximum
    of |entry| over the trailing block scanned column by column
    (icamax_full); stop at an all-zero trailing block; the multipliers are
    the column times 1/pivot and the trailing block gets the rank-1 update
    (lu::gauss_step).  Returns (lu, p_rows, q_cols) with P x Q = L U."""
    a = np.array(x, dtype=np.float64)
    m, n = a.shape
    prow = np.arange(m)
    qcol = np.arange(n)
    for i in range(min(m, n)):
        blk = np.abs(a[i:, i:])
        flat = np.argmax(blk.T.reshape(-1))          # column-major first max
        cj, ri = divmod(int(flat), m - i)
        rp, cp = ri + i, cj + i
        diag = a[rp, cp]
        if diag == 0.0:
            break
        a[:, [i, cp]] = a[:, [cp, i]]
        qcol[[i, cp]] = qcol[[cp, i]]
        if rp != i:
            a[[i, rp], :] = a[[rp, i], :]
            prow[[i, rp]] = prow[[rp, i]]
        a[i + 1:, i] *= 1.0 / diag
        a[i + 1:, i + 1:] -= np.outer(a[i + 1:, i], a[i, i + 1:])
    return a, prow, qcol


def stabilizer(x):
    """L factor (m x min(m,n), unit lower) of full-pivot LU
    (src/lora_helpers.rs:144-146)."""
    lu, _, _ = full_piv_lu(x)
    m, n = lu.shape
    l = np.tril(lu[:, :min(m, n)], -1)
    l[np.arange(min(m, n)), np.arange(min(m, n))] = 1.0
    return l


def cholqr2(y):
    """CholeskyQR2 (no reference counterpart): Q, R with Y = Q R."""
    y = np.asarray(y, dtype=np.float64)
    r1 = np.linalg.cholesky(y.T @ y).T
    q1 = np.linalg.solve(r1.T, y.T).T
    r2 = np.linalg.cholesky(q1.T @ q1).T
    q = np.linalg.solve(r2.T, q1.T).T
    return q, r2 @ r1


# ------------------------------------------------------------- sketches
def reference_stream():
    """The stream every reference sketch uses: ThreeFry2x64Rng::seed_from_u64(0)
    (src/sketch.rs:112)."""
    return R.Stream.from_seed(R.THREEFRY2X64, 0)


def sketch(dist, rows, cols, stream=None, zeta=8):
    return R.sketching_operator(dist, rows, cols, stream or reference_stream(), zeta)


# ------------------------------------------------------------ low rank
def tsog1(a, k, num_passes, passes_per_stab, stream=None):
    """Power-iteration sketch, restated with the reference's even-branch
    behaviour: the Gaussian goes to S but the loop multiplies A by S1, which
    is still zero (src/lora_helpers.rs:58-105, esp. :70-72 and :89)."""
    a = np.asarray(a, dtype=np.float64)
    m, n = a.shape
    passes_done = 0
    s1 = np.zeros((n, k))
    s = np.zeros((n, k))
    if num_passes % 2 == 0:
        s = sketch(0, n, k, stream)                   # discarded below
    else:
        s1 = a.T @ sketch(0, m, k, stream)
        passes_done += 1
        if passes_done % passes_per_stab == 0:
            _ = stabilizer(s1)                        # result unused (:79-82)
    diff = num_passes - passes_done
    while diff >= 2:
        s = a @ s1
        passes_done += 1
        if passes_done % passes_per_stab == 0:
            s = stabilizer(s)
        s = a.T @ s
        passes_done += 1
        if passes_done % passes_per_stab == 0:
            s = stabilizer(s)
        diff -= 2
    return s


def rf1(a, k, stream=None):
    """Y = A tsog1(A, k, 2, 1); Q = Orth(Y) (src/lora_helpers.rs:37-44)."""
    s = tsog1(a, k, 2, 1, stream)
    return orth(np.asarray(a) @ s)


def qb1(a, k, stream=None):
    """Q = RF1(A, k); B = Q^T A (src/lora_helpers.rs:17-23)."""
    q = rf1(a, k, stream)
    return q, q.T @ np.asarray(a)


def _svd_truncate(q, b, k):
    """SVD of B, truncate to r = min(k, ncols Q) (src/lora_drivers.rs:51-68)."""
    ub, sb, vtb = np.linalg.svd(b, full_matrices=False)
    r = min(k, q.shape[1])
    return q @ ub[:, :r], np.diag(sb[:r]), vtb[:r, :]


def rand_svd(a, k, epsilon, s, stream=None):
    """Reference-faithful rand_svd (src/lora_drivers.rs:30-69)."""
    if k == 0 or epsilon <= 0.0 or s == 0:
        raise ValueError("InvalidParameters")
    q, b = qb1(a, k + s, stream)
    return _svd_truncate(q, b, k)


def rand_evd1(a, k, epsilon, s, q=None, stream=None):
    """rand_evd1 (src/lora_drivers.rs:87-151): Q from QB1 (q given, else the
    faithful qb1), C = B Q with B = Q^T A, eigenpairs of C (lower triangle
    read, :129), the k of largest |lambda| by a stable sort (:132-146),
    V = Q U (:148)."""
    if k == 0 or epsilon <= 0.0 or s == 0:
        raise ValueError("InvalidParameters")
    a = np.asarray(a, dtype=np.float64)
    if a.shape[0] != a.shape[1] or not np.array_equal(a, a.T):
        raise ValueError("NotHermitian")
    if q is None:
        q, b = qb1(a, k + s, stream)
    else:
        q = np.asarray(q, dtype=np.float64)
        b = q.T @ a
    c = b @ q
    w, u = np.linalg.eigh(c, UPLO="L")
    order = sorted(range(len(w)), key=lambda i: -abs(w[i]))
    idx = order[:min(k, len(w))]
    return q @ u[:, idx], w[idx]


def rand_evd2(a, k, s, s_mat=None, stream=None):
    """rand_evd2 (src/lora_drivers.rs:167-224): eigenvalue test of A's lower
    triangle (:178-183), S = tsog1(A, k + s, 3, 1) (s_mat given, else the
    faithful tsog1), nu = sqrt(n) eps ||A S||_F (:187-189), Cholesky of
    S^T (A S + nu S) reading its lower triangle (:190-199), B = Y R^{-1}, SVD,
    lambda = sigma^2 for sigma > 0, r = min(k, #{lambda > nu}), lambda - nu
    (:209-221)."""
    if k == 0:
        raise ValueError("InvalidParameters")
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    if np.any(np.linalg.eigvalsh(a, UPLO="L") < 0.0):
        raise ValueError("NotPositiveSemiDefinite")
    sm = tsog1(a, k + s, 3, 1, stream) if s_mat is None else np.asarray(s_mat, dtype=np.float64)
    y = a @ sm
    nu = np.sqrt(n) * np.finfo(np.float64).eps * np.linalg.norm(y)
    y_new = y + nu * sm
    sy = sm.T @ y_new
    sy = np.tril(sy) + np.tril(sy, -1).T
    try:
        lo = np.linalg.cholesky(sy)
    except np.linalg.LinAlgError:
        raise ValueError("MatrixDecompositionError") from None
    bm = np.linalg.solve(lo, y_new.T).T                # Y R^{-1} with R = L^T
    u, sig, _ = np.linalg.svd(bm, full_matrices=False)
    lam = [x * x for x in sig if x > 0.0]
    r = min(k, sum(1 for x in lam if x > nu))
    return u[:, :r], np.array(lam[:r]) - nu


def rand_svd_power(a, k, s, q_iters, omega):
    """The library's rand_svd method: Y = A Omega, q rounds of
    (Z = A^T Q, Y = A Q_z) with re-orthonormalisation after every product,
    Q = orth(Y), B = Q^T A, truncated SVD."""
    a = np.asarray(a, dtype=np.float64)
    y = a @ omega
    qm = orth(y)
    for _ in range(q_iters):
        qz = orth(a.T @ qm)
        qm = orth(a @ qz)
    return _svd_truncate(qm, qm.T @ a, k)


# ------------------------------------------------------- least squares
def _sketch_rows(m, n, sampling_factor):
    """d = min(m, floor(sf * n)) (src/sketch_and_precondition.rs:49)."""
    return m if sampling_factor * n > m else int(np.floor(sampling_factor * n))


def _check_ls(m, n, epsilon, l, sampling_factor):
    if m < n:
        raise ValueError("NotOverdetermined")
    if sampling_factor < 1.0 or epsilon <= 0.0 or l == 0:
        raise ValueError("InvalidParameters")


def blendenpik_overdetermined(a, b, epsilon, l, sampling_factor, s_op=None):
    """src/sketch_and_precondition.rs:26-59.  ``s_op`` is the d x m sketch
    (defaults to the reference's dense Gaussian).  Returns (x, cgls_iters)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    m, n = a.shape
    _check_ls(m, n, epsilon, l, sampling_factor)
    d = _sketch_rows(m, n, sampling_factor)
    s = sketch(0, d, m) if s_op is None else s_op
    a_sk = s @ a
    b_sk = s @ b
    q, r = np.linalg.qr(a_sk, mode="reduced")
    z0 = q.T @ b_sk
    rinv = np.linalg.solve(r, np.eye(n))
    z, it, _ = cgls(lambda v: a @ (rinv @ v), lambda u: rinv.T @ (a.T @ u), n, b,
                    epsilon, l, z0)
    return rinv @ z, it


def blendenpik_lsqr(a, b, epsilon, l, sampling_factor, s_op):
    """blendenpik_overdetermined (src/sketch_and_precondition.rs:26-59) with
    LSQR (src/solvers.rs:115-278; atol = btol = epsilon, conlim 1e8,
    iter_lim = l) in place of CGLS, from the same z0 = Q^T S b: the per-problem
    algorithm of the batched path.  Returns (x, istop, itn)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    m, n = a.shape
    _check_ls(m, n, epsilon, l, sampling_factor)
    q, r = np.linalg.qr(s_op @ a, mode="reduced")
    z0 = q.T @ (s_op @ b)
    rinv = np.linalg.solve(r, np.eye(n))
    out = lsqr(lambda v: a @ (rinv @ v), lambda u: rinv.T @ (a.T @ u), m, n, b,
               atol=epsilon, btol=epsilon, conlim=1e8, iter_lim=l, x0=z0)
    return rinv @ out[0], out[1], out[2]


def lsrn_overdetermined(a, b, epsilon, l, sampling_factor, s_op=None):
    """src/sketch_and_precondition.rs:82-119: N = V Sigma^+ from SVD(SA),
    CGLS on A N from 0."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    m, n = a.shape
    _check_ls(m, n, epsilon, l, sampling_factor)
    d = _sketch_rows(m, n, sampling_factor)
    s = sketch(0, d, m) if s_op is None else s_op
    _, sig, vt = np.linalg.svd(s @ a, full_matrices=False)
    sinv = np.where(sig != 0.0, 1.0 / np.where(sig != 0.0, sig, 1.0), 0.0)
    nmat = vt.T * sinv
    y, it, _ = cgls(lambda v: a @ (nmat @ v), lambda u: nmat.T @ (a.T @ u),
                    nmat.shape[1], b, epsilon, l, np.zeros(nmat.shape[1]))
    return nmat @ y, it


def sketched_least_squares_qr(a, b, s_op=None):
    """src/sketch_and_solve.rs:24-33: S with m // 4 rows (the reference's
    Gaussian unless ``s_op`` is given), QR of S A, x = R^{-1} Q^T S b by
    solve_upper_triangular_system (zero diagonal -> 0).  For m // 4 < n the
    reference's x has m // 4 entries; they are returned zero-padded to n."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    m, n = a.shape
    d = m // 4
    if d <= 0:
        raise ValueError("the sketch needs m // 4 >= 1 rows")
    s = sketch(0, d, m) if s_op is None else s_op
    q, r = np.linalg.qr(s @ a, mode="reduced")
    z = q.T @ (s @ b)
    k = min(d, n)
    x = np.zeros(n)
    x[:k] = np.asarray(solve_upper_triangular_system(r[:k, :k], z[:k])).reshape(-1)
    return x


def sketched_least_squares_svd(a, b, s_op=None):
    """src/sketch_and_solve.rs:53-65: thin SVD S A = U diag(sigma) V^T,
    x = V solve_diagonal_system(sigma, U^T S b)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    m, n = a.shape
    d = m // 4
    if d <= 0:
        raise ValueError("the sketch needs m // 4 >= 1 rows")
    s = sketch(0, d, m) if s_op is None else s_op
    u, sig, vt = np.linalg.svd(s @ a, full_matrices=False)
    return vt.T @ solve_diagonal_system(sig, u.T @ (s @ b))


def sketch_saddle_point_precondition(a, b, c, mu, epsilon, l, sampling_factor, s_op=None):
    """src/sketch_and_precondition.rs:150-216: min ||Ax-b||^2 + mu||x||^2 +
    2<c,x>; returns (x, y = b - A x)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    m, n = a.shape
    _check_ls(m, n, epsilon, l, sampling_factor)
    d = min(max(int(np.floor(sampling_factor * n)), 1), m)
    s = sketch(0, d, m) if s_op is None else s_op
    u, sig, vt = np.linalg.svd(s @ a, full_matrices=False)
    if mu > 0.0:
        mmat = vt.T * (1.0 / np.sqrt(sig ** 2 + mu))
    else:
        kk = 0
        while kk < len(sig) and sig[kk] > 1e-10:
            kk += 1
        mmat = vt.T[:, :kk] * (1.0 / sig[:kk])
    b_mod = b.copy()
    c = np.asarray(c, dtype=np.float64).reshape(-1)
    if c.size:
        b_shift = vt @ c
        if mu > 0.0:
            temp = s.T @ (u @ (b_shift / np.sqrt(sig ** 2 + mu)))
        else:
            temp = s.T @ (u @ (b_shift / sig))
        b_mod = b_mod - temp
    z0 = u.T @ (s @ b_mod)
    z0 = z0[:mmat.shape[1]]
    z, _, _ = cgls(lambda v: a @ (mmat @ v), lambda uu: mmat.T @ (a.T @ uu),
                   mmat.shape[1], b_mod, epsilon, l, z0)
    x = mmat @ z
    return x, b - a @ x


# ------------------------------------------------------- pivoted QR
def qrcp(a):
    """Householder QRCP with recomputed trailing column norms
    (src/pivot_decompositions.rs:106-175).  Q accumulated from the
    reflectors (same matrix as the reference's dense products)."""
    return _qrcp_steps(a, None)


def economic_qrcp(a, k):
    """The first k steps of qrcp, Q m x k, R k x n, full permutation
    (src/pivot_decompositions.rs:194-269; asserts :198-199)."""
    m, n = np.shape(a)
    if k > min(m, n):
        raise AssertionError("k must be <= min(m,n)")
    if k <= 0:
        raise AssertionError("k must be positive")
    q, r, p = _qrcp_steps(a, k)
    return q[:, :k], r[:k, :], p


def _qrcp_steps(a, steps):
    r = np.array(a, dtype=np.float64)
    m, n = r.shape
    q = np.eye(m)
    p = np.arange(n)
    norms = np.linalg.norm(r, axis=0)
    for k in range(min(m, n) if steps is None else steps):
        j = k + int(np.argmax(norms[k:]))            # first maximum, strict >
        if norms[j] <= norms[k]:
            j = k
        if j != k:
            r[:, [k, j]] = r[:, [j, k]]
            p[[k, j]] = p[[j, k]]
            norms[[k, j]] = norms[[j, k]]
        x = r[k:, k].copy()
        nx = np.linalg.norm(x)
        if nx != 0.0:
            v = x
            v[0] += nx if v[0] >= 0.0 else -nx
            v /= np.linalg.norm(v)
            r[k:, k:] -= 2.0 * np.outer(v, v @ r[k:, k:])
            q[:, k:] -= 2.0 * np.outer(q[:, k:] @ v, v)
            norms[k + 1:] = np.linalg.norm(r[k + 1:, k + 1:], axis=0)
    return q, r, p


# ------------------------------------------ interpolative decompositions
ROW, COLUMN = 0, 1


def _check_id_k(k, l, w):
    if k <= 0:
        raise AssertionError("k must be positive)")
    if k > min(l, w):
        raise AssertionError("k must be <= min(l,w)")


def osid_qrcp(y, k, attr):
    """One-sided ID by k QRCP steps (src/id.rs:248-305).  Column: X (k x w)
    with X[:, J[:k]] = I and X[:, J[k:]] = R1^{-1} R2, returns (X, J[:k]).
    Row: the column ID of y^T, transposed."""
    y = np.asarray(y, dtype=np.float64)
    l, w = y.shape
    _check_id_k(k, l, w)
    if attr == ROW:
        x, i = osid_qrcp(y.T, k, COLUMN)
        return x.T, i
    _, r, j = economic_qrcp(y, k)
    r1 = r[:k, :k]
    if np.any(np.diag(r1) == 0.0):
        raise AssertionError("called `Option::unwrap()` on a `None` value")
    t = np.linalg.solve(np.triu(r1), r[:k, k:])
    x = np.zeros((k, w))
    x[:, j[:k]] = np.eye(k)
    x[:, j[k:]] = t
    return x, j[:k]


def osid_randomised(a, k, attr, s_mat=None, stream=None):
    """src/id.rs:205-234.  Column: Y = S A with S the k x m Gaussian sketch
    (s_mat, else the reference stream); Row: Y = A S^T with S = tsog1(A, k, 2, 1)
    (s_mat, else the faithful tsog1), defined for n == k."""
    a = np.asarray(a, dtype=np.float64)
    m, n = a.shape
    _check_id_k(k, m, n)
    if attr == ROW:
        s = tsog1(a, k, 2, 1, stream) if s_mat is None else np.asarray(s_mat, float)
        if n != k:
            raise AssertionError("Matrix multiplication dimensions mismatch")
        return osid_qrcp(a @ s.T, k, ROW)
    s = sketch(0, k, m, stream) if s_mat is None else np.asarray(s_mat, float)
    return osid_qrcp(s @ a, k, COLUMN)


def two_sided_id(a, k, osid=None):
    """src/id.rs:125-132 (osid: the one-sided ID used, osid_qrcp by default;
    src/id.rs:94-101 passes osid_randomised).  Returns (Z, I, J, X)."""
    osid = osid or osid_qrcp
    a = np.asarray(a, dtype=np.float64)
    x, j = osid(a, k, COLUMN)
    z, i = osid(a[:, j], k, ROW)
    return z, i, j, x


def cur(a, k, osid=None):
    """src/id.rs:33-76 (osid_qrcp) and :154-197 (osid_randomised): returns
    (J, U, I) with A ~ A[:, J] U A[I, :]; pinv is nalgebra's
    pseudo_inverse(0.0) (every sigma > 0 inverted)."""
    osid = osid or osid_qrcp
    a = np.asarray(a, dtype=np.float64)
    m, n = a.shape
    if m >= n:
        x, j = osid(a, k, COLUMN)
        _, _, i = economic_qrcp(a[:, j].T, k)
        i = i[:k]
        u = x @ np.linalg.pinv(a[i, :], rcond=0.0)
        return j[:k], u, i
    z, i = osid(a.T, k, COLUMN)
    _, _, j = economic_qrcp(a[i, :], k)
    j = j[:k]
    u = np.linalg.pinv(a[:, j], rcond=0.0) @ z.T
    return j, u, i[:k]


def sap_chol_qrcp(a, d, s_op=None):
    """CQRRPT (src/cqrrpt.rs:27-58).  Returns (Q m x k, R k x n, J, k)."""
    a = np.asarray(a, dtype=np.float64)
    m, n = a.shape
    if not (n <= d <= m):
        raise AssertionError("d must satisfy n ≤ d ≪ m")
    s = sketch(0, d, m) if s_op is None else s_op
    _, r_sk, j = qrcp(s @ a)
    kk = int(np.sum(np.abs(np.diag(r_sk[:min(r_sk.shape), :min(r_sk.shape)])) > 1e-10))
    a_pre = a[:, j[:kk]] @ np.linalg.inv(np.triu(r_sk[:kk, :kk]))
    r_pre = np.linalg.cholesky(a_pre.T @ a_pre).T
    q = a_pre @ np.linalg.inv(r_pre)
    return q, r_pre @ r_sk[:kk, :n], j, kk


# ------------------------------------------------------- test matrices
def lowrank_matrix(m, n, sigma, stream_u, stream_v):
    """U diag(sigma) V^T with U, V the Q factors of seeded Gaussian m x r and
    n x r matrices (r = len(sigma))."""
    r = len(sigma)
    gu = R.gaussian(stream_u, m * r).reshape(r, m).T
    gv = R.gaussian(stream_v, n * r).reshape(r, n).T
    u, _ = np.linalg.qr(gu)
    v, _ = np.linalg.qr(gv)
    return (u * np.asarray(sigma)) @ v.T


def hadamard_index_sign(r, c):
    """Entry (r, c) of the unnormalised Sylvester Hadamard matrix."""
    x = np.bitwise_and(r, c)
    pc = np.zeros_like(x)
    while np.any(x):
        pc += x & 1
        x >>= 1
    return np.where(pc & 1, -1.0, 1.0)


def _substream(s, tag):
    return R.Stream(s.gen, s.key, (s.ctr[0], (s.ctr[1] + tag) & R.M64))


def hadamard_test_matrix(m, n, sigma, noise, stream):
    """The library's test-matrix definition (no reference counterpart; the
    generators of src/test_assist.rs build their matrices from thread_rng):
    A = D_r H_m[:, J] diag(sigma) (H_n D_c)^T / sqrt(m n) + noise * G.
    Sub-streams (high counter word + tag): 1 row signs, 2 column signs,
    3 the map J(j) = (a j + b) mod m (a odd), 4 the noise.  Small m, n only
    (explicit Hadamard matrices)."""
    blk = _substream(stream, 3).blocks([0])[0]
    a = (int(blk[0]) | 1) & (m - 1) or 1
    b = int(blk[1]) & (m - 1)
    J = (a * np.arange(n, dtype=np.uint64) + np.uint64(b)) & np.uint64(m - 1)
    dr = R.rademacher(_substream(stream, 1), m)
    dc = R.rademacher(_substream(stream, 2), n)
    rr = np.arange(m, dtype=np.int64)[:, None]
    cc = np.arange(n, dtype=np.int64)[None, :]
    U = hadamard_index_sign(rr, J.astype(np.int64)[None, :]) * dr[:, None]
    V = hadamard_index_sign(np.arange(n)[:, None], cc) * dc[None, :]
    A = (U * np.asarray(sigma, dtype=np.float64)) @ V.T / np.sqrt(float(m) * n)
    if noise:
        A = A + noise * R.gaussian(_substream(stream, 4), m * n).reshape(n, m).T
    return A
