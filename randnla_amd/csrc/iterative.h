// Device state and launchers of the iterative solvers and small helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rna {

// Scalar state of LSQR (names as in src/solvers.rs:137-187).
struct LsqrState {
  double alfa, beta, rhobar, phibar, rnorm, r1norm, r2norm, anorm, acond;
  double damp, dampsq, ddnorm, res2, xnorm, xxnorm, z, cs2, sn2, arnorm, bnorm, ctol;
  double atol, btol;
  double t1, t2, inv_rho, uscale, vscale;
  double umul;   // fused path: the stored u times umul is LSQR's u (1 at the start)
  int istop, pending_stop, itn, iter_lim;
};

// Scalar state of CGLS (src/cg.rs:29-55).
struct CglsState {
  double norm_s, alpha, beta, tol, last_norm;
  int iters, max_iters, done, converged;
};

hipError_t lsqr_u_update(LsqrState* S, int64_t m, const double* t, double* u, double* part,
                         hipStream_t st);
hipError_t lsqr_v_update(LsqrState* S, int64_t n, const double* t, double* v, const double* w,
                         double* part, double* arnorms, hipStream_t st);
// One pass over A (m x n, n <= 512): p = A w - alfa u (u = umul * stored u),
// beta = ||p|| and anorm advanced as in lsqr_u_update, p stored unscaled in u
// (umul <- 1 / beta), and atu = A^T (p / beta).  part needs
// lsqr_fused_groups(m) doubles, spart lsqr_fused_groups(m) * n.
int lsqr_fused_groups(int64_t m);
hipError_t lsqr_fused_update(LsqrState* S, const double* A, int64_t m, int64_t n, int64_t lda,
                             const double* w, double* u, double* part, double* spart,
                             double* atu, hipStream_t st);
// The rest of one LSQR iteration in one work-group (n <= 1024): v = t - beta v,
// alfa = ||v||, the scalar recurrence, v /= alfa, x += t1 w, var, w = v + t2 w
// and the stop-code commit (lsqr_v_update followed by lsqr_xw_update).
hipError_t lsqr_vphase(LsqrState* S, int64_t n, const double* t, double* v, double* w, double* x,
                       double* var, bool calc_var, double* arnorms, int cap, hipStream_t st);
hipError_t lsqr_xw_update(LsqrState* S, int64_t n, double* x, double* w, const double* v,
                          double* var, bool calc_var, double* arnorms, int cap, hipStream_t st);
hipError_t cgls_alpha(CglsState* S, const double* ap, int64_t m, double* part, hipStream_t st);
hipError_t cgls_xr(CglsState* S, int64_t n, int64_t m, double* x, const double* p, double* r,
                   const double* ap, hipStream_t st);
hipError_t cgls_beta_p(CglsState* S, int64_t n, double* p, const double* s, double* part,
                       hipStream_t st);
hipError_t sqnorm(const double* x, int64_t n, double* out, double* part, hipStream_t st);
hipError_t gather_cols(const double* A, int64_t m, int64_t lda, const int64_t* idx, int64_t k,
                       double* out, int64_t ldo, hipStream_t st);
hipError_t backsub(const double* U, int64_t n, int64_t ldu, const double* y, double* x,
                   double* work, hipStream_t st);
hipError_t backsub_f32(const float* U, int64_t n, int64_t ldu, const float* y, float* x,
                       float* work, hipStream_t st);
hipError_t count_diag_above(const double* R, int64_t k, int64_t ld, double tol, int* out,
                            hipStream_t st);
hipError_t upper_copy(const double* S, int64_t rows, int64_t cols, int64_t lds, double* D,
                      int64_t ldd, hipStream_t st);
// G (n x n, ld ldg) = the symmetric matrix whose lower triangle (diagonal
// included) is A's, plus `shift` on the diagonal.
hipError_t sym_from_lower(const double* A, int64_t n, int64_t lda, double* G, int64_t ldg,
                          double shift, hipStream_t st);
// *flag = 1 if some A(i, j) != A(j, i) (NaN entries compare unequal, the
// diagonal included), else 0.
hipError_t sym_mismatch(const double* A, int64_t n, int64_t lda, int* flag, hipStream_t st);
// Interpolation matrix of a column ID: X (k x w, ld ldx) with column perm[c]
// = e_c for c < k and T[:, c - k] (T k x (w - k), ld k) after.
hipError_t id_scatter(const int64_t* perm, const double* T, int64_t k, int64_t w, double* X,
                      int64_t ldx, hipStream_t st);
// out (k x n, ld ldo) = rows idx[0..k) of A.
hipError_t gather_rows(const double* A, int64_t n, int64_t lda, const int64_t* idx, int64_t k,
                       double* out, int64_t ldo, hipStream_t st);
hipError_t column_norms(const double* Y, int64_t m, int64_t b, int64_t ld, double* out,
                        hipStream_t st);
hipError_t scale_columns(double* X, int64_t rows, int64_t cols, int64_t ld, const double* d,
                         int64_t stride, hipStream_t st);

hipError_t scale_rows(double* X, int64_t rows, int64_t cols, int64_t ld, const double* d,
                      hipStream_t st);
// sg[i] = sign of W[i][i] (-1 if negative, else +1), i < k.
hipError_t diag_signs(const double* W, int64_t k, int64_t ld, double* sg, hipStream_t st);
// R (k x n) = diag(sg) * (upper triangle of W's first k rows).
hipError_t upper_scaled(const double* W, int64_t k, int64_t n, int64_t ldw, const double* sg,
                        double* R, int64_t ldr, hipStream_t st);
hipError_t inv_sigma(const double* s, int64_t n, double mu, double tol, int mode, double* out,
                     int* count, hipStream_t st);

}  // namespace rna
