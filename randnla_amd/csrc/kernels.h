// Internal launch interface of the gfx950 kernels (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rng.h"

namespace rna {

enum class T { F64, F32, BF16 };

// ---- FP64 MFMA GEMM ------------------------------------------------------
// C (M x N, column-major, ldc) = alpha * op(A) * B + beta * C.
// op(A) = A (M x K, lda) when !a_trans, A^T (A is K x M, lda) when a_trans.
// B is K x N (ldb).  Inputs may be FP64 or FP32 (promoted exactly); C may be
// FP64, FP32 or BF16 (rounded on store).  splits > 1 splits K; the partial
// products go to C + z * cstride (beta must be 0) and must be reduced.
struct GemmArgs {
  int64_t M, N, K;
  const void* A; int64_t lda; T ta; bool a_trans;
  const void* B; int64_t ldb; T tb;
  void* C; int64_t ldc; T tc;
  double alpha = 1.0, beta = 0.0;
  int splits = 1;
  int64_t cstride = 0;
  int tri = 0;   // 1: compute only the upper triangle of C; 2: B upper triangular
};
hipError_t gemm_f64(const GemmArgs& g, hipStream_t st);
// G = Y^T Y (l x l, upper triangle incl. the diagonal; zeros below) for a
// column-major FP32 or FP64 panel Y (m x l, l <= 320, 16-byte aligned
// columns); part: syrk_workspace(m, l, ty) bytes of scratch.
bool syrk_supported(T ty, int64_t l, int64_t ldy, const void* Y);
size_t syrk_workspace(int64_t m, int64_t l, T ty);
hipError_t syrk_f64(const void* Y, T ty, int64_t m, int64_t l, int64_t ldy, double* G,
                    int64_t ldg, double* part, hipStream_t st);

// ---- low-precision MFMA GEMM (gfx950) -----------------------------------
// C (fp32, M x N) = alpha * op(A) * B.  mode 0: BF16 operands (A, B stored as
// BF16).  mode 1: FP32 operands split on the fly into FP16 hi + lo pairs
// scaled by sa / sb (powers of two) and multiplied as hi*hi + hi*lo + lo*hi.
// mode 2: op(A) split as in mode 1, B rounded to FP16 once: hi*b + lo*b.
// mode 3: op(A) and B (FP32) each rounded to FP16 once (scaled as in mode 2):
// one product hi*b.  mode 4: op(A) rounded once, B split into hi + lo:
// hi*b_hi + hi*b_lo.
struct LpGemmArgs {
  int mode;
  int64_t M, N, K;
  const void* A; int64_t lda; bool a_trans;
  const void* B; int64_t ldb;
  float* C; int64_t ldc; int64_t cstride;
  float alpha = 1.f, sa = 1.f, sb = 1.f;
  int splits = 1;
  // modes 1-3: when set, sa / sb are derived on the device from {max|op(A)|,
  // max|B|} (doubles) and alpha is divided by sa * sb.
  const double* amax = nullptr;
  // modes 1-3: a caller-provided buffer (leading dimension ldbh >= K, a
  // multiple of 8) of N rows for B rounded to FP16 once (modes 2, 3), or 2 N rows
  // for its FP16 hi and lo planes (mode 1); with it, full-tile products with
  // K-contiguous op(A) load B's tiles straight into LDS.
  uint16_t* bh = nullptr;
  int64_t ldbh = 0;
};
hipError_t gemm_lp(const LpGemmArgs& g, hipStream_t st);
// Sum `splits` fp32 partial matrices (stride cstride) into dst (FP64 or FP32).
hipError_t reduce_partials(const float* src, int64_t M, int64_t N, int64_t lds,
                           int64_t cstride, int splits, void* dst, T tdst, int64_t ldd,
                           double alpha, hipStream_t st);
hipError_t reduce_partials_f64(const double* src, int64_t M, int64_t N, int64_t lds,
                               int64_t cstride, int splits, double* dst, int64_t ldd,
                               double alpha, hipStream_t st);

// ---- conversions / elementwise -------------------------------------------
hipError_t convert(const void* src, T ts, int64_t M, int64_t N, int64_t lds, void* dst,
                   T td, int64_t ldd, double scale, hipStream_t st);
hipError_t absmax(const void* src, T ts, int64_t M, int64_t N, int64_t lds, double* out,
                  hipStream_t st);
// BF16 copy (ldd), FP32 transpose (N x M, ldt) and max |src| of an FP32
// matrix in one pass; M, N and the leading dimensions multiples of 4.
hipError_t convert_bf16_transpose_absmax(const float* src, int64_t M, int64_t N, int64_t lds,
                                         uint16_t* dst, int64_t ldd, float* tr, int64_t ldt,
                                         double* amax, hipStream_t st);
// BF16 copy of an FP32 matrix and max |src| into *amax (one pass over src).
hipError_t convert_bf16_absmax(const float* src, int64_t M, int64_t N, int64_t lds, uint16_t* dst,
                              int64_t ldd, double* amax, hipStream_t st);
hipError_t transpose_f64(const double* src, int64_t M, int64_t N, int64_t lds, double* dst,
                         int64_t ldd, hipStream_t st);
hipError_t fill_f64(double* p, int64_t M, int64_t N, int64_t ld, double v, hipStream_t st);
hipError_t set_identity(double* p, int64_t M, int64_t N, int64_t ld, hipStream_t st);

// Stream-ordered device workspace: a block released with ws_free on stream
// st is handed out again by a later ws_alloc of the same size on st (the
// stream orders the reuse after every earlier use); ws_flush(st) returns the
// idle blocks kept for st (st may be the null stream) to the runtime,
// ws_flush_all those of every stream.
hipError_t ws_alloc(void** p, size_t bytes, hipStream_t st);
void ws_free(void* p, hipStream_t st);
void ws_flush(hipStream_t st);
void ws_flush_all();
// Device -> host copy of at most D2H_STAGE_BYTES through pinned staging: the
// data are in dst when the call returns (the stream is drained).
constexpr size_t D2H_STAGE_BYTES = size_t(1) << 20;
hipError_t d2h_small(hipStream_t st, void* dst, const void* src, size_t bytes);
hipError_t d2h_small_2d(hipStream_t st, void* dst, size_t dpitch, const void* src, size_t spitch,
                        size_t width, size_t height);

// ---- small dense factorizations (single work-group kernels) -------------
// Upper Cholesky of an SPD n x n G (in place into R, upper part).  A column
// whose pivot falls to <= tau * G_jj (dependent column) gets a zero row in R
// and flag[j] = 1.  ndef (device int) counts such columns.
hipError_t cholesky_upper(double* G, int64_t n, int64_t ld, double tau, int* flags,
                          int* ndef, hipStream_t st);
// X = R^{-1} for upper-triangular R; zero diagonal -> zero row/col in X
// (pseudo-inverse on the kept columns).
hipError_t trinv_upper(const double* R, int64_t n, int64_t ldr, double* X, int64_t ldx,
                       hipStream_t st);
// Both of the above in one launch (n <= 512; larger n: the two calls):
// R = chol(G) in G's upper triangle (strictly lower part zeroed), X = R^{-1}.
hipError_t cholesky_inverse_upper(double* G, int64_t n, int64_t ld, double tau, int* flags,
                                  int* ndef, double* X, int64_t ldx, hipStream_t st);
// Replace flagged columns of Q (m x n) with unit vectors e_j.
hipError_t complete_columns(void* Q, T tq, int64_t m, int64_t n, int64_t ldq,
                            const int* flags, hipStream_t st);
// One-sided Jacobi SVD of W (n x n, in place): on return W = U * diag(s)
// (columns scaled), V accumulates the rotations, s = column norms.
// hist (optional, max_sweeps entries): rotations applied in each sweep.
hipError_t jacobi_svd(double* W, int64_t n, int64_t ldw, double* V, int64_t ldv,
                      double* s, int* rot_count, int max_sweeps, double tol,
                      hipStream_t st, int* sweeps_done, int* hist = nullptr);
// Sort columns by s descending (stable), normalise U = W / s (zero s -> e_j).
hipError_t svd_finalize(double* W, int64_t n, int64_t ldw, double* V, int64_t ldv,
                        double* s, double* Uout, int64_t ldu, double* Vout, int64_t ldvo,
                        double* sout, hipStream_t st);
// Householder QR, with column pivoting when `pivot` (pivot = max trailing
// column norm, first maximum), of A (m x n, in place; R in the upper
// triangle, zeros or rounding residue below).  tau (n + m * min(m, n)
// doubles): tau[k] = 1 if reflector k is active; the unit reflectors follow
// as an m x min(m, n) panel.  perm (int64, n) on device.  Without pivoting,
// `extra` columns stored after the n factored ones are multiplied by Q^T
// too (right-hand sides).
// steps >= 0 (pivoted only): stop after that many reflectors.
hipError_t qrcp_householder(double* A, int64_t m, int64_t n, int64_t lda, double* tau,
                            int64_t* perm, bool pivot, hipStream_t st, int64_t extra = 0,
                            int64_t steps = -1);
// Form Q (m x q) from the Householder vectors stored below the diagonal.
hipError_t householder_form_q(const double* A, int64_t m, int64_t n, int64_t lda,
                              const double* tau, double* Q, int64_t ldq, int64_t qcols,
                              hipStream_t st);

// Full-pivot LU of X (m x n, in place; nalgebra 0.33 FullPivLU: first
// maximum of |x| over the trailing block in column-major order, stop at a zero
// maximum) and its unit lower factor L (m x min(m, n)).  work: device doubles,
// full_piv_lu_work(m, n) of them.
int64_t full_piv_lu_work(int64_t m, int64_t n);
hipError_t full_piv_lu_l(double* X, int64_t m, int64_t n, int64_t ld, double* L, int64_t ldl,
                         double* work, hipStream_t st);

// ---- sketches ------------------------------------------------------------
// Dense panel: out (rows x cols, ld) = entries of `dist` from stream s.
hipError_t sketch_panel(int dist, const Stream& s, int64_t rows, int64_t cols, void* out,
                        T tout, int64_t ld, double scale, hipStream_t st);
hipError_t stream_words(const Stream& s, uint64_t first, int64_t n, uint64_t* out,
                        hipStream_t st);
// out (d x n, FP64) = S * A with S (d x m) a sparse-sign operator generated in
// registers (zeta nonzeros +-1/sqrt(zeta) per column).  A FP64 or FP32.
hipError_t sparse_sign_apply(const Stream& s, int64_t d, int zeta, const void* A, T ta,
                             int64_t m, int64_t n, int64_t lda, double* out, int64_t ldo,
                             double* work, int64_t work_elems, hipStream_t st);
// Workspace (doubles) needed by sparse_sign_apply.
int64_t sparse_sign_work(int64_t d, int64_t m, int64_t n);
// Dense sketch fused into the product (operator entries generated in
// registers, never stored).  LEFT: out (d x n) = S (d x m) * A.
// RIGHT: out (m x d) = A * Omega (n x d).
hipError_t dense_sketch_apply(int side, int dist, const Stream& s, const void* A, T ta,
                              int64_t m, int64_t n, int64_t lda, int64_t d, double* out,
                              int64_t ldo, double* work, int64_t work_elems, hipStream_t st);
int64_t dense_sketch_work(int side, int64_t m, int64_t n, int64_t d);
// out (m) = S^T y for a dense generated S (d x m) / a sparse-sign S.
hipError_t dense_sketch_transpose(int dist, const Stream& s, int64_t d, int64_t m,
                                  const double* y, double* out, hipStream_t st);
hipError_t sparse_sign_transpose(const Stream& s, int64_t d, int zeta, int64_t m,
                                 const double* y, double* out, hipStream_t st);

// ---- test matrices -------------------------------------------------------
hipError_t hadamard_test_matrix(int64_t m, int64_t n, const double* sigma_dev,
                                double noise, const Stream& s, double* work, void* out,
                                T tout, int64_t ld, hipStream_t st);

// ---- vector kernels for iterative solvers -------------------------------
// y (m) = alpha * A x + beta * y ; A m x n FP64 column-major.
hipError_t gemv_n(const double* A, int64_t m, int64_t n, int64_t lda, const double* x,
                  double* y, double alpha, double beta, hipStream_t st);
// y (n) = alpha * A^T x + beta * y.
hipError_t gemv_t(const double* A, int64_t m, int64_t n, int64_t lda, const double* x,
                  double* y, double alpha, double beta, double* work, hipStream_t st);
// y = R x (upper triangular, n x n) or y = R^T x.
hipError_t trmv_upper(const double* R, int64_t n, int64_t ldr, bool trans, const double* x,
                      double* y, hipStream_t st);
hipError_t dot(const double* x, const double* y, int64_t n, double* out, double* work,
               hipStream_t st);
hipError_t axpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t st);

// ---- batched uneven least squares (one persistent launch) -----------------
struct BatchProbDesc {
  int64_t m, n, lda;          // A_i: m x n FP32, column-major
  int64_t a_off, b_off, x_off;
  int64_t stream_id;
};
struct BatchPlan;
hipError_t batch_plan_create(const BatchProbDesc* desc, int64_t count, double sf, double tol,
                             int64_t iter_lim, const Stream& s, hipStream_t st, BatchPlan** out);
hipError_t batch_plan_execute(BatchPlan* pl, const float* A, const double* b, double* x,
                              hipStream_t st);
hipError_t batch_plan_autotune(BatchPlan* pl, const float* A, const double* b, double* x,
                               hipStream_t st);
hipError_t batch_plan_info(BatchPlan* pl, void* info, bool info_on_device, int* err_flag,
                           hipStream_t st);
hipError_t batch_plan_sketch(BatchPlan* pl, float* out, hipStream_t st);
int64_t batch_plan_count(const BatchPlan* pl);
int64_t batch_plan_sketch_size(const BatchPlan* pl);
void batch_plan_config(const BatchPlan* pl, int* rshift, int* wgs_per_cu, double* best_ms);
void batch_plan_destroy(BatchPlan* pl);

// ---- learned (k, p, q) proposer: 34 -> 128 -> 128 -> 3 GELU MLP ----------
int mlp_param_count();
int mlp_group_samples();
hipError_t mlp_forward(const float* P, const float* X, int B, float* Y, hipStream_t st);
hipError_t mlp_grad(const float* P, const float* X, const float* T, const float* wo, int B,
                    float* grad, float* loss, float* work, hipStream_t st);
hipError_t mlp_adam(float* P, float* m, float* v, const float* g, int step, float lr, float b1,
                    float b2, float eps, hipStream_t st);

// ---- FP8-e4m3 GEMM with per-128-block power-of-two scales ----------------
// C (FP32) = alpha * sum_k sA(i,k) A8[i,k] sB(k,j) B8[k,j]; A8 row i and B8
// column j are K-contiguous (ld multiples of 16, K padded with zeros to a
// multiple of 128); scales are E8M0 bytes (2^(e-127)):
// sA(i,k) = SA[(i/128) sa_r + (k/128) sa_k], sB(k,j) = SB[j sb_j + k/128].
struct Fp8GemmArgs {
  int64_t M, N, K;
  const uint8_t* A; int64_t lda; const uint8_t* SA; int64_t sa_r, sa_k;
  const uint8_t* B; int64_t ldb; const uint8_t* SB; int64_t sb_j;
  float* C; int64_t ldc; int64_t cstride = 0;
  float alpha = 1.f;
  int splits = 1;
};
hipError_t gemm_fp8(const Fp8GemmArgs& g, hipStream_t st);
// A (FP32, m x n, column-major) -> Acm (column-major FP8, ld ldcm >= m
// rounded up to 128, n rounded up to 128 columns), Arm (row-major FP8, ld
// ldr >= n rounded up to 128, m rounded up to 128 rows), S (one E8M0 byte per
// 128 x 128 block, S[(i/128) * ceil(n/128) + j/128]).
hipError_t quantize_fp8_a(const float* A, int64_t m, int64_t n, int64_t lda, uint8_t* Arm,
                          int64_t ldr, uint8_t* Acm, int64_t ldcm, uint8_t* S, hipStream_t st);
// X (FP32, K x N) -> X8 (column-major, ld ldq >= K rounded up to 128) and
// S[j * ceil(K/128) + k/128].
hipError_t quantize_fp8_panel(const float* X, int64_t K, int64_t N, int64_t ldx, uint8_t* X8,
                              int64_t ldq, uint8_t* S, hipStream_t st);

}  // namespace rna
