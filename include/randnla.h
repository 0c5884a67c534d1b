/*
 * randnla.h -- C ABI of the MI355X (gfx950) sketch-and-factor library.
 *
 * Drop-in boundary for the public API of the reference crate `randblas`
 * 0.1.0 (divijkhaitan/randnla).  Each entry point names the reference
 * function it replaces (file:line, relative to the reference tree).  A Rust
 * wrapper binds these symbols with `extern "C"` and keeps the reference's
 * signatures; see the binding stub in the integration notes.
 *
 * Conventions
 *  - Matrices are column-major with an explicit leading dimension, as
 *    nalgebra's DMatrix (ld = nrows for a contiguous matrix).
 *  - Every matrix/vector pointer may be host memory or device memory of the
 *    context's GPU; the library detects which and stages host buffers.
 *  - Calls are synchronous (the result is complete when the call returns),
 *    like the reference.  A context owns a HIP stream and a workspace; use one
 *    context per host thread.
 *  - Status codes mirror `RandNLAError` (src/errors.rs:3-14); the message is
 *    available from randnla_last_error().  No C++ exception crosses the ABI.
 */
#ifndef RANDNLA_H
#define RANDNLA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: src/errors.rs:3-14 ---------------------------------- */
enum randnla_status {
  RANDNLA_OK = 0,
  RANDNLA_INVALID_PARAMETERS = 1,
  RANDNLA_INVALID_DIMENSIONS = 2,
  RANDNLA_NEGATIVE_DIMENSIONS = 3,
  RANDNLA_NOT_OVERDETERMINED = 4,
  RANDNLA_NOT_SQUARE = 5,
  RANDNLA_SINGULAR_MATRIX = 6,
  RANDNLA_MATRIX_DECOMPOSITION_ERROR = 7,
  RANDNLA_NOT_HERMITIAN = 8,
  RANDNLA_NOT_POSITIVE_SEMI_DEFINITE = 9,
  RANDNLA_COMPUTATION_ERROR = 10,
  /* A violated precondition that the reference reports by panicking
   * (e.g. src/cqrrpt.rs:29); the message is the reference's panic text. */
  RANDNLA_PANIC = 11,
  RANDNLA_HIP_ERROR = 100,
  RANDNLA_NO_DEVICE = 101
};

/* ---- enums ---------------------------------------------------------------*/
/* DistributionType, src/sketch.rs:9-13, plus the sparse-sign operator. */
enum randnla_dist {
  RANDNLA_GAUSSIAN = 0,
  RANDNLA_UNIFORM = 1,
  RANDNLA_RADEMACHER = 2,
  RANDNLA_SPARSE_SIGN = 3
};
/* MatrixAttribute, src/sketch.rs:18-21 */
enum randnla_attr { RANDNLA_ROW = 0, RANDNLA_COLUMN = 1 };
/* Generators of rust-random123/src/rng.rs:89 (ThreeFry2x64Rng) and :85
 * (Philox4x32Rng). */
enum randnla_gen { RANDNLA_THREEFRY2X64 = 0, RANDNLA_PHILOX4X32 = 1 };
enum randnla_precision {
  RANDNLA_FP64 = 0,   /* every product in FP64 */
  RANDNLA_BF16 = 1,   /* range finder in BF16 MFMA, final pass split-FP16 */
  RANDNLA_FP8 = 2     /* range finder in FP8-e4m3 MFMA, final pass split-FP16 */
};
enum randnla_dtype { RANDNLA_F64 = 0, RANDNLA_F32 = 1 };
enum randnla_solver { RANDNLA_CGLS = 0, RANDNLA_LSQR = 1 };
enum randnla_side { RANDNLA_LEFT = 0 /* S*A */, RANDNLA_RIGHT = 1 /* A*Omega */ };

typedef struct randnla_ctx randnla_ctx;

/* A counter-based stream: generator, key and 128-bit starting counter.
 * Block b of the stream is the block function at counter ctr + b. */
typedef struct {
  int32_t gen;
  uint64_t key[2];
  uint64_t ctr[2];
} randnla_rng;

typedef struct {
  int32_t precision;     /* randnla_precision */
  int32_t power_iters;   /* q: rounds of (A^T, A) after Y = A*Omega */
  int32_t sketch;        /* randnla_dist for the LS / CQRRPT sketch */
  int32_t zeta;          /* nonzeros per column of a sparse-sign sketch */
  int32_t solver;        /* randnla_solver for the preconditioned solve */
  int32_t faithful;      /* 0: the intended method (Gaussian test matrix,
                            power_iters rounds, CholeskyQR2); 1: what the
                            reference computes (tsog1's even-branch discard,
                            full-pivot LU Stabilizer, Householder Orth), FP64
                            only; used by rand_svd, QB1, RF1 and tsog1 */
  randnla_rng rng;       /* stream of the random test matrix / sketch */
} randnla_opts;

/* Outcome of an iterative solve. istop/itn follow src/solvers.rs:125-136
 * for LSQR; for CGLS istop = 1 converged (||A^T r|| < tol), 7 = iteration
 * limit (src/cg.rs:46-59). */
typedef struct {
  int32_t istop;
  int32_t itn;
  double r1norm, r2norm, anorm, acond, arnorm, xnorm;
} randnla_solve_info;

/* ---- context ------------------------------------------------------------ */
int randnla_create(randnla_ctx** ctx, int device);
void randnla_destroy(randnla_ctx* ctx);
const char* randnla_last_error(const randnla_ctx* ctx);
/* Run on a caller stream (hipStream_t) instead of the context's own.  The
 * previous stream is drained and its cached workspace released first. */
int randnla_set_stream(randnla_ctx* ctx, void* stream);
/* The stream the context's work runs on (hipStream_t).  Device buffers that
 * other streams produce or consume must be ordered against it by the caller
 * (an event recorded on the producer and waited for on this stream, and the
 * reverse for consumers); host buffers need nothing, as every call returns
 * with its work complete. */
int randnla_get_stream(const randnla_ctx* ctx, void** stream);
void randnla_opts_default(randnla_opts* opts);
/* Per-kernel-class timing with HIP events on the context's stream.  Class
 * names used by rand_svd: "convert_bf16", "gemm_bf16_nn", "gemm_bf16_tn",
 * "gemm_f16", "gemm_f16x2", "gemm_bt", "gemm_fp8_nn", "quantize_fp8",
 * "gram", "cholesky" (factor and triangular inverse), "trmm", "rprod",
 * "jacobi", "svd_u", "svd_v"; by the other entry points: "gemm_bf16",
 * "gemm_f16x3", "gemm_f64", "gemm_fp8", "pass_a", "pass_at", "pass_bt",
 * "sketch_gemm", "sketch_apply", "gemv", "householder", "qrcp", "pinv",
 * "psd_check", "evd_y", "evd_sy", "evd_c", "evd_v", "id_y", "id_t",
 * "lsqr_pass", "batch_ls", "probe", "mlp_forward", "mlp_grad", "mlp_train". */
int randnla_timing_enable(randnla_ctx* ctx, int on);
int randnla_timing_query(randnla_ctx* ctx, const char* kernel_class,
                         double* total_ms, int64_t* launches);
int randnla_timing_reset(randnla_ctx* ctx);

/* ---- counter-based RNG (host) -------------------------------------------
 * Block functions of rust-random123; bit-exact with its KATs. */
void randnla_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]); /* philox.rs:211-223 */
void randnla_philox2x32(const uint32_t ctr[2], uint32_t key, uint32_t out[2]);          /* philox.rs:183-195 */
void randnla_philox2x64(const uint64_t ctr[2], uint64_t key, uint64_t out[2]);          /* philox.rs:197-209 */
void randnla_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]); /* philox.rs:225-237 */
void randnla_threefry2x64(const uint64_t ctr[2], const uint64_t key[2], uint64_t out[2]); /* threefry.rs:69-93 */
/* Rng::seed_from_u64(seed) with counter 0 (rand_core 0.6.4; threefry.rs:23-27,
 * philox.rs:101-105). */
void randnla_rng_from_seed(int gen, uint64_t seed, randnla_rng* out);
/* Words [first, first+n) of the u64 stream, generated on the GPU
 * (BlockRng64 order; threefry.rs:13-21). */
int randnla_stream_words(randnla_ctx* ctx, const randnla_rng* rng, uint64_t first,
                         int64_t n, uint64_t* out);

/* ---- sketch operators: src/sketch.rs --------------------------------------
 * sketching_operator(dist, rows, cols) (src/sketch.rs:102-130): dense
 * rows x cols operator, entry e = i + j*rows; rng NULL = the reference's
 * stream ThreeFry2x64Rng::seed_from_u64(0).  For RANDNLA_SPARSE_SIGN the
 * densified operator is returned (zeta nonzeros +-1/sqrt(zeta) per column). */
int randnla_sketching_operator(randnla_ctx* ctx, int dist, int64_t rows, int64_t cols,
                               const randnla_rng* rng, int zeta, double* out, int64_t ld);
/* haar_sample(rows, cols, attr) (src/sketch.rs:45-85). */
int randnla_haar_sample(randnla_ctx* ctx, int64_t rows, int64_t cols, int attr,
                        const randnla_rng* rng, double* out, int64_t ld);
/* Fused sketch application; the operator is generated inside the product and
 * never stored.  LEFT: out (d x n) = S (d x m) * A.  RIGHT: out (m x d) =
 * A * Omega (n x d).  A is F64 or F32 (dtype); out is F64. */
int randnla_sketch_apply(randnla_ctx* ctx, int side, int dist, int zeta,
                         const randnla_rng* rng, const void* A, int dtype,
                         int64_t m, int64_t n, int64_t lda, int64_t d,
                         double* out, int64_t ldo);

/* No counterpart in the reference.  randnla_sketch_apply for callers that
 * cannot allocate device memory (code under stream capture, tight serving
 * loops, allocation-free harnesses): same operators, arguments and results,
 * but the partial sums live in a fixed 64 MiB array per GPU held by the
 * library (LEFT products need d * n <= 2^23; larger ones return
 * RANDNLA_INVALID_DIMENSIONS).  A and out must be device memory (no staging);
 * the work is queued on the context's stream and the call returns without
 * waiting.  Calls that overlap on one GPU race on the array.  RIGHT with a
 * sparse-sign Omega is not available (RANDNLA_INVALID_PARAMETERS). */
int randnla_sketch_apply_noalloc(randnla_ctx* ctx, int side, int dist, int zeta,
                                 const randnla_rng* rng, const void* A, int dtype,
                                 int64_t m, int64_t n, int64_t lda, int64_t d,
                                 double* out, int64_t ldo);

/* ---- low-rank approximation: src/lora_helpers.rs, src/lora_drivers.rs ----
 * rand_svd(A, k, epsilon, s) (src/lora_drivers.rs:30-69): U (m x r),
 * S (r singular values, descending), Vt (r x n), r = min(k, m, n): the
 * reference's r = min(k, ncols(Q)) with Q of min(k + s, m, n) columns
 * (src/lora_drivers.rs:51); any k > 0 is accepted, as there.
 * opts NULL = defaults (FP64, q = 1 as RF1's tsog1(A, k, 2, 1), Philox4x32
 * seed_from_u64(42)); opts->faithful = 1 gives the reference's own
 * computation (deterministic in A). */
int randnla_rand_svd(randnla_ctx* ctx, const void* A, int dtype, int64_t m, int64_t n,
                     int64_t lda, int64_t k, double epsilon, int64_t s,
                     const randnla_opts* opts, double* U, int64_t ldu, double* S,
                     double* Vt, int64_t ldvt, int64_t* r_out);
/* QB1(A, k, eps) (src/lora_helpers.rs:17-23): Q (m x kc), B = Q^T A (kc x n),
 * kc = min(k, m, n); the rest of the m x k / k x n outputs is zero. */
int randnla_qb1(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                int64_t k, double epsilon, const randnla_opts* opts, double* Q,
                int64_t ldq, double* B, int64_t ldb);
/* RF1(A, k) (src/lora_helpers.rs:37-44): orthonormal Q (m x kc),
 * kc = min(k, m, n); the rest of the m x k output is zero. */
int randnla_rf1(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                int64_t k, const randnla_opts* opts, double* Q, int64_t ldq);
/* tsog1(A, k, num_passes, passes_per_stab) (src/lora_helpers.rs:58-105):
 * intended method: n x k test matrix after num_passes products with A / A^T,
 * stabilised (re-orthonormalised) after every passes_per_stab products.
 * opts->faithful: the reference's result, n x kc with kc = k when the loop
 * does not run (num_passes < 2; the Gaussian for 0, zeros for 1), else the
 * columns left by the last loop iteration's stabilisers (min(n, min(m, k))
 * for passes_per_stab = 1); the rest of the n x k output is zero.  The
 * faithful Gaussian comes from the reference's stream ThreeFry2x64
 * seed_from_u64(0) (src/sketch.rs:112); opts->rng is not used. */
int randnla_tsog1(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                  int64_t k, int num_passes, int passes_per_stab,
                  const randnla_opts* opts, double* S, int64_t lds);
/* Orth(X) (src/lora_helpers.rs:131-133): Q (m x min(m, n)) of the
 * Householder thin QR X = Q R with nalgebra's convention R_ii >= 0 (a zero
 * column gives a unit vector: Orth(0) = I[:, :min(m, n)]); wide X gives an
 * m x m orthogonal Q.  R (min(m, n) x n, may be NULL). */
int randnla_orth(randnla_ctx* ctx, const double* X, int64_t m, int64_t n, int64_t ldx,
                 double* Q, int64_t ldq, double* R, int64_t ldr);
/* CholeskyQR2 of a tall X (m x n, m >= n): Q (m x n), R = R2 R1 (n x n, may
 * be NULL).  Dependent columns (Cholesky pivot <= 16 max(n, 16) eps times the
 * column's diagonal) get a zero row in R1 and a unit-vector completion in Q.
 * No reference counterpart: the orthonormalisation of the range finder. */
int randnla_cholqr2(randnla_ctx* ctx, const double* X, int64_t m, int64_t n, int64_t ldx,
                    double* Q, int64_t ldq, double* R, int64_t ldr);
/* Stabilizer(X) (src/lora_helpers.rs:144-146): L (m x min(m, n)), the unit
 * lower factor of X's full-pivot LU (nalgebra 0.33 FullPivLU: pivot = first
 * maximum of |x| over the trailing block in column-major order, stop at a
 * zero maximum), rows in pivoted order.  Stabilizer(0) = I[:, :min(m, n)]. */
int randnla_stabilizer(randnla_ctx* ctx, const double* X, int64_t m, int64_t n,
                       int64_t ldx, double* Y, int64_t ldy);
/* rand_evd1(A, k, epsilon, s) (src/lora_drivers.rs:87-151): A (m x n) must
 * be square and exactly symmetric (RANDNLA_NOT_HERMITIAN otherwise, NaN
 * entries included).  V (n x r, orthonormal columns) and lambda (r values,
 * |lambda| descending) with r = min(k, min(k + s, n)) approximate A's
 * dominant eigenpairs: Q = RF1(A, min(k + s, n)) (opts as for rand_svd),
 * eigenpairs of Q^T A Q.  epsilon is checked and otherwise unused, as in the
 * reference.  V and lambda may be device or host pointers. */
int randnla_rand_evd1(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                      int64_t k, double epsilon, int64_t s, const randnla_opts* opts,
                      double* V, int64_t ldv, double* lambda, int64_t* r_out);
/* rand_evd2(A, k, s) (src/lora_drivers.rs:167-224): Nystrom approximation of
 * a positive semi-definite A (n x n) from S = tsog1(A, k + s, 3, 1) with the
 * nu = sqrt(n) eps ||A S||_F shift.  RANDNLA_NOT_POSITIVE_SEMI_DEFINITE when
 * the Cholesky factorisation of A's lower-triangle symmetric matrix shifted by
 * n eps max|a_ii| fails (the reference tests its eigenvalues for < 0);
 * RANDNLA_MATRIX_DECOMPOSITION_ERROR when the Cholesky factorisation of
 * S^T (A S + nu S) fails (the zero matrix does).  r = min(k, #{sigma_i^2 >
 * nu}) eigenpairs, lambda descending; k + s must not exceed n. */
int randnla_rand_evd2(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                      int64_t k, int64_t s, const randnla_opts* opts, double* V, int64_t ldv,
                      double* lambda, int64_t* r_out);

/* ---- interpolative decompositions: src/id.rs, src/pivot_decompositions.rs ----
 * The reference's assert! panics return RANDNLA_PANIC with its text
 * ("k must be positive)", "k must be <= min(l,w)"; economic_qrcp: "k must be
 * <= min(m,n)", "k must be positive").  Index outputs are int64 (host or
 * device).  attr: RANDNLA_COLUMN or RANDNLA_ROW. */
/* economic_qrcp(a, k) (src/pivot_decompositions.rs:194-269): the first k
 * steps of Householder QRCP (pivot = first maximum of the recomputed trailing
 * column norms).  Q (m x k, may be NULL), R (k x n), perm (n). */
int randnla_economic_qrcp(randnla_ctx* ctx, const double* A, int64_t m, int64_t n,
                          int64_t lda, int64_t k, double* Q, int64_t ldq, double* R,
                          int64_t ldr, int64_t* perm);
/* osid_qrcp(y, k, attr) (src/id.rs:248-305).  Column: X (k x n) and J (k)
 * with Y ~ Y[:, J] X, X[:, J] = I.  Row: X (m x k), J (k) with Y ~ X Y[J, :]. */
int randnla_osid_qrcp(randnla_ctx* ctx, const double* Y, int64_t m, int64_t n, int64_t ldy,
                      int64_t k, int attr, double* X, int64_t ldx, int64_t* J);
/* osid_randomised(a, k, attr) (src/id.rs:205-234).  Column: the column ID of
 * S A with S a k x m Gaussian (opts NULL: the ThreeFry2x64 seed-0 stream of
 * the dense sketches).  Row: the row ID of A S^T with S = tsog1(A, k, 2, 1),
 * defined for n == k only (RANDNLA_PANIC otherwise, as the reference's
 * product fails). */
int randnla_osid_randomised(randnla_ctx* ctx, const double* A, int64_t m, int64_t n,
                            int64_t lda, int64_t k, int attr, const randnla_opts* opts,
                            double* X, int64_t ldx, int64_t* J);
/* two_sided_id(a, k) (src/id.rs:125-132): A ~ Z A[I, J] X; Z (m x k),
 * X (k x n), I and J (k). */
int randnla_two_sided_id(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                         int64_t k, double* Z, int64_t ldz, int64_t* I, int64_t* J,
                         double* X, int64_t ldx);
/* two_sided_id_randomised(a, k) (src/id.rs:94-101). */
int randnla_two_sided_id_randomised(randnla_ctx* ctx, const double* A, int64_t m, int64_t n,
                                    int64_t lda, int64_t k, const randnla_opts* opts,
                                    double* Z, int64_t ldz, int64_t* I, int64_t* J, double* X,
                                    int64_t ldx);
/* cur(a, k) (src/id.rs:33-76): A ~ A[:, J] U A[I, :]; U (k x k). */
int randnla_cur(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda, int64_t k,
                int64_t* J, double* U, int64_t ldu, int64_t* I);
/* cur_randomised(a, k) (src/id.rs:154-197). */
int randnla_cur_randomised(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                           int64_t k, const randnla_opts* opts, int64_t* J, double* U,
                           int64_t ldu, int64_t* I);

/* ---- least squares: src/sketch_and_precondition.rs, src/cg.rs,
 *      src/solvers.rs ----------------------------------------------------- */
/* blendenpik_overdetermined(a, b, epsilon, l, sampling_factor)
 * (src/sketch_and_precondition.rs:26-59). */
int randnla_blendenpik(randnla_ctx* ctx, const double* A, int64_t m, int64_t n,
                       int64_t lda, const double* b, double epsilon, int64_t l,
                       double sampling_factor, const randnla_opts* opts, double* x,
                       randnla_solve_info* info);
/* lsrn_overdetermined(...) (src/sketch_and_precondition.rs:82-119). */
int randnla_lsrn(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                 const double* b, double epsilon, int64_t l, double sampling_factor,
                 const randnla_opts* opts, double* x, randnla_solve_info* info);
/* sketched_least_squares_qr(a, b) (src/sketch_and_solve.rs:24-33): S with
 * m / 4 rows (Gaussian by default; opts selects the distribution and the
 * stream, default the ThreeFry stream of seed 0), QR of S A, x = R^{-1} Q^T S b
 * with the zero-diagonal rule of solve_upper_triangular_system.  x has n
 * entries; for m / 4 < n the reference returns m / 4 entries, which are the
 * first m / 4 here (the rest 0). */
int randnla_sketched_least_squares_qr(randnla_ctx* ctx, const double* A, int64_t m,
                                      int64_t n, int64_t lda, const double* b,
                                      const randnla_opts* opts, double* x);
/* sketched_least_squares_svd(a, b) (src/sketch_and_solve.rs:53-65): thin SVD
 * S A = U diag(sigma) V^T, x = V diag(1 / sigma) U^T S b with the rule of
 * solve_diagonal_system (sigma == 0 -> 0). */
int randnla_sketched_least_squares_svd(randnla_ctx* ctx, const double* A, int64_t m,
                                       int64_t n, int64_t lda, const double* b,
                                       const randnla_opts* opts, double* x);
/* sketch_saddle_point_precondition(a, b, c, mu, epsilon, l, sampling_factor)
 * (src/sketch_and_precondition.rs:150-216); c may be NULL (empty). */
int randnla_saddle(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                   const double* b, const double* c, double mu, double epsilon,
                   int64_t l, double sampling_factor, const randnla_opts* opts,
                   double* x, double* y, randnla_solve_info* info);
/* cgls(a, b, tolerance, num_iterations, x) (src/cg.rs:18-61); x0 may be NULL. */
int randnla_cgls(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                 const double* b, double tolerance, int64_t num_iterations,
                 const double* x0, double* x, randnla_solve_info* info);
/* lsqr(a, b, damp, atol, btol, conlim, iter_lim, calc_var, x0)
 * (src/solvers.rs:115-278).  iter_lim <= 0 means 2n.  var and arnorms may be
 * NULL; arnorms has room for iter_lim entries, its count is info->itn. */
int randnla_lsqr(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                 const double* b, double damp, double atol, double btol, double conlim,
                 int64_t iter_lim, int calc_var, const double* x0, double* x,
                 double* var, double* arnorms, randnla_solve_info* info);
/* solve_upper_triangular_system(u, y) (src/solvers.rs:22-41); zero
 * diagonal entries leave x_i = 0. */
int randnla_solve_upper_triangular(randnla_ctx* ctx, const double* U, int64_t n,
                                   int64_t ldu, const double* y, double* x);
int randnla_solve_upper_triangular_f32(randnla_ctx* ctx, const float* U, int64_t n,
                                       int64_t ldu, const float* y, float* x);

/* ---- pivoted QR / CQRRPT: src/pivot_decompositions.rs, src/cqrrpt.rs ---- */
/* qrcp(a) (src/pivot_decompositions.rs:106-175): R (m x n, upper), perm
 * (n), and, when Q != NULL, Q (m x m). */
int randnla_qrcp(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                 double* Q, int64_t ldq, double* R, int64_t ldr, int64_t* perm);
/* sap_chol_qrcp(a, d) (src/cqrrpt.rs:27-58): Q (m x k), R (k x n), J (n),
 * k = numerical rank of the sketch.  A d outside [n, m] returns
 * RANDNLA_PANIC with the reference's message "d must satisfy n ≤ d ≪ m". */
int randnla_cqrrpt(randnla_ctx* ctx, const double* A, int64_t m, int64_t n, int64_t lda,
                   int64_t d, const randnla_opts* opts, double* Q, int64_t ldq,
                   double* R, int64_t ldr, int64_t* J, int64_t* k_out);

/* ---- batched uneven least squares ------------------------------------------
 * Thousands of independent overdetermined problems min ||A_i x_i - b_i||_2 in
 * one persistent-kernel launch replayed from a captured hipGraph.  The
 * reference has no batched entry point: each problem is solved the way
 * blendenpik_overdetermined (src/sketch_and_precondition.rs:26-59) solves
 * one with the LSQR solver (src/solvers.rs:115-278): Gaussian sketch with
 * d_i = min(m_i, floor(sampling_factor * n_i)) rows (sampling_factor >= 1,
 * same checks and messages as :33-46), R_i from the sketch, LSQR on
 * A_i R_i^{-1} from z0 = Q^T S b with atol = btol = tol, conlim 1e8, and
 * iter_lim iterations at most.  Precision: A_i FP32 (the sketch product runs
 * on the FP32 matrix cores), b_i, x_i and everything after the sketch FP64.
 * Limits: 1 <= n_i <= 256, n_i <= m_i, lda_i >= m_i.  A, b and x are device
 * pointers of the context's GPU (any size; offsets are in elements). */
typedef struct {
  int64_t m, n, lda;   /* A_i is m x n, column-major, leading dimension lda */
  int64_t a_offset;    /* first element of A_i in the A buffer */
  int64_t b_offset;    /* first element of b_i (m entries) in b */
  int64_t x_offset;    /* first element of x_i (n entries) in x */
  int64_t stream_id;   /* sketch substream of problem i (counter + ((id+1) << 96));
                          equal ids give equal sketches in any batch position */
} randnla_ls_problem;

typedef struct {
  int32_t status;      /* 0 solved; 1 the sketch's Gram is not positive definite
                          (rank-deficient A_i, x_i = 0); 2 internal wait limit */
  int32_t istop;       /* LSQR stop reason 0..7 (src/solvers.rs:264-270) */
  int32_t itn;         /* LSQR iterations */
  int32_t d;           /* sketch rows used */
  double r2norm, arnorm, anorm, acond, xnorm;   /* LSQR outputs for A_i R_i^{-1} */
} randnla_ls_info;

typedef struct randnla_batch_ls randnla_batch_ls;

/* Plan: validates the problems, lays out the workspace, orders the work
 * queue (largest first) and captures the launch.  autotune != 0 makes the
 * first execute time every LSQR chunk size x work-groups-per-CU
 * configuration on that batch and keep the fastest. */
int randnla_batch_ls_create(randnla_ctx* ctx, const randnla_ls_problem* problems, int64_t count,
                            double sampling_factor, double tol, int64_t iter_lim,
                            const randnla_opts* opts, int autotune, randnla_batch_ls** plan);
/* Solve every problem; info (count entries, host memory) may be NULL.
 * Returns RANDNLA_COMPUTATION_ERROR if a problem hit the internal wait limit. */
int randnla_batch_ls_execute(randnla_batch_ls* plan, const float* A, const double* b, double* x,
                             randnla_ls_info* info);
/* The sketches [S_i A_i | S_i b_i] (d_i x (n_i + 1), column-major, packed in
 * problem order; FP32) of the last execute; out may be host or device. */
int randnla_batch_ls_sketch(randnla_batch_ls* plan, float* out, int64_t out_elems);
/* The chosen configuration: halvings of the LSQR rows per wave step (0 = the
 * rows that fill the registers for the problem's n), work-groups per CU and,
 * after autotuning, the best measured time per batch (ms; -1 otherwise). */
int randnla_batch_ls_config(const randnla_batch_ls* plan, int* row_step_shift, int* wgs_per_cu,
                            double* best_ms);
void randnla_batch_ls_destroy(randnla_batch_ls* plan);
/* One-shot form: plan, execute, destroy. */
int randnla_batch_ls_solve(randnla_ctx* ctx, const randnla_ls_problem* problems, int64_t count,
                     const float* A, const double* b, double* x, double sampling_factor,
                     double tol, int64_t iter_lim, const randnla_opts* opts,
                     randnla_ls_info* info);

/* ---- learned (k, p, q) proposer for rand_svd --------------------------------
 * No reference counterpart (the reference leaves k, s and the power-iteration
 * count to the caller, src/lora_drivers.rs:30-69).  A 34 -> 128 -> 128 -> 3
 * MLP with GELU activations; weights and activations rounded to BF16, products
 * on the BF16 matrix cores with FP32 accumulation.  Parameters are FP32,
 * packed W1[128][34], b1[128], W2[128][128], b2[128], W3[3][128], b3[3]
 * (randnla_mlp_param_count() values).  X is B x 34 (row-major: sample-major),
 * targets T and outputs Y are B x 3.  The training loss is
 * sum_s sum_o wo[o] (Y_so - T_so)^2 / B. */
int randnla_mlp_param_count(void);
int randnla_mlp_forward(randnla_ctx* ctx, const float* params, const float* X, int64_t B,
                        float* Y);
int randnla_mlp_grad(randnla_ctx* ctx, const float* params, const float* X, const float* T,
                     const float* wo, int64_t B, float* grad, float* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected with
 * `step` >= 1) on a minibatch; device pointers; loss may be NULL. */
int randnla_mlp_train_step(randnla_ctx* ctx, float* params, float* adam_m, float* adam_v,
                           const float* X, const float* T, const float* wo, int64_t B, int step,
                           float lr, float* loss);
/* The proposer's 34 inputs for A (m x n, F64 or F32): the top 32 singular
 * values of A * Omega0 (Omega0 n x 48 Gaussian from rng; Philox4x32 seed 42
 * when rng is NULL) as log10(s_i / s_0) / 12 floored at -1, then ln(m) / 10
 * and ln(n) / 10.  out: 34 floats, host or device. */
int randnla_probe_features(randnla_ctx* ctx, const void* A, int dtype, int64_t m, int64_t n,
                           int64_t lda, const randnla_rng* rng, float* out);

/* ---- test matrices: src/test_assist.rs ------------------------------------
 * A = D_r H_m[:, J] diag(sigma) (H_n D_c)^T / sqrt(m n) + noise * G, with
 * H Sylvester-Hadamard (m, n powers of two, n <= m), D_r, D_c random signs,
 * J(j) = (a j + b) mod m from the stream, G standard normal; out is F64 or
 * F32 (dtype).  Exact singular values sigma when noise == 0. */
int randnla_test_matrix(randnla_ctx* ctx, int64_t m, int64_t n, const double* sigma,
                        double noise, const randnla_rng* rng, void* out, int dtype,
                        int64_t ld);

/* ---- measurement helpers (no reference counterpart) -----------------------
 * ||A - U diag(S) Vt||_2 (err) and ||A||_2 (anorm) by Golub-Kahan-Lanczos
 * bidiagonalisation with full reorthogonalisation, `iters` steps (at most
 * min(m, n, 512)) from a fixed Gaussian start; each estimate is a lower bound
 * that converges from below.  Either output may be NULL. */
int randnla_spectral_error(randnla_ctx* ctx, const void* A, int dtype, int64_t m, int64_t n,
                           int64_t lda, const double* U, int64_t ldu, const double* S,
                           int64_t r, const double* Vt, int64_t ldvt, int iters, double* err,
                           double* anorm);
/* C (M x N, FP64) = op(A) * B on the library's MFMA kernels; op(A) = A^T
 * when trans_a.  precision: RANDNLA_FP64, RANDNLA_BF16 (operands rounded to
 * BF16, FP32 accumulation), 3 (FP32 operands split into FP16 hi/lo pairs,
 * three FP16 products), 4 (op(A) split into hi/lo, B rounded once to FP16:
 * two products), 5 (both rounded once to FP16: one product) or 6 (op(A)
 * rounded once, B split into hi/lo: two products); FP16 operands
 * are scaled by powers of two that bring their maxima into [2^13, 2^14].
 * Exposes the kernels behind every product above. */
int randnla_gemm(randnla_ctx* ctx, int precision, int trans_a, int64_t M, int64_t N, int64_t K,
                 const void* A, int dtype_a, int64_t lda, const void* B, int dtype_b,
                 int64_t ldb, double* C, int64_t ldc);
/* Diagnostics of the last call: index 0 = Jacobi sweeps of the small SVD. */
int randnla_diag(const randnla_ctx* ctx, int index, double* out);

#ifdef __cplusplus
}
#endif
#endif /* RANDNLA_H */
