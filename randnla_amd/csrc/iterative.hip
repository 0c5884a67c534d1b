// Device-resident recurrences of the iterative least-squares solvers.
//
// LSQR follows src/solvers.rs:115-278 (Paige & Saunders, as translated from
// scipy 1.14.1) and CGLS follows src/cg.rs:18-61.  All scalar state lives in
// device memory and is advanced by single-thread kernels between the
// vector kernels; once a stop condition fires every kernel of the loop
// becomes a no-op, so the host only needs to poll the stop flag every few
// iterations.
#include <algorithm>

#include "common.h"
#include "iterative.h"

namespace rna {
namespace {

constexpr int TPB = 256;

__device__ __forceinline__ double dsign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// Sum of per-block partials into *out.
__global__ void finish_sum(const double* part, int nb, double* out) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) *out = s;
}

// ---- LSQR ----------------------------------------------------------------
// u = t - alfa * u ; partial sums of u^2.
__global__ __launch_bounds__(TPB) void lsqr_u_kernel(LsqrState* S, int64_t m,
                                                     const double* __restrict__ t,
                                                     double* __restrict__ u, double* part) {
  __shared__ double sh[TPB / 64];
  if (S->istop) return;
  const double alfa = S->alfa;
  double q = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < m; i += (int64_t)gridDim.x * TPB) {
    const double v = t[i] - alfa * u[i];
    u[i] = v;
    q += v * v;
  }
  q = block_sum(q, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = q;
}

// beta = ||u||; if beta > 0: anorm update and flag to scale u.
__global__ void lsqr_beta_kernel(LsqrState* S, const double* part, int nb) {
  __shared__ double sh[4];
  if (S->istop) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    const double beta = sqrt(s);
    S->beta = beta;
    if (beta > 0.0) {
      S->anorm = sqrt(S->anorm * S->anorm + S->alfa * S->alfa + beta * beta + S->dampsq);
      S->uscale = 1.0 / beta;
    } else {
      S->uscale = 1.0;
    }
  }
}

__global__ void scale_kernel(const LsqrState* S, int which, int64_t n, double* x) {
  if (S->istop) return;
  const double a = which == 0 ? S->uscale : S->vscale;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= a;
}

// v = t - beta * v (only when beta > 0); partial sums of v^2 and w^2.
__global__ __launch_bounds__(TPB) void lsqr_v_kernel(LsqrState* S, int64_t n,
                                                     const double* __restrict__ t,
                                                     double* __restrict__ v,
                                                     const double* __restrict__ w, double* part,
                                                     int nbp) {
  __shared__ double sh[TPB / 64];
  if (S->istop) return;
  const bool upd = S->beta > 0.0;
  const double beta = S->beta;
  double q = 0.0, r = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    double x = v[i];
    if (upd) { x = t[i] - beta * x; v[i] = x; }
    q += x * x;
    r += w[i] * w[i];
  }
  q = block_sum(q, sh);
  r = block_sum(r, sh);
  if (threadIdx.x == 0) { part[blockIdx.x] = q; part[nbp + blockIdx.x] = r; }
}

// ---- fused LSQR pass (single problem, n <= FP_MAXN) ---------------------
// One pass over A per iteration instead of two: with p = A w - alfa u
// (w = P v, the preconditioned direction), beta = ||p|| and u' = p / beta,
// A^T u' = (A^T p) / beta, so A^T p and ||p||^2 are accumulated while p is
// formed.  Each work-group (persistent, two per CU) walks its rows in chunks
// of FP_RB rows: the chunk (FP_RB x n) is staged in LDS as [column][row]
// (the next chunk is prefetched into registers meanwhile), then
//   phase 1: p_r = sum_j A[r][j] w_j - alfa u_r  (FP_NG column groups of
//            FP_GW, summed through LDS in a fixed order), stored into u;
//   phase 2: s_j += sum_r A[r][j] p_r for the work-group's columns j.
// Outputs: part[g] = sum of p_r^2 over work-group g's rows, spart[g * n + j].
constexpr int FP_THR = 256, FP_RB = 16, FP_MAXN = 512, FP_LD = FP_RB + 1;
constexpr int FP_NG = FP_THR / FP_RB;                    // phase-1 column groups (16)
constexpr int FP_GW = FP_MAXN / FP_NG;                   // columns per group (32)
constexpr int FP_PF = FP_RB * FP_MAXN / FP_THR;          // prefetch registers per thread (32)
constexpr size_t FP_LDS =
    ((size_t)FP_MAXN * FP_LD + FP_MAXN + FP_RB + FP_NG * FP_RB) * sizeof(double);

__global__ __launch_bounds__(FP_THR, 2) void lsqr_fused_kernel(const LsqrState* S,
                                                               const double* __restrict__ A,
                                                               int64_t m, int n, int64_t lda,
                                                               const double* __restrict__ w,
                                                               double* __restrict__ u,
                                                               double* part, double* spart,
                                                               int64_t rows_per_wg) {
  if (S->istop) return;
  extern __shared__ double sm[];
  double* Ac = sm;                                  // [j][r], ld FP_LD
  double* wl = Ac + (size_t)FP_MAXN * FP_LD;       // w
  double* pl = wl + FP_MAXN;                        // p of the chunk
  double* red = pl + FP_RB;                         // [FP_NG][FP_RB] phase-1 partials
  __shared__ double sh[FP_THR / 64];
  const int t = threadIdx.x;
  const double au = S->alfa * S->umul;               // alfa times the pending u scale
  // phase 1 reads columns in fours: w and the (never stored) columns n..n+3
  // of the chunk are zero there
  for (int j = t; j < FP_MAXN; j += FP_THR) wl[j] = j < n ? w[j] : 0.0;
  for (int e = t; e < 4 * FP_LD; e += FP_THR)
    if ((int64_t)n * FP_LD + e < (int64_t)FP_MAXN * FP_LD) Ac[n * FP_LD + e] = 0.0;
  const int64_t r_beg = (int64_t)blockIdx.x * rows_per_wg;
  const int64_t r_end = min(m, r_beg + rows_per_wg);
  const int nchunk = r_end > r_beg ? (int)((r_end - r_beg + FP_RB - 1) / FP_RB) : 0;
  const int nel = FP_RB * n;                        // elements per chunk
  double pf[FP_PF];
  auto prefetch = [&](int c) {
    const int64_t r0 = r_beg + (int64_t)c * FP_RB;
#pragma unroll
    for (int i = 0; i < FP_PF; ++i) {
      const int e = t + FP_THR * i, r = e % FP_RB, j = e / FP_RB;
      const int64_t row = r0 + r;
      pf[i] = (e < nel && row < r_end) ? A[(int64_t)j * lda + row] : 0.0;
    }
  };
  double s0 = 0.0, s1 = 0.0;                        // A^T p for columns t, t + 256
  double q = 0.0;                                   // sum p^2 (thread t < FP_RB)
  if (nchunk > 0) prefetch(0);
  for (int c = 0; c < nchunk; ++c) {
    const int64_t r0 = r_beg + (int64_t)c * FP_RB;
    __syncthreads();                                // previous chunk's phase 2 done
#pragma unroll
    for (int i = 0; i < FP_PF; ++i) {
      const int e = t + FP_THR * i;
      if (e < nel) Ac[(e / FP_RB) * FP_LD + (e % FP_RB)] = pf[i];
    }
    if (c + 1 < nchunk) prefetch(c + 1);
    __syncthreads();
    // phase 1
    {
      const int r = t % FP_RB, g = t / FP_RB;
      const int j0 = g * FP_GW;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
      for (int jj = 0; jj < FP_GW; jj += 4) {
        const int j = j0 + jj;                      // columns past n are zero in Ac
        if (j < n) {
          a0 += Ac[j * FP_LD + r] * wl[j];
          a1 += Ac[(j + 1) * FP_LD + r] * wl[j + 1];
          a2 += Ac[(j + 2) * FP_LD + r] * wl[j + 2];
          a3 += Ac[(j + 3) * FP_LD + r] * wl[j + 3];
        }
      }
      red[g * FP_RB + r] = (a0 + a1) + (a2 + a3);
    }
    __syncthreads();
    if (t < FP_RB) {
      const int64_t row = r0 + t;
      double pr = 0.0;
      if (row < r_end) {
        double a = 0.0;
#pragma unroll
        for (int g = 0; g < FP_NG; ++g) a += red[g * FP_RB + t];
        pr = a - au * u[row];
        u[row] = pr;
        q += pr * pr;
      }
      pl[t] = pr;
    }
    __syncthreads();
    // phase 2
    {
      if (t < n) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int r = 0; r < FP_RB; r += 2) {
          a += Ac[t * FP_LD + r] * pl[r];
          b += Ac[t * FP_LD + r + 1] * pl[r + 1];
        }
        s0 += a + b;
      }
      if (t + FP_THR < n) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int r = 0; r < FP_RB; r += 2) {
          a += Ac[(t + FP_THR) * FP_LD + r] * pl[r];
          b += Ac[(t + FP_THR) * FP_LD + r + 1] * pl[r + 1];
        }
        s1 += a + b;
      }
    }
  }
  if (t < n) spart[(int64_t)blockIdx.x * n + t] = s0;
  if (t + FP_THR < n) spart[(int64_t)blockIdx.x * n + t + FP_THR] = s1;
  q = block_sum(q, sh);
  if (t == 0) part[blockIdx.x] = q;
}

// beta = ||p|| from the per-group sums part[0..ng) (every block forms the
// same sum in the same order; block 0 advances beta, anorm, uscale and
// umul), then out_j = sum_g spart[g * n + j] / beta (= A^T u' of the fused
// pass).  A block of FF_COLS x FF_SL threads owns FF_COLS columns; slice y
// sums the groups y, y + FF_SL, ... and the slices are added in a fixed
// order.
constexpr int FF_COLS = 32, FF_SL = 32;
__global__ __launch_bounds__(FF_COLS* FF_SL) void lsqr_fused_finish(LsqrState* S,
                                                                    const double* part,
                                                                    const double* spart, int ng,
                                                                    int n, double* out) {
  if (S->istop) return;
  __shared__ double red[FF_SL][FF_COLS + 1];
  __shared__ double sh[FF_COLS * FF_SL / WAVE];
  const int x = threadIdx.x % FF_COLS, y = threadIdx.x / FF_COLS;
  double q = 0.0;
  for (int g = threadIdx.x; g < ng; g += FF_COLS * FF_SL) q += part[g];
  q = block_sum(q, sh);
  const double beta = sqrt(q), uscale = beta > 0.0 ? 1.0 / beta : 1.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->beta = beta;
    if (beta > 0.0)
      S->anorm = sqrt(S->anorm * S->anorm + S->alfa * S->alfa + beta * beta + S->dampsq);
    S->uscale = uscale;
    S->umul = uscale;
  }
  const int j = blockIdx.x * FF_COLS + x;
  double s0 = 0.0, s1 = 0.0;
  if (j < n) {
    int g = y;
    for (; g + FF_SL < ng; g += 2 * FF_SL) {
      s0 += spart[(int64_t)g * n + j];
      s1 += spart[(int64_t)(g + FF_SL) * n + j];
    }
    if (g < ng) s0 += spart[(int64_t)g * n + j];
  }
  red[y][x] = s0 + s1;
  __syncthreads();
  if (y == 0 && j < n) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < FF_SL; ++k) s += red[k][x];
    out[j] = uscale * s;
  }
}

__device__ void lsqr_scalar_step(LsqrState* S, double vv, double ww);

// v-phase of one LSQR iteration in one work-group (n <= blockDim.x * k):
// v = t - beta v (beta > 0), alfa = ||v|| and the scalar recurrence, v /=
// alfa, x += t1 w, var += (w / rho)^2, w = v + t2 w, then the commit.
__global__ __launch_bounds__(1024) void lsqr_vphase_kernel(LsqrState* S, int n,
                                                           const double* __restrict__ t,
                                                           double* __restrict__ v,
                                                           double* __restrict__ w,
                                                           double* __restrict__ x,
                                                           double* __restrict__ var, int calc_var,
                                                           double* arnorms, int cap) {
  __shared__ double sh[1024 / WAVE];
  __shared__ double sc[4];
  if (S->istop) return;
  const bool upd = S->beta > 0.0;
  const double beta = S->beta;
  double q = 0.0, r = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double xv = v[i];
    if (upd) { xv = t[i] - beta * xv; v[i] = xv; }
    q += xv * xv;
    r += w[i] * w[i];
  }
  q = block_sum(q, sh);
  r = block_sum(r, sh);
  if (threadIdx.x == 0) {
    lsqr_scalar_step(S, q, r);
    sc[0] = S->vscale; sc[1] = S->t1; sc[2] = S->t2; sc[3] = S->inv_rho;
  }
  __syncthreads();
  const double vs = sc[0], t1 = sc[1], t2 = sc[2], ir = sc[3];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double vi = v[i] * vs, wi = w[i];
    v[i] = vi;
    x[i] += t1 * wi;
    if (calc_var) var[i] += (wi * ir) * (wi * ir);
    w[i] = vi + t2 * wi;
  }
  if (threadIdx.x == 0) {
    S->istop = S->pending_stop;
    if (!S->istop && S->itn < cap) arnorms[S->itn] = S->arnorm;
  }
}

// The scalar part of one LSQR iteration (src/solvers.rs:197-272), from
// vv = ||t - beta v||^2 and ww = ||w||^2 (w before its update).
__device__ void lsqr_scalar_step(LsqrState* S, double vv, double ww) {
  const double eps = 2.220446049250313e-16;
  if (S->beta > 0.0) {
    const double alfa = sqrt(vv);
    S->alfa = alfa;
    S->vscale = alfa > 0.0 ? 1.0 / alfa : 0.0;
  } else {
    S->vscale = 1.0;
  }
  const double alfa = S->alfa, beta = S->beta, damp = S->damp;
  const double rhobar1 = sqrt(S->rhobar * S->rhobar + S->dampsq);
  const double cs1 = S->rhobar / rhobar1;
  const double sn1 = damp / rhobar1;
  const double psi = sn1 * S->phibar;
  S->phibar *= cs1;
  // sym_ortho(rhobar1, beta)  (src/solvers.rs:84-102)
  double cs, sn, rho;
  {
    const double a = rhobar1, b = beta;
    if (b == 0.0) { cs = dsign(a); sn = 0.0; rho = fabs(a); }
    else if (a == 0.0) { cs = 0.0; sn = dsign(b); rho = fabs(b); }
    else if (fabs(b) > fabs(a)) {
      const double tau = a / b;
      sn = dsign(b) / sqrt(1.0 + tau * tau);
      cs = sn * tau;
      rho = b / sn;
    } else {
      const double tau = b / a;
      cs = dsign(a) / sqrt(1.0 + tau * tau);
      sn = cs * tau;
      rho = a / cs;
    }
  }
  const double theta = sn * alfa;
  S->rhobar = -cs * alfa;
  const double phi = cs * S->phibar;
  S->phibar *= sn;
  const double tau = sn * phi;
  S->t1 = phi / rho;
  S->t2 = -theta / rho;
  S->inv_rho = 1.0 / rho;
  // ||dk||^2 = ||w||^2 / rho^2 (w before its update this iteration)
  S->ddnorm += ww / (rho * rho);
  const double delta = S->sn2 * rho;
  const double gambar = -S->cs2 * rho;
  const double rhs = phi - delta * S->z;
  const double zbar = rhs / gambar;
  S->xnorm = sqrt(S->xxnorm + zbar * zbar);
  const double gamma = sqrt(gambar * gambar + theta * theta);
  S->cs2 = gambar / gamma;
  S->sn2 = theta / gamma;
  S->z = rhs / gamma;
  S->xxnorm += S->z * S->z;
  S->acond = S->anorm * sqrt(S->ddnorm);
  const double res1 = S->phibar * S->phibar;
  S->res2 += psi * psi;
  S->rnorm = sqrt(res1 + S->res2);
  S->arnorm = alfa * fabs(tau);
  const double r1sq = S->rnorm * S->rnorm - S->dampsq * S->xxnorm;
  S->r1norm = sqrt(fabs(r1sq));
  S->r2norm = S->rnorm;
  const double bnorm = S->bnorm, anorm = S->anorm, xnorm = S->xnorm, rnorm = S->rnorm;
  const double test1 = rnorm / bnorm;
  const double test2 = S->arnorm / (anorm * rnorm + eps);
  const double test3 = 1.0 / (S->acond + eps);
  const double t1 = test1 / (1.0 + anorm * xnorm / bnorm);
  const double rtol = S->atol + S->btol * (anorm * xnorm / bnorm);
  S->itn += 1;
  int istop = 0;
  if (S->itn >= S->iter_lim) istop = 7;
  if (1.0 + test3 <= 1.0) istop = 6;
  if (1.0 + test2 <= 1.0) istop = 5;
  if (1.0 + t1 <= 1.0) istop = 4;
  if (test3 <= S->ctol) istop = 3;
  if (test2 <= S->atol) istop = 2;
  if (test1 <= rtol) istop = 1;
  S->pending_stop = istop;
}

__global__ void lsqr_scalar_kernel(LsqrState* S, const double* part, int nb, double* arnorms) {
  if (S->istop || threadIdx.x != 0) return;
  double vv = 0.0, ww = 0.0;
  for (int i = 0; i < nb; ++i) { vv += part[i]; ww += part[nb + i]; }
  lsqr_scalar_step(S, vv, ww);
}

// x += t1 w; var += (w / rho)^2; w = v + t2 w.  Then publish the stop code
// and record the next arnorm.
__global__ void lsqr_xw_kernel(LsqrState* S, int64_t n, double* __restrict__ x,
                               double* __restrict__ w, const double* __restrict__ v,
                               double* __restrict__ var, int calc_var) {
  if (S->istop) return;
  const double t1 = S->t1, t2 = S->t2, ir = S->inv_rho;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double wi = w[i];
    x[i] += t1 * wi;
    if (calc_var) var[i] += (wi * ir) * (wi * ir);
    w[i] = v[i] + t2 * wi;
  }
}

__global__ void lsqr_commit_kernel(LsqrState* S, double* arnorms, int cap) {
  if (S->istop) return;
  S->istop = S->pending_stop;
  if (!S->istop && S->itn < cap) arnorms[S->itn] = S->arnorm;
}

// ---- CGLS ----------------------------------------------------------------
// alpha = norm_s / (ap . ap); x += alpha p; r -= alpha ap.
__global__ void cgls_alpha_kernel(CglsState* S, const double* part, int nb) {
  __shared__ double sh[4];
  if (S->done) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) S->alpha = S->norm_s / s;
}

__global__ void cgls_xr_kernel(const CglsState* S, int64_t n, int64_t m, double* x,
                               const double* p, double* r, const double* ap) {
  if (S->done) return;
  const double a = S->alpha;
  const int64_t tot = n + m;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < n) x[i] += a * p[i];
    else r[i - n] -= a * ap[i - n];
  }
}

// norm_s_new = s . s ; stop test ; beta ; p = s + beta p.
__global__ void cgls_beta_kernel(CglsState* S, const double* part, int nb) {
  __shared__ double sh[4];
  if (S->done) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    S->iters += 1;
    if (sqrt(s) < S->tol) {
      S->done = 1;
      S->converged = 1;
    } else {
      S->beta = s / S->norm_s;
      S->norm_s = s;
      if (S->iters >= S->max_iters) S->done = 1;
    }
    S->last_norm = sqrt(s);
  }
}

__global__ void cgls_p_kernel(const CglsState* S, int64_t n, double* p, const double* s) {
  if (S->done) return;
  const double b = S->beta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = s[i] + b * p[i];
}

__global__ __launch_bounds__(TPB) void sqnorm_kernel(const double* __restrict__ x, int64_t n,
                                                     double* part) {
  __shared__ double sh[TPB / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    s += x[i] * x[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---- misc ------------------------------------------------------------------
__global__ void gather_cols_kernel(const double* A, int64_t m, int64_t lda, const int64_t* idx,
                                   double* out, int64_t ldo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < m) out[j * ldo + i] = A[idx[j] * lda + i];
}

template <typename TT>
__global__ __launch_bounds__(1024) void backsub_kernel(const TT* U, int64_t n, int64_t ldu,
                                                       const TT* y, TT* x, TT* b) {
  __shared__ TT xi;
  for (int64_t i = threadIdx.x; i < n; i += 1024) b[i] = y[i];
  __syncthreads();
  for (int64_t i = n - 1; i >= 0; --i) {
    if (threadIdx.x == 0) {
      const TT d = U[i * ldu + i];
      xi = (d != (TT)0) ? b[i] / d : (TT)0;
      x[i] = xi;
    }
    __syncthreads();
    const TT v = xi;
    for (int64_t j = threadIdx.x; j < i; j += 1024) b[j] -= U[i * ldu + j] * v;
    __syncthreads();
  }
}

__global__ void count_diag_kernel(const double* R, int64_t k, int64_t ld, double tol, int* out) {
  if (threadIdx.x != 0) return;
  int c = 0;
  for (int64_t i = 0; i < k; ++i) c += fabs(R[i * ld + i]) > tol;
  *out = c;
}

__global__ void upper_copy_kernel(const double* S, int64_t rows, int64_t cols, int64_t lds,
                                  double* D, int64_t ldd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < rows) D[j * ldd + i] = (i <= j) ? S[j * lds + i] : 0.0;
}

// 32 x 32 output tiles, block (32, 8).  The mirrored tile of A (rows of the
// tile's columns, columns of its rows) is staged through LDS so that both the
// direct and the mirrored reads are coalesced.
template <bool CHECK>
__global__ __launch_bounds__(256) void sym_tile_kernel(const double* A, int64_t n, int64_t lda,
                                                       double* G, int64_t ldg, double shift,
                                                       int* flag) {
  __shared__ double t[32][33];
  // The check compares tile (I, J) with the mirror of tile (J, I): one of
  // each pair suffices (the whole block leaves, so no barrier is skipped).
  if (CHECK && blockIdx.x > blockIdx.y) return;
  const int64_t I = (int64_t)blockIdx.x * 32, J = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int r = ty; r < 32; r += 8) {
    const int64_t row = J + tx, col = I + r;
    t[r][tx] = (row < n && col < n) ? A[col * lda + row] : 0.0;   // A(J + tx, I + r)
  }
  __syncthreads();
  bool bad = false;
  for (int c = ty; c < 32; c += 8) {
    const int64_t i = I + tx, j = J + c;
    if (i < n && j < n) {
      const double mirror = t[tx][c];                              // A(j, i)
      const double own = A[j * lda + i];                           // A(i, j)
      if (CHECK) {
        if (!(own == mirror)) bad = true;
      } else {
        double v = i >= j ? own : mirror;
        if (i == j) v += shift;
        G[j * ldg + i] = v;
      }
    }
  }
  if (CHECK && bad) *flag = 1;
}

__global__ void id_scatter_kernel(const int64_t* perm, const double* T, int64_t k, int64_t w,
                                  double* X, int64_t ldx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  for (int64_t c = blockIdx.y; c < w; c += gridDim.y) {
    const int64_t col = perm[c];
    X[col * ldx + i] = c < k ? (i == c ? 1.0 : 0.0) : T[(c - k) * k + i];
  }
}

__global__ void gather_rows_kernel(const double* A, int64_t n, int64_t lda, const int64_t* idx,
                                   int64_t k, double* out, int64_t ldo) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= k) return;
  const int64_t src = idx[r];
  for (int64_t j = blockIdx.y; j < n; j += gridDim.y) out[j * ldo + r] = A[j * lda + src];
}

__global__ void colnorm_max_kernel(const double* Y, int64_t m, int64_t b, int64_t ld,
                                   double* out) {
  __shared__ double sh[4];
  const int64_t j = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += 256) { const double v = Y[j * ld + i]; s += v * v; }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[j] = sqrt(s);
}

__global__ void diag_scale_kernel(double* X, int64_t rows, int64_t cols, int64_t ld,
                                  const double* d, int64_t stride) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < rows) X[j * ld + i] *= d[j * stride];
}

__global__ void scale_rows_kernel(double* X, int64_t rows, int64_t cols, int64_t ld,
                                  const double* d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < rows) X[j * ld + i] *= d[i];
}

// mode 0: 1/s (0 for s == 0); mode 1: 1/sqrt(s^2 + mu); mode 2: 1/s for
// s > tol else 0, and *count = number of leading entries above tol.
__global__ void inv_sigma_kernel(const double* s, int64_t n, double mu, double tol, int mode,
                                 double* out, int* count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double v = s[i];
    double r;
    if (mode == 0) r = v != 0.0 ? 1.0 / v : 0.0;
    else if (mode == 1) r = 1.0 / sqrt(v * v + mu);
    else r = v > tol ? 1.0 / v : 0.0;
    out[i] = r;
  }
  if (mode == 2 && i == 0) {
    int k = 0;
    while (k < n && s[k] > tol) ++k;
    *count = k;
  }
}

}  // namespace

hipError_t scale_rows(double* X, int64_t rows, int64_t cols, int64_t ld, const double* d,
                      hipStream_t st) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(scale_rows_kernel, dim3((unsigned)((rows + 255) / 256), (unsigned)cols),
                     dim3(256), 0, st, X, rows, cols, ld, d);
  return hipGetLastError();
}

hipError_t inv_sigma(const double* s, int64_t n, double mu, double tol, int mode, double* out,
                     int* count, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(inv_sigma_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, n,
                     mu, tol, mode, out, count);
  return hipGetLastError();
}

// ---- launchers ---------------------------------------------------------------
static int nblocks(int64_t n) {
  int64_t b = (n + TPB - 1) / TPB;
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return (int)b;
}

hipError_t lsqr_u_update(LsqrState* S, int64_t m, const double* t, double* u, double* part,
                         hipStream_t st) {
  const int nb = nblocks(m);
  hipLaunchKernelGGL(lsqr_u_kernel, dim3(nb), dim3(TPB), 0, st, S, m, t, u, part);
  hipLaunchKernelGGL(lsqr_beta_kernel, dim3(1), dim3(256), 0, st, S, part, nb);
  hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(TPB), 0, st, S, 0, m, u);
  return hipGetLastError();
}

int lsqr_fused_groups(int64_t m) {
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int64_t chunks = (m + FP_RB - 1) / FP_RB;
  return (int)std::max<int64_t>(1, std::min<int64_t>(2 * (int64_t)cus, chunks));
}

hipError_t lsqr_fused_update(LsqrState* S, const double* A, int64_t m, int64_t n, int64_t lda,
                             const double* w, double* u, double* part, double* spart,
                             double* atu, hipStream_t st) {
  if (n > FP_MAXN) return hipErrorInvalidValue;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)lsqr_fused_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)FP_LDS);
    if (e != hipSuccess) return e;
    attr = true;
  }
  const int ng = lsqr_fused_groups(m);
  const int64_t chunks = (m + FP_RB - 1) / FP_RB;
  const int64_t rows_per_wg = ((chunks + ng - 1) / ng) * FP_RB;
  hipLaunchKernelGGL(lsqr_fused_kernel, dim3(ng), dim3(FP_THR), FP_LDS, st, S, A, m, (int)n, lda,
                     w, u, part, spart, rows_per_wg);
  hipLaunchKernelGGL(lsqr_fused_finish, dim3((unsigned)((n + FF_COLS - 1) / FF_COLS)),
                     dim3(FF_COLS * FF_SL), 0, st, S, part, spart, ng, (int)n, atu);
  return hipGetLastError();
}

hipError_t lsqr_vphase(LsqrState* S, int64_t n, const double* t, double* v, double* w, double* x,
                       double* var, bool calc_var, double* arnorms, int cap, hipStream_t st) {
  if (n > 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lsqr_vphase_kernel, dim3(1), dim3(n <= 512 ? 512 : 1024), 0, st, S, (int)n, t,
                     v, w, x, var, calc_var ? 1 : 0, arnorms, cap);
  return hipGetLastError();
}

hipError_t lsqr_v_update(LsqrState* S, int64_t n, const double* t, double* v, const double* w,
                         double* part, double* arnorms, hipStream_t st) {
  const int nb = nblocks(n);
  hipLaunchKernelGGL(lsqr_v_kernel, dim3(nb), dim3(TPB), 0, st, S, n, t, v, w, part, nb);
  hipLaunchKernelGGL(lsqr_scalar_kernel, dim3(1), dim3(64), 0, st, S, part, nb, arnorms);
  hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(TPB), 0, st, S, 1, n, v);
  return hipGetLastError();
}

hipError_t lsqr_xw_update(LsqrState* S, int64_t n, double* x, double* w, const double* v,
                          double* var, bool calc_var, double* arnorms, int cap, hipStream_t st) {
  hipLaunchKernelGGL(lsqr_xw_kernel, dim3(nblocks(n)), dim3(TPB), 0, st, S, n, x, w, v, var,
                     calc_var ? 1 : 0);
  hipLaunchKernelGGL(lsqr_commit_kernel, dim3(1), dim3(1), 0, st, S, arnorms, cap);
  return hipGetLastError();
}

hipError_t cgls_alpha(CglsState* S, const double* ap, int64_t m, double* part, hipStream_t st) {
  const int nb = nblocks(m);
  hipLaunchKernelGGL(sqnorm_kernel, dim3(nb), dim3(TPB), 0, st, ap, m, part);
  hipLaunchKernelGGL(cgls_alpha_kernel, dim3(1), dim3(256), 0, st, S, part, nb);
  return hipGetLastError();
}

hipError_t cgls_xr(CglsState* S, int64_t n, int64_t m, double* x, const double* p, double* r,
                   const double* ap, hipStream_t st) {
  hipLaunchKernelGGL(cgls_xr_kernel, dim3(nblocks(n + m)), dim3(TPB), 0, st, S, n, m, x, p, r, ap);
  return hipGetLastError();
}

hipError_t cgls_beta_p(CglsState* S, int64_t n, double* p, const double* s, double* part,
                       hipStream_t st) {
  const int nb = nblocks(n);
  hipLaunchKernelGGL(sqnorm_kernel, dim3(nb), dim3(TPB), 0, st, s, n, part);
  hipLaunchKernelGGL(cgls_beta_kernel, dim3(1), dim3(256), 0, st, S, part, nb);
  hipLaunchKernelGGL(cgls_p_kernel, dim3(nb), dim3(TPB), 0, st, S, n, p, s);
  return hipGetLastError();
}

hipError_t sqnorm(const double* x, int64_t n, double* out, double* part, hipStream_t st) {
  const int nb = nblocks(n);
  hipLaunchKernelGGL(sqnorm_kernel, dim3(nb), dim3(TPB), 0, st, x, n, part);
  hipLaunchKernelGGL(finish_sum, dim3(1), dim3(256), 0, st, part, nb, out);
  return hipGetLastError();
}

hipError_t gather_cols(const double* A, int64_t m, int64_t lda, const int64_t* idx, int64_t k,
                       double* out, int64_t ldo, hipStream_t st) {
  if (m <= 0 || k <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_cols_kernel, dim3((unsigned)((m + 255) / 256), (unsigned)k), dim3(256),
                     0, st, A, m, lda, idx, out, ldo);
  return hipGetLastError();
}

hipError_t backsub(const double* U, int64_t n, int64_t ldu, const double* y, double* x,
                   double* work, hipStream_t st) {
  hipLaunchKernelGGL(backsub_kernel<double>, dim3(1), dim3(1024), 0, st, U, n, ldu, y, x, work);
  return hipGetLastError();
}
hipError_t backsub_f32(const float* U, int64_t n, int64_t ldu, const float* y, float* x,
                       float* work, hipStream_t st) {
  hipLaunchKernelGGL(backsub_kernel<float>, dim3(1), dim3(1024), 0, st, U, n, ldu, y, x, work);
  return hipGetLastError();
}

hipError_t count_diag_above(const double* R, int64_t k, int64_t ld, double tol, int* out,
                            hipStream_t st) {
  hipLaunchKernelGGL(count_diag_kernel, dim3(1), dim3(64), 0, st, R, k, ld, tol, out);
  return hipGetLastError();
}

hipError_t upper_copy(const double* S, int64_t rows, int64_t cols, int64_t lds, double* D,
                      int64_t ldd, hipStream_t st) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(upper_copy_kernel, dim3((unsigned)((rows + 255) / 256), (unsigned)cols),
                     dim3(256), 0, st, S, rows, cols, lds, D, ldd);
  return hipGetLastError();
}

hipError_t id_scatter(const int64_t* perm, const double* T, int64_t k, int64_t w, double* X,
                      int64_t ldx, hipStream_t st) {
  if (k <= 0 || w <= 0) return hipSuccess;
  hipLaunchKernelGGL(id_scatter_kernel,
                     dim3((unsigned)((k + 255) / 256), (unsigned)std::min<int64_t>(w, 65535)),
                     dim3(256), 0, st, perm, T, k, w, X, ldx);
  return hipGetLastError();
}

hipError_t gather_rows(const double* A, int64_t n, int64_t lda, const int64_t* idx, int64_t k,
                       double* out, int64_t ldo, hipStream_t st) {
  if (k <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_rows_kernel,
                     dim3((unsigned)((k + 255) / 256), (unsigned)std::min<int64_t>(n, 65535)),
                     dim3(256), 0, st, A, n, lda, idx, k, out, ldo);
  return hipGetLastError();
}

hipError_t sym_from_lower(const double* A, int64_t n, int64_t lda, double* G, int64_t ldg,
                          double shift, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned nt = (unsigned)((n + 31) / 32);
  hipLaunchKernelGGL(sym_tile_kernel<false>, dim3(nt, nt), dim3(32, 8), 0, st, A, n, lda, G, ldg,
                     shift, nullptr);
  return hipGetLastError();
}

hipError_t sym_mismatch(const double* A, int64_t n, int64_t lda, int* flag, hipStream_t st) {
  hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), st);
  if (e != hipSuccess || n <= 0) return e;
  const unsigned nt = (unsigned)((n + 31) / 32);
  hipLaunchKernelGGL(sym_tile_kernel<true>, dim3(nt, nt), dim3(32, 8), 0, st, A, n, lda, nullptr,
                     0, 0.0, flag);
  return hipGetLastError();
}

hipError_t column_norms(const double* Y, int64_t m, int64_t b, int64_t ld, double* out,
                        hipStream_t st) {
  if (b <= 0) return hipSuccess;
  hipLaunchKernelGGL(colnorm_max_kernel, dim3((unsigned)b), dim3(256), 0, st, Y, m, b, ld, out);
  return hipGetLastError();
}

hipError_t scale_columns(double* X, int64_t rows, int64_t cols, int64_t ld, const double* d,
                         int64_t stride, hipStream_t st) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(diag_scale_kernel, dim3((unsigned)((rows + 255) / 256), (unsigned)cols),
                     dim3(256), 0, st, X, rows, cols, ld, d, stride);
  return hipGetLastError();
}

namespace {
// sg[i] = -1 where W[i][i] < 0, else +1 (i < k).
__global__ void diag_sign_kernel(const double* W, int64_t k, int64_t ld, double* sg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) sg[i] = W[i + i * ld] < 0.0 ? -1.0 : 1.0;
}
// R (k x n) = diag(sg) * upper(W[:k, :]) (zeros below the diagonal).
__global__ void upper_scaled_kernel(const double* W, int64_t k, int64_t n, int64_t ldw,
                                    const double* sg, double* R, int64_t ldr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= k || j >= n) return;
  R[i + j * ldr] = i <= j ? sg[i] * W[i + j * ldw] : 0.0;
}
}  // namespace

hipError_t diag_signs(const double* W, int64_t k, int64_t ld, double* sg, hipStream_t st) {
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(diag_sign_kernel, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, st, W, k,
                     ld, sg);
  return hipGetLastError();
}

hipError_t upper_scaled(const double* W, int64_t k, int64_t n, int64_t ldw, const double* sg,
                        double* R, int64_t ldr, hipStream_t st) {
  if (k <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(upper_scaled_kernel, dim3((unsigned)((k + 255) / 256), (unsigned)n),
                     dim3(256), 0, st, W, k, n, ldw, sg, R, ldr);
  return hipGetLastError();
}

}  // namespace rna
