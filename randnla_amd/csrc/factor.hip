// Small dense factorizations (FP64) and elementwise helpers.
//
// The factorizations work on the l x l (or d x n) matrices of the path:
// Gram matrices of the tall panels, the R factors, the sketch of A.  They are
// latency-bound; each runs as one work-group (Cholesky, triangular inverse,
// SVD finalisation) or as a short sequence of launches (Jacobi rounds,
// Householder steps).
#include <algorithm>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace rna {
namespace {

constexpr int BIG = 1024;

// ---------------------------------------------------------------- Cholesky
// Right-looking, strips of NB rows staged in LDS.  G is column-major; only
// its upper triangle is read.  On exit the upper triangle holds R (G = R^T R)
// and the strictly lower triangle is zero.
constexpr int CH_NB = 16;
constexpr int64_t CH_OUTER = 512;   // rows per panel of the blocked factorisation
constexpr int CH_MAXN = 1024;

__global__ __launch_bounds__(BIG) void cholesky_kernel(double* G, int n, int64_t ld, double tau,
                                                       int* flags, int* ndef) {
  __shared__ double S[CH_NB][CH_MAXN];
  __shared__ double d0[CH_MAXN];
  __shared__ double rjj;
  __shared__ int cnt;
  const int t = threadIdx.x;
  for (int j = t; j < n; j += BIG) { d0[j] = G[(int64_t)j * ld + j]; flags[j] = 0; }
  if (t == 0) cnt = 0;
  __syncthreads();
  for (int J = 0; J < n; J += CH_NB) {
    const int nb = min(CH_NB, n - J), w = n - J;
    for (int idx = t; idx < nb * w; idx += BIG) {
      const int p = idx % nb, c = idx / nb;           // row J+p, column J+c
      S[p][c] = (c >= p) ? G[(int64_t)(J + c) * ld + (J + p)] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
      if (t == 0) {
        const double d = S[j][j];
        if (d > tau * d0[J + j] && d > 0.0) {
          rjj = sqrt(d);
        } else {
          rjj = 0.0;
          flags[J + j] = 1;
          ++cnt;
        }
      }
      __syncthreads();
      const double r = rjj;
      const double inv = (r > 0.0) ? 1.0 / r : 0.0;
      for (int c = j + t; c < w; c += BIG) S[j][c] = (c == j) ? r : S[j][c] * inv;
      __syncthreads();
      // Update the remaining rows of the strip.
      for (int idx = t; idx < (nb - j - 1) * w; idx += BIG) {
        const int p = j + 1 + idx / w, c = idx % w;
        if (c >= p) S[p][c] -= S[j][p] * S[j][c];
      }
      __syncthreads();
    }
    // Write the strip back (upper part of R rows J..J+nb-1).
    for (int idx = t; idx < nb * w; idx += BIG) {
      const int p = idx % nb, c = idx / nb;
      if (c >= p) G[(int64_t)(J + c) * ld + (J + p)] = S[p][c];
    }
    // Trailing update of rows/cols >= J+nb (upper triangle).
    const int r0 = nb, rem = w - nb;
    const int wv = t >> 6, ln = t & 63;
    for (int cc = wv; cc < rem; cc += BIG / 64) {
      const int c = r0 + cc;
      for (int pp = ln; pp <= cc; pp += 64) {
        const int p = r0 + pp;
        double s = 0.0;
#pragma unroll 4
        for (int q = 0; q < nb; ++q) s += S[q][p] * S[q][c];
        G[(int64_t)(J + c) * ld + (J + p)] -= s;
      }
    }
    __syncthreads();
  }
  // Zero the strictly lower triangle.
  for (int64_t idx = t; idx < (int64_t)n * n; idx += BIG) {
    const int i = (int)(idx % n), c = (int)(idx / n);
    if (i > c) G[(int64_t)c * ld + i] = 0.0;
  }
  if (t == 0) *ndef = cnt;
}

// Blocked right-looking Cholesky, step J: factor the nb x nb diagonal block
// (nb <= 64) in LDS with the dependent-column rule (one wave), write R_JJ
// (zeroing the block's lower part) and R_JJ^{-1} (zero rows for dependent
// columns).  The panel and trailing updates run on the FP64 GEMM.
constexpr int PB = 64;

__global__ void chol_init_kernel(const double* G, int n, int64_t ld, double* d0, int* flags,
                                 int* ndef) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) { d0[j] = G[(int64_t)j * ld + j]; flags[j] = 0; }
  if (j == 0) *ndef = 0;
}

// Broadcast lane `lane`'s value (uniform lane index).
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// X = R^{-1} for an upper-triangular block held one column per lane (lane c:
// col[p] = R[p][c] for p <= c): lane c back-substitutes R x = e_c.  A zero
// diagonal entry gives a zero row and column.
__device__ __forceinline__ void invert_cols_wave(const double (&col)[PB], double (&x)[PB], int nb) {
  const int c = threadIdx.x;
#pragma unroll
  for (int p = PB - 1; p >= 0; --p) {
    double s = (p == c) ? 1.0 : 0.0;
#pragma unroll
    for (int k = p + 1; k < PB; ++k) s -= readlane_f64(col[p], k) * x[k];
    const double rpp = readlane_f64(col[p], p);
    x[p] = (rpp != 0.0 && p < nb && c < nb) ? s / rpp : 0.0;
  }
}

// Cholesky of the nb x nb diagonal block at (J, J) with the dependent-column
// rule, and its inverse into Rinv (64 x 64, zero padded).  One wave; lane c
// keeps column c in registers (loops fully unrolled, so the register index is
// static) and row j of R reaches the other lanes through v_readlane.  Entries
// below the diagonal of a lane's column become scratch and are never read.
__global__ __launch_bounds__(64) void potrf_wave_kernel(double* G, int64_t ld, int J, int nb,
                                                        double tau, const double* d0, int* flags,
                                                        int* ndef, double* Rinv) {
  const int c = threadIdx.x;
  double col[PB];
#pragma unroll
  for (int p = 0; p < PB; ++p)
    col[p] = (c < nb && p <= c) ? G[(int64_t)(J + c) * ld + (J + p)] : 0.0;
  int nd = 0;
#pragma unroll
  for (int j = 0; j < PB; ++j) {
    if (j < nb) {
      const double d = readlane_f64(col[j], j);
      const bool ok = d > tau * d0[J + j] && d > 0.0;
      const double r = ok ? sqrt(d) : 0.0, inv = ok ? 1.0 / r : 0.0;
      if (!ok) {
        ++nd;
        if (c == 0) flags[J + j] = 1;
      }
      col[j] = (c == j) ? r : col[j] * inv;
#pragma unroll
      for (int p = j + 1; p < PB; ++p) col[p] -= readlane_f64(col[j], p) * col[j];
    }
  }
  if (c == 0 && nd) atomicAdd(ndef, nd);
  if (c < nb) {
#pragma unroll
    for (int p = 0; p < PB; ++p)
      if (p < nb) G[(int64_t)(J + c) * ld + (J + p)] = (p <= c) ? col[p] : 0.0;
  }
  double x[PB];
  invert_cols_wave(col, x, nb);
#pragma unroll
  for (int p = 0; p < PB; ++p) Rinv[(int64_t)c * PB + p] = x[p];
}

// Inverse of each 64 x 64 diagonal block of R into the same block of X (one
// wave per block, columns in registers).
__global__ __launch_bounds__(64) void diag_inv_wave_kernel(const double* R, int n, int64_t ldr,
                                                           double* X, int64_t ldx) {
  const int J = blockIdx.x * PB, nb = min(PB, n - J), c = threadIdx.x;
  double col[PB], x[PB];
#pragma unroll
  for (int p = 0; p < PB; ++p)
    col[p] = (c < nb && p <= c) ? R[(int64_t)(J + c) * ldr + (J + p)] : 0.0;
  invert_cols_wave(col, x, nb);
  if (c < nb) {
#pragma unroll
    for (int p = 0; p < PB; ++p)
      if (p < nb) X[(int64_t)(J + c) * ldx + (J + p)] = x[p];
  }
}

// 1/x and 1/sqrt(x) from the hardware estimates refined by two Newton steps
// (each step doubles the correct bits), for normal, finite positive x.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double rsq_nr(double x) {
  const double h = 0.5 * x;
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

// ------------------------------------------- fused Cholesky + inverse (n <= 512)
// One work-group factors G = R^T R (upper R, right-looking over 32-row block
// rows, dependent-column rule as above) and writes the inverses of the
// diagonal blocks into X; trinv_cols_kernel completes X = R^{-1}.  LDS holds
// the current 32-row strip, the 32 x 32 inverse of the diagonal block and the
// original diagonal.  Per block row:
//   wave 0:   Cholesky of the 32 x 32 diagonal block, columns in registers,
//             rows broadcast by readlane; its inverse by back substitution;
//   all:      panel R[J, >J] = R_JJ^{-T} G[J, >J] and trailing update
//             G[>J, >J] -= R[J, >J]^T R[J, >J] (upper tiles) on FP64 MFMA.
// Wave 0 updates the next diagonal block and factors it while the other
// waves update the rest.  Each wave loads its tiles of G before the panel
// (their latency overlaps it); the updated tiles of the next block row stay
// in registers and are written into LDS as the next strip, the others go
// back to G.
constexpr int CB = 32;
constexpr int CI_MAXN = 512;
constexpr int CI_SLD = CI_MAXN + 2;   // strip row stride (doubles)
constexpr size_t CI_LDS = ((size_t)CB * CI_SLD + 2 * CB * (CB + 1) + CI_MAXN) * sizeof(double);
constexpr int CI_THR = 1024;
constexpr int CI_NW = CI_THR / 64;

// Wave 0's part of a block row of chol_inv_kernel: Cholesky of the nb x nb
// diagonal block at (j0, j0) with the dependent-column rule, read from src
// (column-major, leading dimension sld; upper triangle), and its inverse.
// Lane c < CB holds column c of the block in registers; step j broadcasts
// R_jj and row j of R with v_readlane (no LDS round trip on the chain).
// Lanes CB + cc start as column cc of the identity and go through the same
// row operations (scale row j by 1 / R_jj, subtract R[j][p] times row j from
// row p > j): that is forward elimination with L = R^T, so they end as
// column cc of L^{-1} = R^{-T}, i.e. row cc of X_JJ = R_JJ^{-1}, with no
// separate substitution pass.  R_JJ goes to G (zero below the diagonal),
// X_JJ to DI and X; a dependent column has a zero row in R and a zero row and
// column in X.  src may be LDS that the caller reuses: it is read first.
// Not inlined: it keeps its own register allocation.
template <bool FULL>
__device__ __attribute__((noinline)) void chol_diag_block(const double* src, int64_t sld,
                                                          double* G, int64_t ld, int j0, int nb,
                                                          double tau, const double* d0,
                                                          int* flags, int* cnt, double* DI,
                                                          double* X, int64_t ldx) {
  // Arguments arrive in VGPRs: make the uniform ones scalar, so that the
  // per-step tests on them are scalar branches, not lane masks.
  nb = __builtin_amdgcn_readfirstlane(nb);
  j0 = __builtin_amdgcn_readfirstlane(j0);
  tau = readlane_f64(tau, 0);
  const int c = threadIdx.x & 63;
  const bool xl = c >= CB;
  const int cc = c - CB;
  double col[CB];
#pragma unroll
  for (int p = 0; p < CB; ++p)
    col[p] = xl ? (p == cc ? 1.0 : 0.0)
                : ((c < nb && p <= c) ? src[(int64_t)c * sld + p] : 0.0);
  const double dz = c < nb ? d0[j0 + c] : 1.0;   // lane j: original diagonal of row j
  uint32_t dep = 0;
#pragma unroll
  for (int j = 0; j < CB; ++j) {
    if (FULL || j < nb) {
      const double d = readlane_f64(col[j], j);
      const double dref = readlane_f64(dz, j);
      const bool ok = d > tau * dref && d > 0.0;
      // Computed for any d and then selected: no branch splits the steps,
      // so the next step's chain can be scheduled among this step's updates.
      const double r = rsq_nr(d);
      const double inv = ok ? r : 0.0;
      dep |= ok ? 0u : (1u << j);
      col[j] = (c == j) ? d * inv : col[j] * inv;    // R[j][c] for c >= j; row j of L^{-1}
#pragma unroll
      for (int p = j + 1; p < CB; ++p) col[p] = fma(-readlane_f64(col[j], p), col[j], col[p]);
    }
  }
  if (c == 0 && dep) {
    *cnt += __builtin_popcount(dep);
    for (int j = 0; j < CB; ++j)
      if (dep & (1u << j)) flags[j0 + j] = 1;
  }
  if (c < nb) {
#pragma unroll
    for (int p = 0; p < CB; ++p)
      if (FULL || p < nb) G[(int64_t)(j0 + c) * ld + j0 + p] = (p <= c) ? col[p] : 0.0;
  }
  if (xl) {
    // col[p] = L^{-1}[p][cc] = X[cc][p] (zero for p < cc).
#pragma unroll
    for (int p = 0; p < CB; ++p) DI[cc * (CB + 1) + p] = col[p];
    if (cc < nb) {
#pragma unroll
      for (int p = 0; p < CB; ++p)
        if (FULL || p < nb) X[(int64_t)(j0 + p) * ldx + j0 + cc] = col[p];
    }
  }
}

// The diagonal block of chol_inv_kernel (full or partial).
__device__ __forceinline__ void chol_diag(const double* src, int64_t sld, double* G, int64_t ld,
                                          int j0, int nb, double tau, const double* d0,
                                          int* flags, int* cnt, double* DI, double* X,
                                          int64_t ldx) {
  // A full block has no per-step test on nb (one basic block, freely scheduled).
  if (nb == CB)
    chol_diag_block<true>(src, sld, G, ld, j0, nb, tau, d0, flags, cnt, DI, X, ldx);
  else
    chol_diag_block<false>(src, sld, G, ld, j0, nb, tau, d0, flags, cnt, DI, X, ldx);
}

// Zeros below the diagonal block of block column J in G and X.
__device__ __forceinline__ void chol_zero_below(double* G, int n, int64_t ld, double* X,
                                                int64_t ldx, int J) {
  const int t = threadIdx.x;
  const int j0 = J * CB, nb = min(CB, n - j0), c0 = j0 + nb, w2 = n - c0;
  for (int idx = t; idx < nb * w2; idx += CI_THR) {
    const int cj = idx / w2, r = c0 + idx % w2;
    G[(int64_t)(j0 + cj) * ld + r] = 0.0;
    X[(int64_t)(j0 + cj) * ldx + r] = 0.0;
  }
}

// Strip of block row J of G into S (rows >= nb zero) and zeros below the
// diagonal block of block column J in G and X (chol_inv_kernel).
__device__ __attribute__((noinline)) void chol_load_strip(double* G, int n, int64_t ld, double* X,
                                                          int64_t ldx, double* S, int J) {
  const int t = threadIdx.x;
  const int j0 = J * CB, nb = min(CB, n - j0), c0 = j0 + nb, w2 = n - c0;
  for (int idx = t; idx < CB * w2; idx += CI_THR) {
    const int p = idx % CB, cc = idx / CB;
    S[p * CI_SLD + cc] = p < nb ? G[(int64_t)(c0 + cc) * ld + j0 + p] : 0.0;
  }
  chol_zero_below(G, n, ld, X, ldx, J);
}

// The q-th upper tile (a, b) of a trailing update with T tile columns, in
// row-major order without the tiles of the next diagonal block ((0,0), (0,1),
// (1,1) when T >= 2, else (0,0)).
__device__ __forceinline__ void chol_tile_ab(int q, int T, int& a, int& b) {
  a = 0;
  if (T >= 2) {
    if (q < T - 2) { b = q + 2; return; }
    q -= T - 2;
    if (q < T - 2) { a = 1; b = q + 2; return; }
    q -= T - 2;
    a = 2;
  } else {
    a = 1;
  }
  while (q >= T - a) { q -= T - a; ++a; }
  b = a + q;
}

// Tile layout of chol_inv_kernel's 16 x 16 FP64 MFMA results: the product
// is formed transposed, so that lane ln's value v is element
// (row 16a + ln % 16, column 16b + 4v + ln / 16) of tile (a, b), and one
// load or store instruction covers 16 consecutive rows (one cache line of
// the column-major G) in each of 4 columns.
__device__ __forceinline__ int ct_row(int a, int ln) { return a * 16 + (ln & 15); }
__device__ __forceinline__ int ct_col(int b, int v, int ln) { return b * 16 + 4 * v + (ln >> 4); }

// G values of upper tile (a, b) of the trailing matrix at c0, in the tile
// layout.  The loads are unconditional so that they issue back to back: an
// element outside w2 (or of an invalid tile) reads G[0], and its value is
// never used (every store of a tile element is masked the same way).
__device__ __forceinline__ void chol_tile_load(const double* G, int64_t ld, int c0, int w2, int a,
                                               int b, bool valid, int ln, double (&g)[4]) {
  const int r = ct_row(a, ln);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int cl = ct_col(b, v, ln);
    const bool ok = valid && r < w2 && cl < w2;
    g[v] = G[ok ? (int64_t)(c0 + cl) * ld + c0 + r : 0];
  }
}

// (P^T P) tile (a, b) of the trailing update in the tile layout, P the
// 32 x w2 panel in S: the MFMA forms P_b^T P_a (operand A from the columns
// of tile column b, B from those of tile row a).  The LDS reads are
// unconditional (columns below 16 T < CI_SLD exist in S) and masked
// afterwards.
__device__ __forceinline__ f64x4 chol_tile_mfma(const double* S, int w2, int a, int b, int ln) {
  const int bc = b * 16 + (ln & 15), ar = a * 16 + (ln & 15);
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < CB / 4; ++kk) {
    const int k = 4 * kk + (ln >> 4);
    const double x = S[k * CI_SLD + bc], y = S[k * CI_SLD + ar];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(bc < w2 ? x : 0.0, ar < w2 ? y : 0.0, acc, 0, 0,
                                               0);
  }
  return acc;
}

// Work-group barriers: LDS traffic only, and LDS plus global memory (the
// waves' global stores complete before any wave passes).
__device__ __forceinline__ void ci_lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}
__device__ __forceinline__ void ci_mem_barrier() {
  asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// The trailing update's tiles go to the waves w > 0: the wave's tiles are
// q = w - 1 + NWO i (NWO = CI_THR / 64 - 1 waves).  The first CI_PF0 are
// loaded before the panel; every tile of the next strip's rows (a = 0, 1) is
// among them, so the strip is built from registers.  The others are loaded
// one tile ahead, before the current tile's stores (the wait for a load then
// does not include earlier stores).
constexpr int CI_PF0 = 4;
constexpr int CI_NWO = CI_THR / 64 - 1;
static_assert(2 * ((CI_MAXN - CB) / 16 - 2) <= CI_NWO * CI_PF0,
              "the next strip's tiles must be in the first group");

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Updates tile (a, b) held in g; rows a >= 2 also go back to G, through the
// buffer resource grs over G: an element outside w2 gets an offset past the
// resource's end and the hardware drops it, so the stores need no branch
// (a branch would hold the wait for g's loads inside it, and the wait
// analysis would then assume the loads still pending after it).
__device__ __forceinline__ void chol_tile_finish(__amdgpu_buffer_rsrc_t grs, int64_t ld,
                                                 const double* S, int c0, int w2, int a, int b,
                                                 int ln, double (&g)[4]) {
  const f64x4 acc = chol_tile_mfma(S, w2, a, b, ln);
#pragma unroll
  for (int v = 0; v < 4; ++v) g[v] -= acc[v];
  if (a >= 2) {
    const int r = ct_row(a, ln);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int cl = ct_col(b, v, ln);
      const bool ok = r < w2 && cl < w2;
      const int boff = (int)(((int64_t)(c0 + cl) * ld + c0 + r) * 8);
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, g[v]), grs,
                                            ok ? boff : (int)0x80000000, 0, 0);
    }
  }
}

__global__ __launch_bounds__(CI_THR) void chol_inv_kernel(double* G, int n, int64_t ld, double tau,
                                                       int* flags, int* ndef, double* X,
                                                       int64_t ldx) {
  extern __shared__ double shm[];
  double* S = shm;                             // strip / panel [CB][CI_SLD]
  double* DI = S + (size_t)CB * CI_SLD;        // [CB][CB + 1]: inverse of the diagonal block
  double* RJ = DI + CB * (CB + 1);             // [CB][CB + 1]: next diagonal block / R_JJ
  double* d0 = RJ + CB * (CB + 1);             // original diagonal
  __shared__ int cnt;
  const int t = threadIdx.x, wv0 = __builtin_amdgcn_readfirstlane(t >> 6), ln0 = t & 63;
  // G as a buffer resource (n <= CI_MAXN: byte offsets fit in 31 bits).
  const __amdgpu_buffer_rsrc_t grs = __builtin_amdgcn_make_buffer_rsrc(
      (void*)G, (short)0, (int)(((int64_t)(n - 1) * ld + n) * 8), 0x00020000);
  const int nt = (n + CB - 1) / CB;
  for (int j = t; j < n; j += CI_THR) { d0[j] = G[(int64_t)j * ld + j]; flags[j] = 0; }
  if (t == 0) cnt = 0;
  __syncthreads();
  if (wv0 == 0) chol_diag(G, ld, G, ld, 0, min(CB, n), tau, d0, flags, &cnt, DI, X, ldx);
  __syncthreads();
  chol_load_strip(G, n, ld, X, ldx, S, 0);
  __syncthreads();

  for (int J = 0; J < nt - 1; ++J) {
    const int j0 = J * CB, c0 = j0 + CB, w2 = n - c0;   // J < nt - 1: full block
    const int T = (w2 + 15) / 16;
    const int nd = T >= 2 ? 3 : 1;                    // tiles of the next diagonal block
    const int nq = T * (T + 1) / 2 - nd;              // the other upper tiles
    // Lane and wave index made opaque per block row: addresses derived from
    // them are not hoisted out of the loop (they would hold registers across
    // all of it).
    int ln = ln0, wv = wv0;
    asm volatile("" : "+v"(ln));
    asm volatile("" : "+s"(wv));
    // ---- the wave's first tiles of G, loaded now so that their latency
    // overlaps the panel: wave 0 the next diagonal block's tiles ((0,0),
    // (0,1), (1,1)), wave w > 0 its group 0.
    // The loads sit outside the wave-dependent branches, which only choose
    // the tile: loads in both arms would land in different registers and
    // the copies at the join would wait for them.
    double gv[CI_PF0][4];
    const int q0 = wv - 1;
#pragma unroll
    for (int i = 0; i < CI_PF0; ++i) {
      int a = 0, b = 0;
      bool valid;
      if (wv == 0) {
        valid = i < nd;
        a = i == 2 ? 1 : 0;
        b = i == 0 ? 0 : 1;
      } else {
        const int q = q0 + CI_NWO * i;
        valid = q < nq;
        if (valid) chol_tile_ab(q, T, a, b);
      }
      chol_tile_load(G, ld, c0, w2, a, b, valid, ln, gv[i]);
    }
    // ---- panel P = X_JJ^T S, in registers, then into S and G
    {
      constexpr int PQ = (2 * ((CI_MAXN - CB + 15) / 16) + CI_NW - 1) / CI_NW;   // tiles per wave
      f64x4 acc[PQ];
      int cntt = 0;
#pragma unroll
      for (int q = 0; q < PQ; ++q) {
        const int tile = wv + CI_NW * q;              // 2T tiles: (ti, tj)
        acc[q] = f64x4{0.0, 0.0, 0.0, 0.0};
        if (tile < 2 * T) {
          // P tile (ti, tj) in the tile layout: the MFMA forms
          // S_tj^T X_JJ(:, ti).
          const int ti = tile & 1, tj = tile >> 1;
          const int row = ti * 16 + (ln & 15), col = tj * 16 + (ln & 15);
#pragma unroll
          for (int kk = 0; kk < CB / 4; ++kk) {
            const int k = 4 * kk + (ln >> 4);
            const double y = S[k * CI_SLD + col];     // col < 16 T < CI_SLD
            acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(col < w2 ? y : 0.0,
                                                          DI[k * (CB + 1) + row], acc[q], 0, 0, 0);
          }
          cntt = q + 1;
        }
      }
      ci_lds_barrier();
#pragma unroll
      for (int q = 0; q < PQ; ++q) {
        const int tile = wv + CI_NW * q;
        if (q < cntt) {
          const int ti = tile & 1, tj = tile >> 1;
          const int row = ct_row(ti, ln);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int col = ct_col(tj, v, ln);
            if (col < w2) {
              S[row * CI_SLD + col] = acc[q][v];
              G[(int64_t)(c0 + col) * ld + j0 + row] = acc[q][v];
            }
          }
        }
      }
    }
    // The panel's stores to G are not read again in this kernel.
    ci_lds_barrier();
    // ---- trailing update G[>J, >J] -= P^T P on the upper tiles (a <= b).
    if (wv == 0) {
      // The next diagonal block, staged in RJ (its G copy is overwritten by
      // R_JJ) and factored right away.
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        if (i < nd) {
          const int ta = i == 2 ? 1 : 0, tb = i == 0 ? 0 : 1;
          const f64x4 acc = chol_tile_mfma(S, w2, ta, tb, ln);
          const int r = ct_row(ta, ln);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int cl = ct_col(tb, v, ln);
            if (cl < w2 && r < w2) RJ[cl * (CB + 1) + r] = gv[i][v] - acc[v];
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      // The factorisation is the critical path: its instructions go first
      // on its SIMD, ahead of the trailing-update waves there.
      __builtin_amdgcn_s_setprio(3);
      chol_diag(RJ, CB + 1, G, ld, c0, min(CB, w2), tau, d0, flags, &cnt, DI, X, ldx);
      __builtin_amdgcn_s_setprio(0);
      // Wave 0's global stores (R_JJ, X_JJ, flags) are not read in this
      // kernel: it needs no wait for them.
      ci_lds_barrier();
    } else {
      // Groups g >= 1 load while the previous group updates.  Rows a >= 2
      // go back to G (the next blocks load them); group 0's rows 0 and 1
      // stay in registers and become the next strip.
      // The tiles after group 0 come two at a time through two buffers, each
      // reloaded right after its tile is done: two tiles' loads are in
      // flight while one updates (a copy between buffers would wait).
      const auto load_tile = [&](int qq, double (&g)[4]) {
        int a = 0, b = 0;
        if (qq < nq) chol_tile_ab(qq, T, a, b);
        chol_tile_load(G, ld, c0, w2, a, b, qq < nq, ln, g);
      };
      const auto finish_tile = [&](int qq, double (&g)[4]) {
        if (qq < nq) {
          int a, b;
          chol_tile_ab(qq, T, a, b);                   // a >= 2 after group 0
          chol_tile_finish(grs, ld, S, c0, w2, a, b, ln, g);
        }
      };
      int q = q0 + CI_NWO * CI_PF0;                    // first tile after group 0
      double nx[4], ny[4];
      load_tile(q, nx);
      load_tile(q + CI_NWO, ny);
#pragma unroll
      for (int i = 0; i < CI_PF0; ++i) finish_tile(q0 + CI_NWO * i, gv[i]);
      for (; q < nq; q += 2 * CI_NWO) {
        finish_tile(q, nx);
        load_tile(q + 2 * CI_NWO, nx);
        finish_tile(q + CI_NWO, ny);
        load_tile(q + 3 * CI_NWO, ny);
      }
      ci_mem_barrier();
      // The strip of block row J + 1 (rows c0 .. c0 + 31, columns from
      // c0 + CB) from the tiles (0, b) and (1, b), b >= 2.  Such tiles exist
      // only when w2 > 32, i.e. when block J + 1 is full.
#pragma unroll
      for (int i = 0; i < CI_PF0; ++i) {
        const int q = q0 + CI_NWO * i;
        if (q < nq) {
          int ta, tb;
          chol_tile_ab(q, T, ta, tb);
          if (ta < 2) {
            const int r = ct_row(ta, ln);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int cl = ct_col(tb, v, ln);
              if (cl < w2) S[r * CI_SLD + cl - CB] = gv[i][v];
            }
          }
        }
      }
    }
    chol_zero_below(G, n, ld, X, ldx, J + 1);
    ci_lds_barrier();
  }
  if (t == 0) *ndef = cnt;
}

// X = R^{-1} above the diagonal blocks, one work-group per 32-column block
// J >= 1 of X (independent): the block column X[:, J] is built in LDS from the
// bottom, X[I, J] = -X_II sum_{I < K <= J} R[I, K] X[K, J], with the
// diagonal blocks X_KK already in X (written by chol_inv_kernel).  Per block
// row I, wave w computes the 16-row half w & 1 of T = R[I, >I] X[>I, J] over
// the K slice w >> 1 (both 16-column halves, so each R value feeds two
// MFMAs); the eight slice products go to eight LDS buffers and are summed in
// slice order by all threads (n <= 320), or slices 0-3 and 4-7 are summed in
// order into two buffers (n <= 512, less LDS); X[I, J] = -X_II T on four
// waves.  R[I - 1, >I - 1] and X_{I-1,I-1} are loaded during block row I,
// into the other of two register sets: the loop's barriers wait on LDS only,
// so those loads stay in flight across them.
constexpr int XI_NS = 8;   // K slices
// MS: 4-row K steps per slice, at least ceil((n - CB) / (4 XI_NS)) (every
// step is a load per lane; the instantiations cover n <= 160, 320, 512,
// i.e. n <= CB (MS + 1)).  The block column takes CB (MS + 1) + 1 rows per
// column (odd stride), NB slice buffers of CB x (CB + 1) and X_II one more.
template <int MS>
struct XiCfg {
  static constexpr int LD = CB * (MS + 1) + 1;
  static constexpr size_t lds(int nb) {
    return ((size_t)CB * LD + (size_t)(nb + 1) * CB * (CB + 1)) * sizeof(double);
  }
  static constexpr int NB = lds(XI_NS) <= 160 * 1024 ? XI_NS : 2;
  static constexpr size_t LDS = lds(NB);
};
static_assert(XiCfg<15>::LD >= CI_MAXN + 1 && XiCfg<15>::LDS <= 160 * 1024, "LDS");
template <int MS>
__global__ __launch_bounds__(BIG) void trinv_cols_kernel(const double* R, int n, int64_t ldr,
                                                         double* X, int64_t ldx) {
  constexpr int XI_LD = XiCfg<MS>::LD, NB = XiCfg<MS>::NB, RND = XI_NS / NB;
  static_assert(BIG == CB * CB, "one element of T per thread");
  extern __shared__ double shm[];
  double* XC = shm;                          // X[0 : j0 + 32, J] as [col][row], ld XI_LD
  double* T0 = XC + (size_t)CB * XI_LD;      // NB slice buffers, [row][col], ld CB + 1
  double* XD = T0 + NB * CB * (CB + 1);      // X_II as [col][row]
  const int t = threadIdx.x, wv = __builtin_amdgcn_readfirstlane(t >> 6), ln = t & 63;
  const int J = blockIdx.x + 1, j0 = J * CB, nb = min(CB, n - j0), je = j0 + nb;
  // R as a buffer resource (byte offsets fit in 31 bits, checked by the
  // caller): a load past the resource's end returns 0, so the slice ends need
  // no branch.
  const __amdgpu_buffer_rsrc_t rrs = __builtin_amdgcn_make_buffer_rsrc(
      (void*)R, (short)0, (int)(((int64_t)(n - 1) * ldr + n) * 8), 0x00020000);
  const int ti = wv & 1, sl = wv >> 1, kq = ln >> 4, lc = ln & 15;
  const int row = ti * 16 + lc;
  // slice [kb, ke) of the K range of block row I (kw = je - (I + 1) * CB rows)
  auto slice = [&](int I, int& kb, int& ke) {
    const int kw = je - (I + 1) * CB;
    const int ksl = (kw + 4 * XI_NS - 1) / (4 * XI_NS) * 4;
    kb = sl * ksl;
    ke = min(kw, kb + ksl);
  };
  auto load = [&](int I, double (&a)[MS], double& xd) {
    const int i0 = I * CB, k0 = i0 + CB;
    int kb, ke;
    slice(I, kb, ke);
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      const int k = kb + 4 * s + kq;
      const int off = (int)(((int64_t)(k0 + k) * ldr + i0 + row) * 8);
      a[s] = __builtin_bit_cast(
          double, __builtin_amdgcn_raw_buffer_load_b64(rrs, k < ke ? off : (int)0x80000000, 0, 0));
    }
    xd = X[(int64_t)(i0 + (t >> 5)) * ldx + i0 + (t & 31)];
  };
  auto step = [&](int I, const double (&a)[MS], double xd) {
    const int i0 = I * CB, k0 = i0 + CB, kl = je - k0 - 1;
    int kb, ke;
    slice(I, kb, ke);
    XD[(t >> 5) * (CB + 1) + (t & 31)] = xd;
    f64x4 c0 = {0.0, 0.0, 0.0, 0.0}, c1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      if (kb + 4 * s < ke) {
        // rows past the slice hold a = 0; clamped into the written rows of XC
        const int k = k0 + min(kb + 4 * s + kq, kl);
        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], XC[lc * XI_LD + k], c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], XC[(16 + lc) * XI_LD + k], c1, 0, 0, 0);
      }
    }
    // slice sl into buffer sl / RND, in round sl % RND
    double* Tb = T0 + (sl / RND) * CB * (CB + 1);
#pragma unroll
    for (int s = 0; s < RND; ++s) {
      if (sl % RND == s) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double* d = &Tb[(ti * 16 + 4 * v + kq) * (CB + 1) + lc];
          d[0] = (s == 0 ? 0.0 : d[0]) + c0[v];
          d[16] = (s == 0 ? 0.0 : d[16]) + c1[v];
        }
      }
      ci_lds_barrier();
    }
    if (NB > 2) {                              // T = sum of the buffers, in order
      double* d = &T0[(t >> 5) * (CB + 1) + (t & 31)];
      double v = d[0];
#pragma unroll
      for (int b = 1; b < NB; ++b) v += d[b * CB * (CB + 1)];
      d[0] = v;
      ci_lds_barrier();
    }
    // XC[I rows] = -X_II T: wave q < 4 takes one 16 x 16 tile
    if (wv < 4) {
      const int col = (wv >> 1) * 16 + lc;
      // two accumulators (even and odd K steps) halve the dependent chain
      f64x4 o[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
      for (int kk = 0; kk < CB / 4; ++kk) {
        const int k = 4 * kk + kq;
        o[kk & 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(
            XD[k * (CB + 1) + row],
            NB > 2 ? T0[k * (CB + 1) + col]
                   : T0[k * (CB + 1) + col] + T0[(CB + k) * (CB + 1) + col],
            o[kk & 1], 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
        XC[col * XI_LD + i0 + ti * 16 + 4 * v + kq] = -(o[0][v] + o[1][v]);
    }
    ci_lds_barrier();
  };
  // X_JJ into the block column
  for (int idx = t; idx < CB * CB; idx += BIG) {
    const int p = idx % CB, c = idx / CB;
    XC[c * XI_LD + j0 + p] = (p < nb && c < nb) ? X[(int64_t)(j0 + c) * ldx + j0 + p] : 0.0;
  }
  double a0[MS], a1[MS], x0, x1;
  load(J - 1, a0, x0);
  __syncthreads();
  // two block rows per trip, alternating the register sets (a copy between
  // them would wait for the loads)
  for (int I = J - 1;;) {
    load(max(I - 1, 0), a1, x1);
    step(I, a0, x0);
    if (--I < 0) break;
    load(max(I - 1, 0), a0, x0);
    step(I, a1, x1);
    if (--I < 0) break;
  }
  // block column J of X, rows 0 .. j0 - 1 (the diagonal block is in place)
  for (int idx = t; idx < nb * j0; idx += BIG) {
    const int r = idx % j0, c = idx / j0;
    X[(int64_t)(j0 + c) * ldx + r] = XC[c * XI_LD + r];
  }
}

__global__ void zero_lower_kernel(double* G, int64_t n, int64_t ld) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y;
  if (i < n && i > c) G[c * ld + i] = 0.0;
}


// -------------------------------------------------------- column completion
template <typename TQ>
// A few work-groups per column (grid-stride over the rows): nearly every
// column is unflagged, and its work-groups only read the flag.
__global__ void complete_kernel(TQ* Q, int64_t m, int64_t n, int64_t ldq, const int* flags) {
  const int64_t j = blockIdx.y;
  if (!flags[j]) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x)
    Q[j * ldq + i] = (TQ)((i == j) ? 1.0 : 0.0);
}
__global__ void complete_bf16_kernel(uint16_t* Q, int64_t m, int64_t n, int64_t ldq,
                                     const int* flags) {
  const int64_t j = blockIdx.y;
  if (!flags[j]) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x)
    Q[j * ldq + i] = (i == j) ? (uint16_t)0x3f80 : (uint16_t)0;
}

// ------------------------------------------------------ one-sided Jacobi
// Block pairs of 2*JB columns; one work-group per pair per round.  The
// 2*JB columns of W live in LDS; the accumulated rotation J (2JB x 2JB) is
// applied to V at the end of the launch.
constexpr int JB = 16;
constexpr int JW = 2 * JB;   // 32 columns per work-group

// One block pair (bi, bj): the 2*JB columns of W are staged in LDS, one
// inner sweep of rotations (circle-method order, 16 disjoint column pairs per
// sub-round, one wave per pair) is applied to them and accumulated in Jm,
// then W is written back and V[:, cols] = V[:, cols] * Jm.
__device__ void jacobi_pair(double* W, int n, int64_t ldw, double* V, int64_t ldv, int bi, int bj,
                            double tol, int* rot_count, int nbpad, double* Wl,
                            double (*Jm)[JW + 1], int* cols, int* nrot, double* nrm) {
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  if (t < JW) {
    const int blk = (t < JB) ? bi : bj;
    const int c = blk * JB + (t % JB);
    cols[t] = (c < n && blk < nbpad) ? c : -1;
  }
  if (t == 0) *nrot = 0;
  for (int idx = t; idx < JW * JW; idx += BIG) Jm[idx % JW][idx / JW] = (idx % JW == idx / JW) ? 1.0 : 0.0;
  __syncthreads();
  for (int idx = t; idx < n * JW; idx += BIG) {
    const int i = idx % n, c = idx / n;
    Wl[idx] = cols[c] >= 0 ? W[(int64_t)cols[c] * ldw + i] : 0.0;
  }
  __syncthreads();
  // Squared column norms, tracked through the rotations below
  // (||a'||^2 = ||a||^2 - t g, ||b'||^2 = ||b||^2 + t g) so that each rotation
  // needs a single reduction (g = a . b); recomputed every round.
  for (int c = wv; c < JW; c += BIG / 64) {
    double q = 0.0;
    for (int i = ln; i < n; i += 64) { const double x = Wl[c * n + i]; q += x * x; }
    q = wave_sum(q);
    if (ln == 0) nrm[c] = q;
  }
  __syncthreads();
  for (int r = 0; r < JW - 1; ++r) {
    for (int pi = wv; pi < JW / 2; pi += BIG / 64) {
      int a, b;
      if (pi == 0) { a = r; b = JW - 1; }
      else { a = (r + pi) % (JW - 1); b = (r - pi + JW - 1) % (JW - 1); }
      if (a > b) { const int tmp = a; a = b; b = tmp; }
      double ga = 0.0;
      for (int i = ln; i < n; i += 64) ga += Wl[a * n + i] * Wl[b * n + i];
      ga = wave_sum(ga);
      const double al = nrm[a], be = nrm[b];
      if (fabs(ga) > tol * sqrt(al * be) && ga != 0.0) {
        const double zeta = (be - al) / (2.0 * ga);
        const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
        for (int i = ln; i < n; i += 64) {
          const double x = Wl[a * n + i], y = Wl[b * n + i];
          Wl[a * n + i] = cs * x - sn * y;
          Wl[b * n + i] = sn * x + cs * y;
        }
        if (ln < JW) {
          const double x = Jm[ln][a], y = Jm[ln][b];
          Jm[ln][a] = cs * x - sn * y;
          Jm[ln][b] = sn * x + cs * y;
        }
        if (ln == 0) {
          nrm[a] = al - tt * ga;
          nrm[b] = be + tt * ga;
          if (cols[a] >= 0 && cols[b] >= 0) atomicAdd(nrot, 1);
        }
      }
    }
    __syncthreads();
  }
  for (int idx = t; idx < n * JW; idx += BIG) {
    const int i = idx % n, c = idx / n;
    if (cols[c] >= 0) W[(int64_t)cols[c] * ldw + i] = Wl[idx];
  }
  __syncthreads();
  // V[:, cols] = V[:, cols] * Jm, with V[:, cols] staged in LDS.
  for (int idx = t; idx < n * JW; idx += BIG) {
    const int i = idx % n, c = idx / n;
    Wl[idx] = cols[c] >= 0 ? V[(int64_t)cols[c] * ldv + i] : 0.0;
  }
  __syncthreads();
  for (int idx = t; idx < n * JW; idx += BIG) {
    const int i = idx % n, c = idx / n;
    if (cols[c] < 0) continue;
    double acc = 0.0;
#pragma unroll 8
    for (int q = 0; q < JW; ++q) acc += Wl[q * n + i] * Jm[q][c];
    V[(int64_t)cols[c] * ldv + i] = acc;
  }
  if (t == 0 && *nrot) atomicAdd(rot_count, *nrot);
  __syncthreads();
}

// One round per launch (W too large for the Gram-based pair step in LDS).
__global__ __launch_bounds__(BIG) void jacobi_round_kernel(double* W, int n, int64_t ldw,
                                                           double* V, int64_t ldv,
                                                           const int* pairs, double tol,
                                                           int* rot_count, int nbpad) {
  extern __shared__ double shm[];
  __shared__ double Jm[JW][JW + 1];
  __shared__ int cols[JW];
  __shared__ int nrot;
  __shared__ double nrm[JW];
  jacobi_pair(W, n, ldw, V, ldv, pairs[2 * blockIdx.x], pairs[2 * blockIdx.x + 1], tol, rot_count,
              nbpad, shm, Jm, cols, &nrot, nrm);
}

// Same sweep structure as above, but each block-pair step works on the
// 32 x 32 Gram matrix G = W_C^T W_C of its columns instead of on the columns:
//   1. W_C is staged in LDS and G is formed with v_mfma_f64_16x16x4f64
//      (4 tiles x 4 K-quarters, partial sums added in a fixed order);
//   2. if no off-diagonal entry passes the rotation test the step ends here
//      (converged pairs cost one Gram);
//   3. parallel two-sided rotations on G (16 disjoint pairs per sub-round,
//      one thread per entry) accumulate the 32 x 32 rotation Jm;
//   4. W_C <- W_C Jm and V_C <- V_C Jm with MFMA, one wave per 16-row tile.
// In exact arithmetic step 3 applies the same rotations as the one-sided
// sweep (the Gram entries are the column dot products).
constexpr int JG_LD = JW + 1;
constexpr int JG_MAT = JW * JG_LD;   // doubles per 32 x 33 LDS matrix

// Partner of slot x in sub-round r of the circle schedule over JW slots
// (slot JW - 1 meets r; the others pair up as x + x' = 2r mod JW - 1).
__device__ __forceinline__ int circle_partner(int x, int r) {
  if (x == JW - 1) return r;
  if (x == r) return JW - 1;
  int p = (2 * r - x) % (JW - 1);
  return p < 0 ? p + JW - 1 : p;
}

// Pair idx (0 .. JW/2 - 1) of sub-round r as (a, b), a < b.
__device__ __forceinline__ void circle_pair(int idx, int r, int& a, int& b) {
  const int x = idx == 0 ? JW - 1 : (r + idx) % (JW - 1);
  const int y = circle_partner(x, r);
  a = min(x, y);
  b = max(x, y);
}

// The circle schedule as a table: entry r * JW/2 + idx = a | b << 8 for pair
// idx of sub-round r.  Ends with a barrier.
__device__ void fill_pair_table(uint16_t* ptab) {
  for (int i = threadIdx.x; i < (JW - 1) * (JW / 2); i += blockDim.x) {
    int a, b;
    circle_pair(i % (JW / 2), i / (JW / 2), a, b);
    ptab[i] = (uint16_t)(a | (b << 8));
  }
  __syncthreads();
}

// Rotation that zeroes G[a][b] of the 2 x 2 block [al ga; ga be] (a < b):
// [c s; -s c] applied to columns a, b; ta, tb are the new diagonal entries.
// t = tan(theta) is the smaller root of t^2 + 2 zeta t - 1 = 0.
// For |d|, |2 ga| in [1e-140, 1e150] (d = be - al) the short form
// t = sign(zeta) |2 ga| / (|d| + hypot(d, 2 ga)), zeta = d / (2 ga), keeps the
// dependent chain to one 1/sqrt, one reciprocal and one 1/sqrt; outside that
// range t comes from zeta with overflow guards.
__device__ __forceinline__ bool jacobi_rotation(double al, double be, double ga, double tol,
                                                double& c, double& s, double& ta, double& tb) {
  c = 1.0; s = 0.0; ta = al; tb = be;
  const double d = be - al, g2 = ga + ga;
  const double mx = fmax(fabs(d), fabs(g2));
  if (mx < 1e150 && mx > 1e-140) {
    const double x = fma(d, d, g2 * g2);
    const double u = fabs(g2) * rcp_nr(fabs(d) + x * rsq_nr(x));
    const double t = (d == 0.0 || (d > 0.0) == (g2 > 0.0)) ? u : -u;
    const double cc = rsq_nr(fma(t, t, 1.0));
    // |ga| > tol sqrt(al be) squared: an overflow of al be reads as "no
    // rotation", an underflow as "rotate", as the unsquared test decides there
    if (!(ga * ga > (tol * tol) * (al * be)) || ga == 0.0) return false;
    c = cc;
    s = cc * t;
    ta = al - t * ga;
    tb = be + t * ga;
    return true;
  }
  if (!(fabs(ga) > tol * sqrt(al * be)) || ga == 0.0) return false;
  double t;
  if (fabs(ga) > 1e-290) {
    const double zeta = (be - al) * rcp_nr(2.0 * ga);
    const double az = fabs(zeta);
    if (az > 1e150) {
      t = 0.5 * rcp_nr(az);
    } else {
      const double w = fma(az, az, 1.0);
      t = rcp_nr(az + w * rsq_nr(w));
    }
    t = zeta >= 0.0 ? t : -t;
  } else {
    const double zeta = (be - al) / (2.0 * ga);
    t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  }
  c = rsq_nr(fma(t, t, 1.0));
  s = c * t;
  ta = al - t * ga;
  tb = be + t * ga;
  return true;
}

// The step's row tiles of [W_C; V_C] (16 rows each) are updated by S work-
// groups (part = 0 .. S-1), which all load W_C, form G and run the same
// rotations (bit-identical Jm), and each write only the tiles
// u = part + S (wv + 16 it).  pctr (S > 1) counts the pair's work-groups that
// have W_C in LDS; none writes W before it reaches ptgt, so no work-group can
// read a W column another has already updated.  Only part 0 counts rotations.
// With chk, the step only tests: it forms G and, if a pair is above the
// tolerance, raises *chk to chkv (atomic max); W and V are not touched.
__device__ __forceinline__ void jacobi_pair_gram(double* W, int n, int64_t ldw, double* V, int64_t ldv, int bi,
                                 int bj, double tol, int* rot_count, int nbpad, double* shm,
                                 int* cols, int* flag, const uint16_t* ptab, int part, int S,
                                 unsigned* pctr, unsigned ptgt, unsigned* giveup,
                                 unsigned* chk = nullptr, unsigned chkv = 0u) {
  // The thread index through an opaque copy: in the all-sweeps kernel this
  // step sits inside two loops, and addresses derived from threadIdx would
  // otherwise be hoisted out of them and held in registers throughout.
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  const int wv = t >> 6, ln = t & 63;
  // W_C in LDS with its rows zero-padded to whole 16-row tiles (npad) and an
  // odd column stride (conflict-free column reads); ghost columns are zero.
  const int npad = (n + 15) / 16 * 16, ldl = npad + 1;
  double* Wl = shm;
  double* Gb = shm + (size_t)ldl * JW;   // G[0], G[1], Jm[0], Jm[1]
  double* G0 = Gb;
  double* P1 = Gb + JG_MAT;              // K-quarter partials 1..3 alias G[1], Jm[0], Jm[1]
  if (t < JW) {
    const int blk = (t < JB) ? bi : bj;
    const int c = blk * JB + (t % JB);
    cols[t] = (c < n && blk < nbpad) ? c : -1;
  }
  if (t == 0) { flag[0] = 0; flag[1] = 0; }
  __syncthreads();
  // V operands of this wave's V tiles (update step 4), loaded now so that
  // their latency hides behind the Gram and the rotations.  Wave wv owns the
  // 16-row tiles u = part + S (wv + 16 it) (it = 0..3) of [W_C; V_C]; V tiles
  // are u >= nrt, at most two per wave (nrt <= 29 since n <= 468).
  const int nrt = (n + 15) / 16;
  const int need = (nrt - part + S - 1) / S;     // first j = wv + 16 it with a V tile
  const int itv = need > wv ? (need - wv + 15) / 16 : 0;
  double v0[JW / 4], v1[JW / 4];
#pragma unroll
  for (int sl = 0; sl < 2; ++sl) {
    const int u = part + S * (wv + 16 * (itv + sl));
    const int i = (u - nrt) * 16 + (ln & 15);
#pragma unroll
    for (int kk = 0; kk < JW / 4; ++kk) {
      const int k = 4 * kk + (ln >> 4);
      const double x = (u < 2 * nrt && i < n && cols[k] >= 0) ? V[(int64_t)cols[k] * ldv + i] : 0.0;
      if (sl == 0) v0[kk] = x; else v1[kk] = x;
    }
  }
  // W_C into LDS: thread t loads rows (t % 32) + 32 r of column t / 32, in
  // batches of JL loads (one memory round trip per batch)
  static_assert(BIG == 32 * JW, "one 32-thread group per column");
  {
    constexpr int JL = 5;
    const int c = t >> 5, i0 = t & 31, cc = cols[c];
#pragma unroll 1
    for (int r0 = 0; 32 * r0 < npad; r0 += JL) {
      double x[JL];
#pragma unroll
      for (int u = 0; u < JL; ++u) {
        const int i = i0 + 32 * (r0 + u);
        x[u] = (i < n && cc >= 0) ? W[(int64_t)cc * ldw + i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < JL; ++u) {
        const int i = i0 + 32 * (r0 + u);
        if (i < npad) Wl[c * ldl + i] = x[u];
      }
    }
  }
  __syncthreads();
  if (pctr && t == 0) atomicAdd(pctr, 1u);      // this work-group's W_C is read
  // 1. Gram: the three 16 x 16 tiles (0,0), (0,1), (1,1) of the symmetric G,
  // wave wv < 15 -> tile wv % 3, K-slice wv / 3 of five; slice 0 lands in G,
  // slices 1-4 in 256-double partial tiles, summed in slice order.
  if (wv < 15) {
    const int tile = wv % 3, q = wv / 3;
    const int ta = tile == 2 ? 1 : 0, tb = tile == 0 ? 0 : 1;
    const int ks = npad / 4;
    const int k0 = q * ks / 5, k1 = (q + 1) * ks / 5;
    f64x4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    const double* wa = Wl + (ta * 16 + (ln & 15)) * ldl + (ln >> 4);
    const double* wb = Wl + (tb * 16 + (ln & 15)) * ldl + (ln >> 4);
    int kk = k0;
    for (; kk + 1 < k1; kk += 2) {           // two independent accumulators
      const double a0 = wa[4 * kk], b0 = wb[4 * kk], a1 = wa[4 * kk + 4], b1 = wb[4 * kk + 4];
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc1, 0, 0, 0);
    }
    if (kk < k1) acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(wa[4 * kk], wb[4 * kk], acc0, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 4 * v + (ln >> 4), cl = ln & 15;
      const double x = acc0[v] + acc1[v];
      if (q == 0) G0[(ta * 16 + r) * JG_LD + tb * 16 + cl] = x;
      else P1[((q - 1) * 3 + tile) * 256 + r * 16 + cl] = x;
    }
  }
  __syncthreads();
  const int gi = t >> 5, gj = t & 31;   // this thread's entry of the 32 x 32 matrices
  {
    // entry (gi, gj) of G; below the diagonal blocks it is (gj, gi) of (0,1)
    const int lo = gi >= 16 && gj < 16;
    const int ri = lo ? gj : gi, cj = lo ? gi : gj;
    const int ta = ri >> 4, tb = cj >> 4, tile = ta == 0 ? tb : 2;
    const int lr = ri & 15, lc = cj & 15;
    double g = G0[ri * JG_LD + cj];
#pragma unroll
    for (int q = 1; q < 5; ++q) g += P1[((q - 1) * 3 + tile) * 256 + lr * 16 + lc];
    __syncthreads();
    G0[gi * JG_LD + gj] = g;
  }
  __syncthreads();
  // 2. Anything to rotate?
  if (gi < gj && cols[gi] >= 0 && cols[gj] >= 0) {
    const double g = G0[gi * JG_LD + gj];
    if (fabs(g) > tol * sqrt(G0[gi * JG_LD + gi] * G0[gj * JG_LD + gj]) && g != 0.0) flag[0] = 1;
  }
  __syncthreads();
  if (chk) {
    if (t == 0 && flag[0]) atomicMax(chk, chkv);
    return;
  }
  if (!flag[0]) return;
  // 3a. Nearly diagonal G (the late sweeps): the pair's rotation without the
  // sweep.  K_ij = G_ij / (G_jj - G_ii) (i != j) is skew-symmetric and zeroes
  // G's off-diagonal to first order; Newton-Schulz steps J <- J (3 I - J^T J)
  // / 2 from J = I + K make J orthogonal to rounding.  A second pass repeats
  // this on J^T G J (J <- J (I + K'), Newton-Schulz), which leaves an
  // off-diagonal of third order.  Taken when max |K_ij| < 0.4 and
  // ||K||_F^2 < 0.6 (the steps then converge quadratically from
  // ||J^T J - I|| <= ||K||^2 < 0.6; thresholds measured on the 320-column
  // R factor of A 65536 x 16384: 0.3 / 0.5 took 5.50 ms, 0.4 / 0.6 5.30 ms
  // with the same sweeps and error, 0.5 / 0.7 one sweep more); otherwise, or
  // if J^T J is not within 1e-14 of I after the steps, the rotation sweep
  // (3b) runs instead.  Rotations are
  // counted as there: the pairs above the tolerance.  The 32 x 32 products
  // run on FP64 MFMA, one 16 x 16 tile per wave of waves 0-3.
  __shared__ double jred[2 * (BIG / 64)];
  const double* Jf = Gb + 2 * JG_MAT;     // the step's J (3b: Jm)
  bool fast;
  {
    double* Ja = Gb + 2 * JG_MAT;
    const double gij = G0[gi * JG_LD + gj];
    const double gii = G0[gi * JG_LD + gi], gjj = G0[gj * JG_LD + gj];
    double k = 0.0;
    if (gi != gj && gij != 0.0) {
      const double den = gjj - gii;
      k = den != 0.0 ? gij / den : HUGE_VAL;
    }
    const double am = wave_max(fabs(k)), s2 = wave_sum(k * k);
    if (ln == 0) { jred[wv] = am; jred[BIG / 64 + wv] = s2; }
    Ja[gi * JG_LD + gj] = gi == gj ? 1.0 : k;
    __syncthreads();
    double mx = 0.0, ss = 0.0;
#pragma unroll
    for (int w = 0; w < BIG / 64; ++w) { mx = fmax(mx, jred[w]); ss += jred[BIG / 64 + w]; }
    fast = mx < 0.4 && ss < 0.6;
    __syncthreads();
  }
  if (fast) {
    const int ta = wv & 1, tb = (wv >> 1) & 1, lr = ln & 15, lk = ln >> 4;
    // C = op(X) Y (op: X^T if xt) for 32 x 32 matrices (ld JG_LD); with e,
    // max |C - I| of the tile goes to jred[wv].  Waves 0-3; no barrier.
    auto mm = [&](const double* X, bool xt, const double* Y, double* C, bool e) {
      if (wv >= 4) return;
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < JW / 4; ++kk) {
        const int k = 4 * kk + lk;
        const double x = xt ? X[k * JG_LD + 16 * ta + lr] : X[(16 * ta + lr) * JG_LD + k];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(x, Y[k * JG_LD + 16 * tb + lr], acc, 0, 0, 0);
      }
      double d = 0.0, nz = 0.0;             // nz: 0, or NaN for a non-finite entry
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * ta + 4 * v + lk, c = 16 * tb + lr;
        C[r * JG_LD + c] = acc[v];
        d = fmax(d, fabs(acc[v] - (r == c ? 1.0 : 0.0)));
        nz += acc[v] - acc[v];
      }
      if (e) {
        d = wave_max(d) + wave_sum(nz);
        if (ln == 0) jred[wv] = d;
      }
    };
    double* B0 = Gb + JG_MAT;             // three 32 x 33 buffers, roles rotate
    double* J = Gb + 2 * JG_MAT;          // I + K
    double* B1 = Gb + 3 * JG_MAT;
    double err = 1.0;
#pragma unroll 1
    for (int pass = 0; pass < 2 && fast; ++pass) {
      if (pass == 1) {
        mm(G0, false, J, B0, false);      // T = G J
        __syncthreads();
        mm(J, true, B0, B1, false);       // G' = J^T T
        __syncthreads();
        {                                 // B1 <- I + K' (in place, one entry per thread)
          const double g = B1[gi * JG_LD + gj];
          const double den = B1[gj * JG_LD + gj] - B1[gi * JG_LD + gi];
          const double k = (gi == gj || g == 0.0) ? 0.0 : (den != 0.0 ? g / den : HUGE_VAL);
          const double am = wave_max(fabs(k)), s2 = wave_sum(k * k);
          if (ln == 0) { jred[wv] = am; jred[BIG / 64 + wv] = s2; }
          __syncthreads();
          B1[gi * JG_LD + gj] = gi == gj ? 1.0 : k;
        }
        __syncthreads();
        {                                 // the same bounds as the first pass, or keep J
          double mx = 0.0, ss = 0.0;
#pragma unroll
          for (int w = 0; w < BIG / 64; ++w) { mx = fmax(mx, jred[w]); ss += jred[BIG / 64 + w]; }
          __syncthreads();
          if (!(mx < 0.4 && ss < 0.6)) break;
        }
        mm(J, false, B1, B0, false);      // J (I + K')
        __syncthreads();
        double* tmp = J; J = B0; B0 = tmp;
      }
#pragma unroll 1
      for (int it = 0; it < 7; ++it) {    // Newton-Schulz on J
        mm(J, true, J, B0, true);         // P = J^T J
        __syncthreads();
        err = jred[0] + jred[1] + jred[2] + jred[3];   // NaN stays NaN
        err = fmax(fmax(jred[0], jred[1]), fmax(jred[2], jred[3])) + (err - err);
        if (err <= 2e-15 || it == 6) break;
        {                                 // B0 <- 1.5 I - 0.5 P (in place)
          const double pv = B0[gi * JG_LD + gj];
          B0[gi * JG_LD + gj] = (gi == gj ? 1.5 : 0.0) - 0.5 * pv;
        }
        __syncthreads();
        mm(J, false, B0, B1, false);      // J (1.5 I - 0.5 P)
        __syncthreads();
        double* tmp = J; J = B1; B1 = tmp;
      }
      if (!(err <= 1e-14)) fast = false;   // also NaN
    }
    if (fast) {
      Jf = J;
      const bool rot = gi < gj && cols[gi] >= 0 && cols[gj] >= 0 &&
                       fabs(G0[gi * JG_LD + gj]) >
                           tol * sqrt(G0[gi * JG_LD + gi] * G0[gj * JG_LD + gj]) &&
                       G0[gi * JG_LD + gj] != 0.0;
      const uint64_t bm = __ballot(rot);
      if (ln == 0 && bm) atomicAdd(&flag[1], (int)__popcll(bm));
    }
    __syncthreads();
  }
  if (!fast) {
  // 3b. One inner sweep of two-sided rotations on G (circle schedule, 16
  // disjoint pairs per sub-round), one barrier per sub-round: G is double
  // buffered and thread (p, q), p <= q, of waves 0-3 owns the 2 x 2 block of
  // G at rows of pair p and columns of pair q and its transpose.  It computes
  // the rotations of both pairs itself from their 2 x 2 diagonal blocks (no
  // thread waits for another's parameters); the diagonal thread (p, p)
  // publishes pair p's (c, s), which waves 4-11 apply to Jm (row i, pair q,
  // in place) one sub-round later, off the critical path.  Pair indices come
  // from ptab (pair idx of sub-round r at r * 16 + idx, a | b << 8).
  double* Jm = Gb + 2 * JG_MAT;
  double* prm = Gb + 3 * JG_MAT;          // [sub-round parity][pair][c, s]
  Jm[gi * JG_LD + gj] = gi == gj ? 1.0 : 0.0;
  // bit j: column j of the pair is a real column (rotations of ghost
  // columns are not counted)
  const uint64_t vmask = __ballot(ln < JW && cols[ln & (JW - 1)] >= 0);
  auto jm_apply = [&](int r) {            // Jm <- Jm R_r
    const int it = t - 256, i = it >> 4, q = it & 15;
    const unsigned e = ptab[r * (JW / 2) + q];
    const int a = e & 255, b = e >> 8;
    const double c = prm[(r & 1) * JW + 2 * q], sn = prm[(r & 1) * JW + 2 * q + 1];
    const double x = Jm[i * JG_LD + a], y = Jm[i * JG_LD + b];
    Jm[i * JG_LD + a] = c * x - sn * y;
    Jm[i * JG_LD + b] = sn * x + c * y;
  };
  // Every lane of waves 0-3 computes the rotation of pair q = t % 16 (the
  // four 16-lane groups of a wave redundantly); pair p's comes from lane p of
  // the same wave (whose q is p) by a lane shuffle, so each lane evaluates
  // one rotation instead of two.
  const int p = (t >> 4) & 15, q = t & 15;
  const bool rotw = wv < 4;
  const bool quad = rotw && p <= q;
  unsigned ep = 0, eq = 0;
  if (rotw) { ep = ptab[p]; eq = ptab[q]; }
  int cur = 0;
  __syncthreads();
#pragma unroll 1
  for (int r = 0; r < JW - 1; ++r) {
    if (rotw) {
      const double* G = Gb + cur * JG_MAT;
      double* Gn = Gb + (cur ^ 1) * JG_MAT;
      const int ap = ep & 255, bp = ep >> 8, aq = eq & 255, bq = eq >> 8;
      const double alq = G[aq * JG_LD + aq], beq = G[bq * JG_LD + bq], gaq = G[aq * JG_LD + bq];
      double g00 = 0.0, g01 = 0.0, g10 = 0.0, g11 = 0.0;
      if (quad) {
        g00 = G[ap * JG_LD + aq]; g01 = G[ap * JG_LD + bq];
        g10 = G[bp * JG_LD + aq]; g11 = G[bp * JG_LD + bq];
      }
      if (r + 1 < JW - 1) { ep = ptab[(r + 1) * (JW / 2) + p]; eq = ptab[(r + 1) * (JW / 2) + q]; }
      double cq, sq, taq, tbq;
      const bool onq = jacobi_rotation(alq, beq, gaq, tol, cq, sq, taq, tbq);
      const double cp = __shfl(cq, p), sp = __shfl(sq, p);
      if (quad && p == q) {
        Gn[aq * JG_LD + aq] = taq;
        Gn[bq * JG_LD + bq] = tbq;
        Gn[aq * JG_LD + bq] = 0.0;
        Gn[bq * JG_LD + aq] = 0.0;
        prm[(r & 1) * JW + 2 * q] = cq;
        prm[(r & 1) * JW + 2 * q + 1] = sq;
        if (onq && ((vmask >> aq) & (vmask >> bq) & 1)) atomicAdd(&flag[1], 1);
      } else if (quad) {
        // columns (a' = c a - s b, b' = s a + c b), then the same on rows
        const double h00 = cq * g00 - sq * g01, h01 = sq * g00 + cq * g01;
        const double h10 = cq * g10 - sq * g11, h11 = sq * g10 + cq * g11;
        const double n00 = cp * h00 - sp * h10, n01 = cp * h01 - sp * h11;
        const double n10 = sp * h00 + cp * h10, n11 = sp * h01 + cp * h11;
        Gn[ap * JG_LD + aq] = n00; Gn[aq * JG_LD + ap] = n00;
        Gn[ap * JG_LD + bq] = n01; Gn[bq * JG_LD + ap] = n01;
        Gn[bp * JG_LD + aq] = n10; Gn[aq * JG_LD + bp] = n10;
        Gn[bp * JG_LD + bq] = n11; Gn[bq * JG_LD + bp] = n11;
      }
    } else if (wv >= 4 && wv < 12 && r > 0) {
      jm_apply(r - 1);
    }
    __syncthreads();
    cur ^= 1;
  }
  if (wv >= 4 && wv < 12) jm_apply(JW - 2);
  __syncthreads();
  }
  // 4. W_C <- W_C J (from LDS), V_C <- V_C J (prefetched), MFMA per 16-row tile.
  if (pctr) {                                 // every part has read W_C
    if (t == 0) {
      unsigned spins = 0;
      while (__hip_atomic_load(pctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < ptgt) {
        if (__hip_atomic_load(giveup, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
            ++spins > (1u << 24)) {
          atomicExch(giveup, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
  }
  const int c0 = cols[ln & 15], c1 = cols[16 + (ln & 15)];
  auto store_tile = [&](double* dst, int64_t ld, int rt, const f64x4& acc0, const f64x4& acc1) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = rt * 16 + 4 * v + (ln >> 4);
      if (row < n) {
        if (c0 >= 0) dst[(int64_t)c0 * ld + row] = acc0[v];
        if (c1 >= 0) dst[(int64_t)c1 * ld + row] = acc1[v];
      }
    }
  };
#pragma unroll
  for (int it = 0; it < 2; ++it) {          // W tiles u < nrt (nrt <= 29)
    const int rt = part + S * (wv + 16 * it);
    if (rt < nrt) {
      const int i = rt * 16 + (ln & 15);
      f64x4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < JW / 4; ++kk) {
        const int k = 4 * kk + (ln >> 4);
        const double a = Wl[k * ldl + i];    // rows < npad; ghost columns are zero
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Jf[k * JG_LD + (ln & 15)], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Jf[k * JG_LD + 16 + (ln & 15)], acc1, 0, 0, 0);
      }
      store_tile(W, ldw, rt, acc0, acc1);
    }
  }
#pragma unroll
  for (int sl = 0; sl < 2; ++sl) {          // V tiles, operands in v0 / v1
    const int u = part + S * (wv + 16 * (itv + sl));
    if (u < 2 * nrt) {
      f64x4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < JW / 4; ++kk) {
        const int k = 4 * kk + (ln >> 4);
        const double a = sl == 0 ? v0[kk] : v1[kk];
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Jf[k * JG_LD + (ln & 15)], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Jf[k * JG_LD + 16 + (ln & 15)], acc1, 0, 0, 0);
      }
      store_tile(V, ldv, u - nrt, acc0, acc1);
    }
  }
  __syncthreads();
  if (t == 0 && flag[1] && part == 0) atomicAdd(rot_count, flag[1]);
  __syncthreads();
}

__global__ __launch_bounds__(BIG) void jacobi_gram_round_kernel(double* W, int n, int64_t ldw,
                                                                double* V, int64_t ldv,
                                                                const int* pairs, double tol,
                                                                int* rot_count, int nbpad) {
  extern __shared__ double shm[];
  __shared__ int cols[JW], flag[2];
  __shared__ uint16_t ptab[(JW - 1) * (JW / 2)];
  fill_pair_table(ptab);
  jacobi_pair_gram(W, n, ldw, V, ldv, pairs[2 * blockIdx.x], pairs[2 * blockIdx.x + 1], tol,
                   rot_count, nbpad, shm, cols, flag, ptab, 0, 1, nullptr, 0u, nullptr);
}

// All sweeps in one cooperative launch: rounds separated by a grid barrier
// (bounded wait), rotations of sweep s counted in hist[s]; the launch ends
// after the first sweep without rotations or after max_sweeps.  ctl[0] is
// the barrier counter, ctl[1] the give-up flag, ctl[2] the sweep count,
// ctl[3] the last sweep whose test pass found a pair to rotate (plus one);
// ctl[4 + max_sweeps + i] counts the W_C reads of pair slot i.  Work-group g
// takes pair slot g / S, part g % S.
// A sweep after one with at most 4 n rotations starts with a test pass: the
// work-groups share out all its block pairs and only form G and test it (no
// writes to W).  If no pair is above the tolerance, the sweep's rounds would
// see the same W and rotate nothing, so the sweep is counted as one without
// rotations and the launch ends without running them.
__global__ __launch_bounds__(BIG) void jacobi_gram_sweep_kernel(double* W, int n, int64_t ldw,
                                                                double* V, int64_t ldv,
                                                                const int* pairs, int rounds,
                                                                int per, double tol, int* hist,
                                                                int max_sweeps, int nbpad,
                                                                unsigned* ctl, int S) {
  extern __shared__ double shm[];
  __shared__ int cols[JW], flag[2];
  __shared__ int bad, stop, tst;
  __shared__ uint16_t ptab[(JW - 1) * (JW / 2)];
  fill_pair_table(ptab);
  unsigned gen = 0;   // grid barriers passed
  unsigned rnd = 0;   // pair rounds run (the W_C read counters count these)
  // Grid barrier; false when a wait gave up (then every work-group returns).
  auto grid_sync = [&]() -> bool {
    ++gen;
    if (threadIdx.x == 0) {
      __threadfence();
      atomicAdd(ctl, 1u);
      const unsigned target = gen * gridDim.x;
      unsigned spins = 0;
      int e = 0;
      // poll without acquire (an acquire invalidates the L2 on every poll),
      // then one acquire
      while (__hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        e = (int)__hip_atomic_load(ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (e || ++spins > (1u << 24)) { atomicExch(ctl + 1, 1u); e = 1; break; }
        __builtin_amdgcn_s_sleep(2);
      }
      __threadfence();
      bad = e;
    }
    __syncthreads();
    return !bad;
  };
  int sw = 0;
  while (sw < max_sweeps) {
    // (tst, not stop: a wave may still be reading the last sweep's stop)
    if (threadIdx.x == 0)
      tst = sw > 0 && __hip_atomic_load(hist + sw - 1, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT) <= 4 * n;
    __syncthreads();
    if (tst) {                                    // test pass
#pragma unroll 1
      for (int e = blockIdx.x; e < rounds * per; e += gridDim.x)
        jacobi_pair_gram(W, n, ldw, V, ldv, pairs[2 * e], pairs[2 * e + 1], tol, nullptr, nbpad,
                         shm, cols, flag, ptab, 0, 1, nullptr, 0u, ctl + 1, ctl + 3,
                         (unsigned)sw + 1);
      if (!grid_sync()) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl[2] = (unsigned)(sw + 1);
        return;
      }
      if (threadIdx.x == 0)
        stop = __hip_atomic_load(ctl + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
               (unsigned)sw + 1;
      __syncthreads();
      if (stop) { ++sw; break; }                  // hist[sw] stays 0
    }
#pragma unroll 1
    for (int r = 0; r < rounds; ++r) {
      const int slot = blockIdx.x / S, part = blockIdx.x % S;
      const int* pr = pairs + 2 * (r * per + slot);
      jacobi_pair_gram(W, n, ldw, V, ldv, pr[0], pr[1], tol, hist + sw, nbpad, shm, cols, flag,
                       ptab, part, S, S > 1 ? ctl + 4 + max_sweeps + slot : nullptr, (rnd + 1) * (unsigned)S,
                       ctl + 1);
      ++rnd;
      if (!grid_sync()) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl[2] = (unsigned)(sw + 1);
        return;
      }
    }
    if (threadIdx.x == 0)
      stop = __hip_atomic_load(hist + sw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    __syncthreads();
    ++sw;
    if (stop) break;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl[2] = (unsigned)sw;
}

// Norms, stable descending sort, normalised U, permuted V (n <= SF_MAXN).
constexpr int SF_MAXN = 8192;
__global__ __launch_bounds__(BIG) void svd_finalize_kernel(const double* W, int n, int64_t ldw,
                                                           const double* V, int64_t ldv,
                                                           double* s, double* U, int64_t ldu,
                                                           double* Vo, int64_t ldvo,
                                                           double* so) {
  __shared__ int src[SF_MAXN];
  __shared__ double sl[SF_MAXN];                   // the column norms
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  // Loads in chunks of eight per lane, issued together (rows past n read
  // row 0 and are masked).
  constexpr int CH = 8;
  for (int c = wv; c < n; c += BIG / 64) {
    const double* wc = W + (int64_t)c * ldw;
    double a = 0.0;
    for (int i0 = ln; i0 < n; i0 += 64 * CH) {
      double x[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) x[u] = wc[i0 + 64 * u < n ? i0 + 64 * u : 0];
#pragma unroll
      for (int u = 0; u < CH; ++u) a += i0 + 64 * u < n ? x[u] * x[u] : 0.0;
    }
    a = wave_sum(a);
    if (ln == 0) { s[c] = sqrt(a); sl[c] = sqrt(a); }
  }
  __syncthreads();
  // rank of column c (descending, ties by index) -> src[rank]; then wave wv
  // writes output columns r = wv, wv + 16, ... with its lanes along the rows.
  // (NaN norms can give equal ranks: src starts as the identity, so every
  // entry names a column.)
  for (int r = t; r < n; r += BIG) src[r] = r;
  __syncthreads();
  for (int c = t; c < n; c += BIG) {
    const double sc = sl[c];
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      const double si = sl[i];                     // one LDS broadcast per step
      rank += (si > sc) || (si == sc && i < c);
    }
    so[rank] = sc;
    src[rank] = c;
  }
  __syncthreads();
  for (int r = wv; r < n; r += BIG / 64) {
    const int c = src[r];
    const double sc = sl[c];
    const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
    const double* wc = W + (int64_t)c * ldw;
    const double* vc = V + (int64_t)c * ldv;
    for (int i0 = ln; i0 < n; i0 += 64 * CH) {
      double x[CH], y[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int i = i0 + 64 * u < n ? i0 + 64 * u : 0;
        x[u] = wc[i];
        y[u] = vc[i];
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int i = i0 + 64 * u;
        if (i < n) {
          U[(int64_t)r * ldu + i] = sc > 0.0 ? x[u] * inv : (i == r ? 1.0 : 0.0);
          Vo[(int64_t)r * ldvo + i] = y[u];
        }
      }
    }
  }
}

// ------------------------------------------------- Householder QR (pivoted)
// Step k, part 1 (one work-group): optional pivot (first maximum of the
// recomputed trailing norms), then the reflector v (unit norm) for column k.
__global__ __launch_bounds__(BIG) void hh_step1(double* A, int64_t m, int64_t n, int64_t lda,
                                                int64_t k, double* norms, int64_t* perm,
                                                double* vbuf, double* flag, int pivot,
                                                double* Vh, int64_t ldv, double* tau) {
  __shared__ double sh[16];
  __shared__ int64_t piv;
  __shared__ double bv[BIG];
  __shared__ int64_t bi[BIG];
  const int t = threadIdx.x;
  if (pivot) {
    // First maximum of norms[k..n): each thread keeps the first maximum of
    // its strided subset, the tree keeps the larger value and, on ties, the
    // smaller index.  A NaN never wins; a NaN at k keeps k (no norm compares
    // greater than it).
    double best = -1.0;
    int64_t bj = n;
    for (int64_t c = k + t; c < n; c += BIG) {
      const double v = norms[c];
      if (v > best) { best = v; bj = c; }
    }
    bv[t] = best;
    bi[t] = bj;
    __syncthreads();
    for (int off = BIG / 2; off > 0; off >>= 1) {
      if (t < off) {
        const double v = bv[t + off];
        const int64_t j2 = bi[t + off];
        if (v > bv[t] || (v == bv[t] && j2 < bi[t])) { bv[t] = v; bi[t] = j2; }
      }
      __syncthreads();
    }
    if (t == 0) {
      const double nk = norms[k];
      const int64_t j = (nk != nk || bi[0] >= n) ? k : bi[0];
      piv = j;
      if (j != k) {
        const int64_t tp = perm[k]; perm[k] = perm[j]; perm[j] = tp;
        const double tn = norms[k]; norms[k] = norms[j]; norms[j] = tn;
      }
    }
    __syncthreads();
    const int64_t j = piv;
    if (j != k)
      for (int64_t i = t; i < m; i += BIG) {
        const double a = A[k * lda + i];
        A[k * lda + i] = A[j * lda + i];
        A[j * lda + i] = a;
      }
    __syncthreads();
  }
  double s = 0.0;
  for (int64_t i = k + t; i < m; i += BIG) { const double x = A[k * lda + i]; s += x * x; }
  s = block_sum(s, sh);
  const double nx = sqrt(s);
  // The reflector also goes to column k of the panel Vh (zeros above row k,
  // and everywhere for a zero column) with tau[k] = 1 (active) or 0.
  double* vk = Vh + k * ldv;
  for (int64_t i = t; i < k; i += BIG) vk[i] = 0.0;
  if (nx == 0.0) {
    for (int64_t i = k + t; i < m; i += BIG) vk[i] = 0.0;
    if (t == 0) { flag[0] = 0.0; tau[k] = 0.0; }
    return;
  }
  const double x0 = A[k * lda + k];
  const double v0 = x0 + (x0 >= 0.0 ? nx : -nx);
  // ||v||^2 = s - x0^2 + v0^2
  const double nv = sqrt(s - x0 * x0 + v0 * v0);
  for (int64_t i = k + t; i < m; i += BIG) {
    const double v = ((i == k) ? v0 : A[k * lda + i]) / nv;
    vbuf[i] = v;
    vk[i] = v;
  }
  if (t == 0) { flag[0] = 1.0; tau[k] = 1.0; }
}

// Part 2 (one work-group per column c >= k): A[k:, c] -= 2 v (v . A[k:, c]);
// then recompute the trailing norm of column c (rows k+1..m-1).
__global__ __launch_bounds__(256) void hh_step2(double* A, int64_t m, int64_t n, int64_t lda,
                                                int64_t k, const double* vbuf,
                                                const double* flag, double* norms) {
  __shared__ double sh[4];
  if (flag[0] == 0.0) return;
  const int64_t c = k + blockIdx.x;
  const int t = threadIdx.x;
  double* col = A + c * lda;
  double s = 0.0;
  for (int64_t i = k + t; i < m; i += 256) s += vbuf[i] * col[i];
  s = block_sum(s, sh);
  double q = 0.0;
  for (int64_t i = k + t; i < m; i += 256) {
    const double v = col[i] - 2.0 * vbuf[i] * s;
    col[i] = v;
    if (i > k) q += v * v;
  }
  q = block_sum(q, sh);
  if (t == 0 && c > k) norms[c] = sqrt(q);
}

__global__ void init_norms(const double* A, int64_t m, int64_t n, int64_t lda, double* norms,
                           int64_t* perm) {
  __shared__ double sh[4];
  const int64_t c = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += 256) { const double x = A[c * lda + i]; s += x * x; }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) { norms[c] = sqrt(s); perm[c] = c; }
}

// Q[k:, j] -= 2 v (v . Q[k:, j]) for all columns j (one work-group each).
__global__ __launch_bounds__(256) void hh_apply_q(const double* Vh, int64_t m, int64_t ldv,
                                                  int64_t k, const double* tau, double* Q,
                                                  int64_t ldq) {
  __shared__ double sh[4];
  if (tau[k] == 0.0) return;
  const int64_t j = blockIdx.x;
  const double* v = Vh + k * ldv;
  double* col = Q + j * ldq;
  double s = 0.0;
  for (int64_t i = k + threadIdx.x; i < m; i += 256) s += v[i] * col[i];
  s = block_sum(s, sh);
  for (int64_t i = k + threadIdx.x; i < m; i += 256) col[i] -= 2.0 * v[i] * s;
}

// ------------------------------------- Householder QR (unpivoted), one launch
// The same reflectors as hh_step1/hh_step2 (H = I - 2 v v^T with v of unit
// norm, here scaled by the reciprocal of its norm; tau[k] = 1, or 0 and
// v = 0 for a zero column), in one cooperative launch.  Blocks of b columns
// are dealt round-robin to the G work-groups, which keep their columns in
// LDS (rows padded with zeros to a multiple of HQ_THR) for the whole
// factorisation.  Only v_k crosses work-groups, through the reflector panel
// Vh, which holds HQ_EMPTY in rows >= k (zeros above) before the launch:
// the owner of column k stores v_k with agent-scope atomic stores, and at
// step k every work-group that still owns a column c > k polls its rows of
// Vh column k until none is HQ_EMPTY (each element is its own flag, so no
// fences are needed), then applies H_k to those columns.  The owner of
// column k + 1 updates that column first, forms and stores v_{k+1}, and only
// then updates its other columns, so one step's critical path is one column
// update plus one reflector (plus the hand-over once per block).  Columns
// nf..nu-1 (right-hand sides) receive every H_k but are not factored.  Polls
// are bounded: the first to give up raises ctl[0] and every work-group
// leaves.
//
// Thread t owns the rows t + HQ_THR * j (row group j) of every LDS column
// and of v for the whole launch, so no thread reads LDS another wrote except
// through hq_block_sums.  v is zero outside rows k..m-1, so the column
// updates run unmasked over the row groups from k / HQ_THR on.
constexpr int HQ_THR = 256, HQ_NC = 4;
constexpr int HQ_RB = 8;                   // row groups per unrolled batch
constexpr int HQ_B = 2;                    // columns per ownership block
constexpr size_t HQ_LDS_MAX = 150 * 1024;
constexpr unsigned long long HQ_EMPTY = 0x7FF4DEADBEEF0001ull;   // a NaN no arithmetic yields

// DPP move of a double (both halves) with the row-local pattern CTRL.
template <int CTRL>
__device__ __forceinline__ double hq_dpp(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Wave sum, the same bits in every lane: pairs, quads, 8 and 16 lanes by
// row-local DPP moves (each lane adds the mirrored partner, so all lanes of
// a row hold the same value), then the four rows added in a fixed order.
__device__ __forceinline__ double hq_wave_sum(double x) {
  x += hq_dpp<0xB1>(x);    // quad_perm [1, 0, 3, 2]
  x += hq_dpp<0x4E>(x);    // quad_perm [2, 3, 0, 1]
  x += hq_dpp<0x141>(x);   // row_half_mirror
  x += hq_dpp<0x140>(x);   // row_mirror
  return ((readlane_f64(x, 0) + readlane_f64(x, 16)) + readlane_f64(x, 32)) + readlane_f64(x, 48);
}

// Work-group sums of NC values per thread, added in a fixed order; every
// thread receives the totals.  Two halves of red (each HQ_THR / WAVE x
// HQ_NC) alternate with the parity ph, so one barrier per call suffices: a
// half is written again only after every thread has passed the next call's
// barrier, i.e. after all reads of this call.
template <int NC>
__device__ __forceinline__ void hq_block_sums(double (&x)[NC], double* red, int& ph) {
  const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  double* r = red + ph * (HQ_THR / WAVE) * HQ_NC;
  ph ^= 1;
#pragma unroll
  for (int q = 0; q < NC; ++q) x[q] = hq_wave_sum(x[q]);
  if (l == 0) {
#pragma unroll
    for (int q = 0; q < NC; ++q) r[w * HQ_NC + q] = x[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < HQ_THR / WAVE; ++i) s += r[i * HQ_NC + q];
    x[q] = s;
  }
}

// Apply H = I - 2 v v^T to the NC LDS columns from q0 (ng row groups).
template <int NC>
__device__ __forceinline__ void hq_apply_n(double* cols, int mp, int ng, int q0, int k,
                                           const double* v, double* red, int& ph) {
  double x[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) x[q] = 0.0;
  const int j0 = k / HQ_THR;
  for (int jb = j0; jb < ng; jb += HQ_RB) {
#pragma unroll
    for (int r = 0; r < HQ_RB; ++r) {
      if (jb + r < ng) {
        const int i = (jb + r) * HQ_THR + (int)threadIdx.x;
        const double vi = v[i];
#pragma unroll
        for (int q = 0; q < NC; ++q) x[q] += vi * cols[(q0 + q) * mp + i];
      }
    }
  }
  hq_block_sums<NC>(x, red, ph);
#pragma unroll
  for (int q = 0; q < NC; ++q) x[q] *= 2.0;
  for (int jb = j0; jb < ng; jb += HQ_RB) {
#pragma unroll
    for (int r = 0; r < HQ_RB; ++r) {
      if (jb + r < ng) {
        const int i = (jb + r) * HQ_THR + (int)threadIdx.x;
        const double vi = v[i];
#pragma unroll
        for (int q = 0; q < NC; ++q) cols[(q0 + q) * mp + i] -= x[q] * vi;
      }
    }
  }
}

// The same for the LDS columns qa..qb-1, in groups of 4, 2 and 1.
__device__ __forceinline__ void hq_apply(double* cols, int mp, int ng, int qa, int qb, int k,
                                         const double* v, double* red, int& ph) {
  int q0 = qa;
  for (; q0 + 4 <= qb; q0 += 4) hq_apply_n<4>(cols, mp, ng, q0, k, v, red, ph);
  if (q0 + 2 <= qb) { hq_apply_n<2>(cols, mp, ng, q0, k, v, red, ph); q0 += 2; }
  if (q0 < qb) hq_apply_n<1>(cols, mp, ng, q0, k, v, red, ph);
}

// Reflector of LDS column q at step k: v into the LDS buffer v (zero outside
// rows k..m-1) and into Vh column k (rows k..m-1), tau[k]; the column
// becomes R's column (R_kk = -sign(x0) ||x||, zeros below).  x0 reaches
// every thread through the work-group sum (only row k's thread adds it).
__device__ __forceinline__ void hq_reflector(double* cols, int m, int mp, int ng, int q, int k,
                                             double* v, double* Vh, int64_t ldv, double* tau,
                                             double* red, int& ph) {
  double* col = cols + q * mp;
  double x[2] = {0.0, 0.0};
  const int j0 = k / HQ_THR;
  for (int jb = j0; jb < ng; jb += HQ_RB) {
#pragma unroll
    for (int r = 0; r < HQ_RB; ++r) {
      if (jb + r < ng) {
        const int i = (jb + r) * HQ_THR + (int)threadIdx.x;
        const double c = col[i];
        x[0] += i >= k ? c * c : 0.0;
        x[1] += i == k ? c : 0.0;
      }
    }
  }
  hq_block_sums<2>(x, red, ph);
  const double s = x[0], x0 = x[1], nx = sqrt(s);
  const double v0 = x0 + (x0 >= 0.0 ? nx : -nx);
  const double inv = nx != 0.0 ? 1.0 / sqrt(s - x0 * x0 + v0 * v0) : 0.0;
  const double rkk = x0 >= 0.0 ? -nx : nx;
  unsigned long long* vg = reinterpret_cast<unsigned long long*>(Vh + k * ldv);
  for (int jb = j0; jb < ng; jb += HQ_RB) {
#pragma unroll
    for (int r = 0; r < HQ_RB; ++r) {
      if (jb + r < ng) {
        const int i = (jb + r) * HQ_THR + (int)threadIdx.x;
        const double c = col[i];
        const double vi = i > k ? c * inv : (i == k ? v0 * inv : 0.0);
        v[i] = vi;
        col[i] = (i < k || nx == 0.0) ? c : (i == k ? rkk : 0.0);
        if (i >= k && i < m)
          __hip_atomic_store(vg + i, (unsigned long long)__double_as_longlong(vi),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  if (threadIdx.x == 0) tau[k] = nx != 0.0 ? 1.0 : 0.0;
}

// Column of local index q in work-group g (blocks of b columns dealt
// round-robin over the G work-groups).
__device__ __forceinline__ int64_t hq_col(int q, int g, int G, int b) {
  return ((int64_t)(q / b) * G + g) * b + q % b;
}

__global__ __launch_bounds__(HQ_THR) void hh_qr_coop_kernel(double* A, int m, int mp, int nf,
                                                            int nu, int b, int64_t lda, double* Vh,
                                                            int64_t ldv, double* tau,
                                                            unsigned* ctl) {
  extern __shared__ double hq_sm[];
  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const int ng = mp / HQ_THR;                            // row groups
  // blocks of b columns: block j belongs to work-group j % G
  const int nblk = (nu + b - 1) / b;
  const int cnt = g < nblk ? (nblk - 1 - g) / G + 1 : 0;
  int nown = cnt * b;
  if (cnt > 0 && g + (cnt - 1) * G == nblk - 1) nown = (cnt - 1) * b + (nu - (nblk - 1) * b);
  const int nmax = ((nblk + G - 1) / G) * b;
  double* cols = hq_sm;                                  // [nown][mp]
  double* vb = cols + nmax * mp;                         // [2][mp]: v_k in vb[k & 1]
  double* red = vb + 2 * mp;                             // [2][HQ_THR / WAVE][HQ_NC]
  int ph = 0;
  for (int q = 0; q < nown; ++q) {
    const int64_t c = hq_col(q, g, G, b);
    for (int i = t; i < mp; i += HQ_THR) cols[q * mp + i] = i < m ? A[c * lda + i] : 0.0;
  }
  const int64_t cmax = nown > 0 ? hq_col(nown - 1, g, G, b) : -1;
  if (nf > 0 && g == 0) hq_reflector(cols, m, mp, ng, 0, 0, vb, Vh, ldv, tau, red, ph);
  for (int k = 0; k < nf; ++k) {
    if (cmax <= k) break;                                // no column > k left here
    double* v = vb + (k & 1) * mp;
    if ((k / b) % G != g) {
      // v_k from Vh: a thread's rows are loaded as one batch (one round trip)
      // and the batch is reloaded while any of them is still empty
      const unsigned long long* vg = reinterpret_cast<const unsigned long long*>(Vh + k * ldv);
      int e = 0;
      for (int jb = k / HQ_THR; jb < ng && !e; jb += HQ_RB) {
        unsigned long long bits[HQ_RB];
        unsigned spins = 0;
        for (;;) {
          bool empty = false;
#pragma unroll
          for (int r = 0; r < HQ_RB; ++r) {
            const int i = (jb + r) * HQ_THR + t;
            bits[r] = (jb + r < ng && i >= k && i < m)
                          ? __hip_atomic_load(vg + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                          : 0ull;
            empty |= bits[r] == HQ_EMPTY;
          }
          if (!empty) break;
          ++spins;
          if (spins > (1u << 22) ||
              ((spins & 255u) == 0 &&
               __hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            e = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (!e) {
#pragma unroll
          for (int r = 0; r < HQ_RB; ++r)
            if (jb + r < ng) v[(jb + r) * HQ_THR + t] = __longlong_as_double((long long)bits[r]);
        }
      }
      if (e) atomicExch(ctl, 1u);
      if (__syncthreads_or(e)) return;
    }
    // qs: owned columns <= k, i.e. the local index of the first column > k
    const int kb = k / b;
    const int cb = kb >= g ? (kb - g) / G + 1 : 0;       // owned blocks <= kb
    const int qs = min(nown, kb % G == g ? (cb - 1) * b + k % b + 1 : cb * b);
    if (k + 1 < nf && ((k + 1) / b) % G == g) {
      hq_apply(cols, mp, ng, qs, qs + 1, k, v, red, ph);
      hq_reflector(cols, m, mp, ng, qs, k + 1, vb + ((k + 1) & 1) * mp, Vh, ldv, tau, red, ph);
      hq_apply(cols, mp, ng, qs + 1, nown, k, v, red, ph);
    } else {
      hq_apply(cols, mp, ng, qs, nown, k, v, red, ph);
    }
  }
  for (int q = 0; q < nown; ++q) {
    const int64_t c = hq_col(q, g, G, b);
    for (int i = t; i < m; i += HQ_THR) A[c * lda + i] = cols[q * mp + i];
  }
}

// Vh (m x nf, leading dimension m) before the launch: zeros above the
// diagonal, HQ_EMPTY on and below it.
__global__ void hq_fill_vh_kernel(double* Vh, int64_t m, int64_t nf) {
  const int64_t n = m * nf;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e / m, i = e % m;
    reinterpret_cast<unsigned long long*>(Vh)[e] = i < k ? 0ull : HQ_EMPTY;
  }
}

__global__ void zero_u32_kernel(unsigned* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ------------------------------------------------------------ elementwise
template <typename TS, typename TD>
__global__ void convert_kernel(const TS* __restrict__ s, int64_t M, int64_t lds, TD* __restrict__ d,
                               int64_t ldd, double scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= M) return;
  d[j * ldd + i] = (TD)(scale * (double)s[j * lds + i]);
}
// FP32 -> BF16, four elements per thread.
__global__ void f32_to_bf16_kernel(const float* __restrict__ s, int64_t M, int64_t lds,
                                   uint16_t* __restrict__ d, int64_t ldd, float scale) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  const int64_t j = blockIdx.y;
  if (i >= M) return;
  const float* sp = s + j * lds + i;
  uint16_t* dp = d + j * ldd + i;
  if (i + 3 < M && ((lds & 3) == 0) && ((ldd & 3) == 0)) {
    const f32x4 v = *(const f32x4*)sp;
    *(u32x2v*)dp = u32x2v{bf16_bits(v[0] * scale) | (bf16_bits(v[1] * scale) << 16),
                          bf16_bits(v[2] * scale) | (bf16_bits(v[3] * scale) << 16)};
  } else {
    for (int e = 0; e < 4 && i + e < M; ++e) dp[e] = (uint16_t)bf16_bits(sp[e] * scale);
  }
}
// BF16 copy of an FP32 matrix that also returns max |s| (atomicMax on the
// bits of a non-negative double, one per work-group).  Work-group g handles
// columns g, g + grid, ...; needs M, lds, ldd multiples of 4.
__global__ __launch_bounds__(256) void f32_to_bf16_amax_kernel(const float* __restrict__ s,
                                                               int64_t M, int64_t N, int64_t lds,
                                                               uint16_t* __restrict__ d,
                                                               int64_t ldd,
                                                               unsigned long long* amax) {
  __shared__ float sh[4];
  float mx = 0.f;
  for (int64_t j = blockIdx.x; j < N; j += gridDim.x) {
    const float* sp = s + j * lds;
    uint16_t* dp = d + j * ldd;
    for (int64_t i = 4 * (int64_t)threadIdx.x; i < M; i += 1024) {
      const f32x4 v = *(const f32x4*)(sp + i);
      mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))));
      *(u32x2v*)(dp + i) = u32x2v{bf16_bits(v[0]) | (bf16_bits(v[1]) << 16),
                                  bf16_bits(v[2]) | (bf16_bits(v[3]) << 16)};
    }
  }
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double r = (double)fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
    atomicMax(amax, (unsigned long long)__double_as_longlong(r));
  }
}

// One pass over an FP32 matrix S (M x N, column-major, M and N multiples of
// 4): its BF16 copy D (column-major, ldd), its FP32 transpose T (N x M
// column-major, ldt: the row-major copy, K-contiguous for S * X) and
// max |S| (atomicMax on the bits, one per work-group).  64 x 64 tiles go
// through LDS (stride 65) so that both the reads of S and the writes of T
// are coalesced 16-byte accesses.
__global__ __launch_bounds__(256) void f32_bf16_transpose_kernel(const float* __restrict__ s,
                                                                 int64_t M, int64_t N,
                                                                 int64_t lds,
                                                                 uint16_t* __restrict__ d,
                                                                 int64_t ldd,
                                                                 float* __restrict__ t,
                                                                 int64_t ldt,
                                                                 unsigned long long* amax) {
  __shared__ float tile[64][65];
  __shared__ float sh[4];
  const int tx = threadIdx.x;
  const int r4 = (tx & 15) * 4, c0 = tx >> 4;
  const int64_t mt = (M + 63) / 64, ntiles = mt * ((N + 63) / 64);
  float mx = 0.f;
  for (int64_t tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
  const int64_t i0 = (tix % mt) * 64, j0 = (tix / mt) * 64;
  __syncthreads();                 // the previous tile's reads of `tile` are done
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int c = c0 + 16 * q;
    const int64_t i = i0 + r4, j = j0 + c;
    f32x4 v = {0.f, 0.f, 0.f, 0.f};
    if (i < M && j < N) {
      v = *(const f32x4*)(s + j * lds + i);
      mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))));
      *(u32x2v*)(d + j * ldd + i) = u32x2v{bf16_bits(v[0]) | (bf16_bits(v[1]) << 16),
                                           bf16_bits(v[2]) | (bf16_bits(v[3]) << 16)};
    }
    tile[c][r4] = v[0];
    tile[c][r4 + 1] = v[1];
    tile[c][r4 + 2] = v[2];
    tile[c][r4 + 3] = v[3];
  }
  __syncthreads();
  // T[j, i] for i = i0 + row, j = j0 + 4 (tx & 15) .. + 3
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = c0 + 16 * q;
    const int64_t i = i0 + r, j = j0 + r4;
    if (i < M && j < N)
      *(f32x4*)(t + i * ldt + j) = f32x4{tile[r4][r], tile[r4 + 1][r], tile[r4 + 2][r],
                                         tile[r4 + 3][r]};
  }
  }
  mx = wave_max(mx);
  if ((tx & 63) == 0) sh[tx >> 6] = mx;
  __syncthreads();
  if (tx == 0) {
    const double rr = (double)fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
    atomicMax(amax, (unsigned long long)__double_as_longlong(rr));
  }
}

__global__ void f64_to_bf16_kernel(const double* __restrict__ s, int64_t M, int64_t lds,
                                   uint16_t* __restrict__ d, int64_t ldd, double scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= M) return;
  d[j * ldd + i] = (uint16_t)bf16_bits((float)(scale * s[j * lds + i]));
}

// max |x| over an M x N matrix: block (x, y) takes rows [AM_CH x, AM_CH (x+1))
// of the columns y, y + gridDim.y, ... (sixteen independent loads per thread
// and column), reduces in registers and LDS, and one atomic per block
// combines the result (non-negative doubles order like their bit patterns).
constexpr int AM_CH = 4096;
template <typename TS>
__global__ __launch_bounds__(256) void absmax_kernel(const TS* __restrict__ s, int64_t M, int64_t N,
                                                     int64_t lds, unsigned long long* out) {
  __shared__ double sh[4];
  double mx = 0.0;
  const int64_t i0 = (int64_t)blockIdx.x * AM_CH + threadIdx.x;
  for (int64_t j = blockIdx.y; j < N; j += gridDim.y) {
    const TS* col = s + j * lds;
    // All loads first, unconditional (a row past M reads row 0 and is masked
    // below): a conditional load converted on the spot waits for itself
    // before the next one is issued.
    TS x[AM_CH / 256];
#pragma unroll
    for (int u = 0; u < AM_CH / 256; ++u) {
      const int64_t i = i0 + 256 * u;
      x[u] = col[i < M ? i : 0];
    }
#pragma unroll
    for (int u = 0; u < AM_CH / 256; ++u)
      mx = fmax(mx, i0 + 256 * u < M ? fabs((double)x[u]) : 0.0);
  }
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    atomicMax(out, (unsigned long long)__double_as_longlong(r));
  }
}

__global__ void transpose_kernel(const double* __restrict__ s, int64_t M, int64_t N, int64_t lds,
                                 double* __restrict__ d, int64_t ldd) {
  __shared__ double tile[32][33];
  const int64_t i0 = (int64_t)blockIdx.x * 32, j0 = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 256 threads: 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int64_t i = i0 + tx, j = j0 + r;
    tile[r][tx] = (i < M && j < N) ? s[j * lds + i] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int64_t j = j0 + tx, i = i0 + r;   // d (N x M): d[j, i] at i*ldd + j
    if (i < M && j < N) d[i * ldd + j] = tile[tx][r];
  }
}

__global__ void fill_kernel(double* p, int64_t M, int64_t ld, double v, int ident) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= M) return;
  p[j * ld + i] = ident ? (i == j ? 1.0 : 0.0) : v;
}

inline dim3 grid2(int64_t M, int64_t N, int b = 256) {
  return dim3((unsigned)((M + b - 1) / b), (unsigned)N);
}

}  // namespace

hipError_t cholesky_upper(double* G, int64_t n, int64_t ld, double tau, int* flags, int* ndef,
                          hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (n <= 2 * CH_NB) {
    // Small: one work-group does everything.
    hipLaunchKernelGGL(cholesky_kernel, dim3(1), dim3(BIG), 0, st, G, (int)n, ld, tau, flags,
                       ndef);
    return hipGetLastError();
  }
  double* ws = nullptr;
  hipError_t e = ws_alloc((void**)&ws, sizeof(double) * (n + PB * PB), st);
  if (e != hipSuccess) return e;
  double* d0 = ws;
  double* Rinv = ws + n;
  hipLaunchKernelGGL(chol_init_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, G,
                     (int)n, ld, d0, flags, ndef);
  // Two levels: 64-row steps inside CH_OUTER-row panels.  A step updates only
  // the rest of its panel's rows; the rows below the panel get one rank-
  // CH_OUTER update per panel, which moves CH_OUTER / 64 times less trailing
  // matrix through HBM than rank-64 updates would.
  for (int64_t J0 = 0; J0 < n && e == hipSuccess; J0 += CH_OUTER) {
    const int64_t jend = std::min<int64_t>(n, J0 + CH_OUTER);
    for (int64_t J = J0; J < jend && e == hipSuccess; J += PB) {
      const int nb = (int)std::min<int64_t>(PB, jend - J);
      hipLaunchKernelGGL(potrf_wave_kernel, dim3(1), dim3(PB), 0, st, G, ld, (int)J, nb, tau, d0,
                         flags, ndef, Rinv);
      const int64_t w = n - J - nb;
      if (w <= 0) break;
      double* panel = G + (J + nb) * ld + J;         // rows J..J+nb, cols J+nb..n
      GemmArgs g;                                    // panel = R_JJ^{-T} panel (in place)
      g.M = nb; g.N = w; g.K = nb;
      g.A = Rinv; g.lda = PB; g.ta = T::F64; g.a_trans = true;
      g.B = panel; g.ldb = ld; g.tb = T::F64;
      g.C = panel; g.ldc = ld; g.tc = T::F64;
      e = gemm_f64(g, st);
      const int64_t rows = jend - (J + nb);          // the panel's rows still to factor
      if (e != hipSuccess || rows <= 0) continue;
      GemmArgs h;                                    // those rows -= panel^T panel
      h.M = rows; h.N = w; h.K = nb;
      h.A = panel; h.lda = ld; h.ta = T::F64; h.a_trans = true;
      h.B = panel; h.ldb = ld; h.tb = T::F64;
      h.C = G + (J + nb) * ld + (J + nb); h.ldc = ld; h.tc = T::F64;
      h.alpha = -1.0; h.beta = 1.0;
      h.tri = 1;                                     // only the upper triangle is read later
      e = gemm_f64(h, st);
    }
    const int64_t w2 = n - jend;
    if (e != hipSuccess || w2 <= 0) continue;
    double* P = G + jend * ld + J0;                  // R rows J0..jend, cols jend..n
    GemmArgs h;                                      // trailing -= P^T P
    h.M = w2; h.N = w2; h.K = jend - J0;
    h.A = P; h.lda = ld; h.ta = T::F64; h.a_trans = true;
    h.B = P; h.ldb = ld; h.tb = T::F64;
    h.C = G + jend * ld + jend; h.ldc = ld; h.tc = T::F64;
    h.alpha = -1.0; h.beta = 1.0;
    h.tri = 1;
    e = gemm_f64(h, st);
  }
  if (e == hipSuccess)
    hipLaunchKernelGGL(zero_lower_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256),
                       0, st, G, n, ld);
  ws_free(ws, st);
  return e != hipSuccess ? e : hipGetLastError();
}

hipError_t cholesky_inverse_upper(double* G, int64_t n, int64_t ld, double tau, int* flags,
                                  int* ndef, double* X, int64_t ldx, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  // chol_inv_kernel addresses G through a buffer resource with 32-bit byte
  // offsets: a G beyond 2 GB (a very large ld) takes the blocked path.
  if (n > CI_MAXN || ((n - 1) * ld + n) * (int64_t)sizeof(double) >= ((int64_t)1 << 31)) {
    hipError_t e = cholesky_upper(G, n, ld, tau, flags, ndef, st);
    return e != hipSuccess ? e : trinv_upper(G, n, ld, X, ldx, st);
  }
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)chol_inv_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)CI_LDS);
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)trinv_cols_kernel<4>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)XiCfg<4>::LDS);
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)trinv_cols_kernel<9>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)XiCfg<9>::LDS);
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)trinv_cols_kernel<15>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)XiCfg<15>::LDS);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(CI_THR), CI_LDS, st, G, (int)n, ld, tau, flags,
                     ndef, X, ldx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int nt = (int)((n + CB - 1) / CB);
  // K steps per slice: ceil((n - CB) / (4 XI_NS)) <= nt - 1
  static_assert((CI_MAXN - CB + 4 * XI_NS - 1) / (4 * XI_NS) <= 15, "trinv_cols_kernel<15>");
  if (nt > 10)
    hipLaunchKernelGGL(trinv_cols_kernel<15>, dim3(nt - 1), dim3(BIG), XiCfg<15>::LDS, st, G,
                       (int)n, ld, X, ldx);
  else if (nt > 5)
    hipLaunchKernelGGL(trinv_cols_kernel<9>, dim3(nt - 1), dim3(BIG), XiCfg<9>::LDS, st, G,
                       (int)n, ld, X, ldx);
  else if (nt > 1)
    hipLaunchKernelGGL(trinv_cols_kernel<4>, dim3(nt - 1), dim3(BIG), XiCfg<4>::LDS, st, G,
                       (int)n, ld, X, ldx);
  return hipGetLastError();
}

hipError_t trinv_upper(const double* R, int64_t n, int64_t ldr, double* X, int64_t ldx,
                       hipStream_t st) {
  // X = R^{-1}: diagonal blocks inverted in parallel, then block rows from the
  // bottom: X[I, >I] = -X_II (R[I, >I] X[>I, >I]) on the FP64 GEMM.
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, st,
                     X, n, ldx, 0.0, 0);
  const int nblk = (int)((n + PB - 1) / PB);
  hipLaunchKernelGGL(diag_inv_wave_kernel, dim3(nblk), dim3(PB), 0, st, R, (int)n, ldr, X, ldx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nblk == 1) return e;
  double* T = nullptr;
  e = ws_alloc((void**)&T, sizeof(double) * PB * n, st);
  if (e != hipSuccess) return e;
  for (int I = nblk - 2; I >= 0 && e == hipSuccess; --I) {
    const int64_t r0 = (int64_t)I * PB, c0 = r0 + PB, w = n - c0;
    GemmArgs g;                       // T = R[r0:r0+64, c0:] X[c0:, c0:]
    g.M = PB; g.N = w; g.K = w;
    g.A = R + c0 * ldr + r0; g.lda = ldr; g.ta = T::F64; g.a_trans = false;
    g.B = X + c0 * ldx + c0; g.ldb = ldx; g.tb = T::F64;
    g.C = T; g.ldc = PB; g.tc = T::F64;
    e = gemm_f64(g, st);
    if (e != hipSuccess) break;
    GemmArgs h;                       // X[r0:r0+64, c0:] = -X_II T
    h.M = PB; h.N = w; h.K = PB;
    h.A = X + r0 * ldx + r0; h.lda = ldx; h.ta = T::F64; h.a_trans = false;
    h.B = T; h.ldb = PB; h.tb = T::F64;
    h.C = X + c0 * ldx + r0; h.ldc = ldx; h.tc = T::F64;
    h.alpha = -1.0;
    e = gemm_f64(h, st);
  }
  ws_free(T, st);
  return e;
}

hipError_t complete_columns(void* Q, T tq, int64_t m, int64_t n, int64_t ldq, const int* flags,
                            hipStream_t st) {
  if (m <= 0 || n <= 0) return hipSuccess;
  const dim3 grid((unsigned)std::min<int64_t>((m + 255) / 256, 4), (unsigned)n);
  switch (tq) {
    case T::F64:
      hipLaunchKernelGGL(complete_kernel<double>, grid, dim3(256), 0, st, (double*)Q, m, n, ldq,
                         flags);
      break;
    case T::F32:
      hipLaunchKernelGGL(complete_kernel<float>, grid, dim3(256), 0, st, (float*)Q, m, n, ldq,
                         flags);
      break;
    case T::BF16:
      hipLaunchKernelGGL(complete_bf16_kernel, grid, dim3(256), 0, st, (uint16_t*)Q, m, n, ldq,
                         flags);
      break;
  }
  return hipGetLastError();
}

hipError_t jacobi_svd(double* W, int64_t n, int64_t ldw, double* V, int64_t ldv, double* s,
                      int* rot_count, int max_sweeps, double tol, hipStream_t st,
                      int* sweeps_done, int* hist) {
  (void)s;
  *sweeps_done = 0;
  if (n <= 1 || max_sweeps <= 0) return hipSuccess;
  const int nb = (int)((n + JB - 1) / JB);
  const int nbp = nb + (nb & 1);            // pad to even with a ghost block
  // Gram-based pair steps when W_C (odd stride) and four 32 x 33 matrices
  // fit in LDS; the column-rotation kernel otherwise.
  const int64_t ldl = (n + 15) / 16 * 16 + 1;   // as in jacobi_pair_gram
  const size_t lds_gram = ((size_t)ldl * JW + 4 * JG_MAT) * sizeof(double);
  const bool gram = lds_gram <= 150 * 1024;
  const size_t lds = gram ? lds_gram : (size_t)n * JW * sizeof(double);
  if (lds > 150 * 1024) return hipErrorInvalidValue;
  static bool attr = false;
  if (!attr) {
    const void* fns[] = {(const void*)jacobi_round_kernel, (const void*)jacobi_gram_round_kernel,
                         (const void*)jacobi_gram_sweep_kernel};
    for (const void* f : fns) {
      hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
      if (e != hipSuccess) return e;
    }
    attr = true;
  }
  // Circle-method schedule of block pairs, all rounds of one sweep.
  const int rounds = nbp - 1, per = nbp / 2;
  std::vector<int> hpairs((size_t)2 * rounds * per);
  for (int r = 0; r < rounds; ++r)
    for (int i = 0; i < per; ++i) {
      int a, b;
      if (i == 0) { a = r; b = nbp - 1; }
      else { a = (r + i) % (nbp - 1); b = (r - i + nbp - 1) % (nbp - 1); }
      hpairs[2 * (r * per + i)] = a;
      hpairs[2 * (r * per + i) + 1] = b;
    }
  // Device block: pairs, then ctl[4] (barrier, give-up flag, sweep count),
  // then the per-sweep rotation counts, then one W_C-read counter per pair
  // slot (all-sweeps kernel with S > 1).
  const size_t npair = hpairs.size();
  int* dblk = nullptr;
  std::vector<int> hout(4 + max_sweeps);     // ctl[4] + hist, read back
  hipError_t e = ws_alloc((void**)&dblk, sizeof(int) * (npair + 4 + max_sweeps + per), st);
  if (e != hipSuccess) return e;
  int* dpairs = dblk;
  unsigned* ctl = (unsigned*)(dblk + npair);
  int* dhist = dblk + npair + 4;
  e = hipMemcpyAsync(dpairs, hpairs.data(), sizeof(int) * npair, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(ctl, 0, sizeof(int) * (4 + max_sweeps + per), st);
  static int coop = -1;   // cooperative launch usable on this device
  static int cus = 0;
  if (coop < 0 && e == hipSuccess) {
    int dev = 0, attr_coop = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&attr_coop, hipDeviceAttributeCooperativeLaunch, dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    coop = attr_coop ? 1 : 0;
  }
  // Work-groups per pair in the all-sweeps kernel: each pair step's W/V
  // update (FP64 MFMA, 2 n x 32 x 32 per pair) is split S ways; the load,
  // Gram and rotations are repeated by each part.  One work-group per CU.
  int S = 4;
  while (S > 1 && per * S > cus) S /= 2;
  bool done = false;
  if (e == hipSuccess && gram && coop == 1) {
    // All sweeps in one launch.
    double* Wp = W; int nn = (int)n; int64_t ldw_ = ldw; double* Vp = V; int64_t ldv_ = ldv;
    const int* pp = dpairs; int rr = rounds, pe = per; double tl = tol; int* hp = dhist;
    int ms = max_sweeps, nb_ = nb; unsigned* cp = ctl; int sp = S;
    void* args[] = {&Wp, &nn, &ldw_, &Vp, &ldv_, &pp, &rr, &pe, &tl, &hp, &ms, &nb_, &cp, &sp};
    if (hipLaunchCooperativeKernel((const void*)jacobi_gram_sweep_kernel, dim3(per * S), dim3(BIG),
                                   args, (unsigned)lds, st) == hipSuccess) {
      e = d2h_small(st, hout.data(), ctl, sizeof(int) * (4 + max_sweeps));
      if (e == hipSuccess) {
        const int sw = (int)hout[2];
        *sweeps_done = sw;
        if (hist) for (int i = 0; i < sw && i < max_sweeps; ++i) hist[i] = hout[4 + i];
        if (hout[1]) e = hipErrorLaunchTimeOut;   // a barrier wait gave up
      }
      done = true;
    } else {
      (void)hipGetLastError();
      coop = 0;
    }
  }
  // Otherwise one launch per round, the convergence test on the host.
  for (int sw = 0; !done && sw < max_sweeps && e == hipSuccess; ++sw) {
    int* rc = dhist + sw;
    for (int r = 0; r < rounds && e == hipSuccess; ++r) {
      if (gram)
        hipLaunchKernelGGL(jacobi_gram_round_kernel, dim3(per), dim3(BIG), lds, st, W, (int)n,
                           ldw, V, ldv, dpairs + 2 * r * per, tol, rc, nb);
      else
        hipLaunchKernelGGL(jacobi_round_kernel, dim3(per), dim3(BIG), lds, st, W, (int)n, ldw,
                           V, ldv, dpairs + 2 * r * per, tol, rc, nb);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = d2h_small(st, hout.data(), rc, sizeof(int));
    if (e != hipSuccess) break;
    *sweeps_done = sw + 1;
    if (hist) hist[sw] = hout[0];
    if (hout[0] == 0) break;
  }
  ws_free(dblk, st);
  (void)hipStreamSynchronize(st);           // hpairs (host) was a copy source
  (void)rot_count;
  return e;
}

hipError_t svd_finalize(double* W, int64_t n, int64_t ldw, double* V, int64_t ldv, double* s,
                        double* Uout, int64_t ldu, double* Vout, int64_t ldvo, double* sout,
                        hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (n > SF_MAXN) return hipErrorInvalidValue;
  hipLaunchKernelGGL(svd_finalize_kernel, dim3(1), dim3(BIG), 0, st, W, (int)n, ldw, V, ldv, s,
                     Uout, ldu, Vout, ldvo, sout);
  return hipGetLastError();
}

// Single-launch unpivoted QR when the columns fit the work-groups' LDS and
// cooperative launches are available; false when not used (nothing launched).
static bool hh_qr_coop(double* A, int64_t m, int64_t nf, int64_t nu, int64_t lda, double* tau,
                       int64_t n, hipStream_t st, hipError_t* err) {
  static int cus = -1, coop = 0;
  if (cus < 0) {
    int dev = 0, c = 0, a = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipDeviceGetAttribute(&a, hipDeviceAttributeCooperativeLaunch, dev);
    cus = c > 0 ? c : 1;
    coop = a;
    if (coop && hipFuncSetAttribute((const void*)hh_qr_coop_kernel,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)HQ_LDS_MAX) != hipSuccess) {
      (void)hipGetLastError();
      coop = 0;
    }
  }
  if (!coop || nf <= 0 || nu > INT32_MAX) return false;
  // blocks of HQ_B columns per work-group: one hand-over between
  // work-groups per HQ_B steps; rows padded to a multiple of HQ_THR
  const int b = HQ_B;
  const int64_t nblk = (nu + b - 1) / b;
  const int64_t G = std::min<int64_t>(nblk, cus);
  const int64_t nmax = ((nblk + G - 1) / G) * b;
  const int64_t mp = (m + HQ_THR - 1) / HQ_THR * HQ_THR;
  const size_t lds = ((size_t)(nmax + 2) * mp + 2 * (HQ_THR / WAVE) * HQ_NC) * sizeof(double);
  if (lds > HQ_LDS_MAX || mp > INT32_MAX / 8) return false;
  unsigned* ctl = nullptr;
  hipError_t e = ws_alloc((void**)&ctl, sizeof(unsigned), st);
  if (e != hipSuccess) { *err = e; return true; }
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, st, ctl, (int64_t)1);
  double* Vh = tau + n;
  hipLaunchKernelGGL(hq_fill_vh_kernel,
                     dim3((unsigned)std::min<int64_t>((m * nf + 255) / 256, 4096)), dim3(256), 0,
                     st, Vh, m, nf);
  double* Ap = A; int m_ = (int)m, mp_ = (int)mp, nf_ = (int)nf, nu_ = (int)nu, b_ = b;
  int64_t lda_ = lda, ldv = m; double* tp = tau; unsigned* cp = ctl;
  void* args[] = {&Ap, &m_, &mp_, &nf_, &nu_, &b_, &lda_, &Vh, &ldv, &tp, &cp};
  e = hipLaunchCooperativeKernel((const void*)hh_qr_coop_kernel, dim3((unsigned)G),
                                 dim3(HQ_THR), args, (unsigned)lds, st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ws_free(ctl, st);
    return false;
  }
  unsigned gave_up = 0;
  e = d2h_small(st, &gave_up, ctl, sizeof(unsigned));
  ws_free(ctl, st);
  if (e == hipSuccess && gave_up) e = hipErrorLaunchTimeOut;   // a wait gave up
  *err = e;
  return true;
}

__global__ void identity_perm_kernel(int64_t* perm, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) perm[i] = i;
}

hipError_t qrcp_householder(double* A, int64_t m, int64_t n, int64_t lda, double* tau,
                            int64_t* perm, bool pivot, hipStream_t st, int64_t extra,
                            int64_t steps) {
  // Reflectors: tau[0..kmax) (1, or 0 for a zero column), unit v_k in the
  // panel at tau + n (m x kmax, leading dimension m, zeros above row k).
  // `extra` columns after the n factored ones receive every reflector
  // (unpivoted only).  steps >= 0 stops a pivoted factorisation after that
  // many reflectors; tau[steps..kmax) are then 0 (inactive).
  const int64_t kfull = m < n ? m : n;
  const int64_t kmax = steps >= 0 && steps < kfull ? steps : kfull;
  if (extra < 0 || (extra > 0 && pivot) || (kmax < kfull && !pivot)) return hipErrorInvalidValue;
  if (kmax < kfull) {
    hipError_t e = hipMemsetAsync(tau + kmax, 0, sizeof(double) * (kfull - kmax), st);
    if (e != hipSuccess) return e;
  }
  if (!pivot) {
    hipError_t e = hipSuccess;
    if (hh_qr_coop(A, m, kmax, n + extra, lda, tau, n, st, &e)) {
      if (e == hipSuccess && perm)
        hipLaunchKernelGGL(identity_perm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           st, perm, n);
      return e != hipSuccess ? e : hipGetLastError();
    }
  }
  const int64_t nu = n + extra;
  double *norms = nullptr, *vbuf = nullptr, *flag = nullptr;
  hipError_t e = ws_alloc((void**)&norms, sizeof(double) * (nu + m + 1), st);
  if (e != hipSuccess) return e;
  vbuf = norms + nu;
  flag = vbuf + m;
  hipLaunchKernelGGL(init_norms, dim3((unsigned)n), dim3(256), 0, st, A, m, n, lda, norms, perm);
  for (int64_t k = 0; k < kmax; ++k) {
    hipLaunchKernelGGL(hh_step1, dim3(1), dim3(BIG), 0, st, A, m, n, lda, k, norms, perm, vbuf,
                       flag, pivot ? 1 : 0, tau + n, m, tau);
    hipLaunchKernelGGL(hh_step2, dim3((unsigned)(nu - k)), dim3(256), 0, st, A, m, nu, lda, k,
                       vbuf, flag, norms);
  }
  e = hipGetLastError();
  ws_free(norms, st);
  return e;
}

hipError_t householder_form_q(const double* A, int64_t m, int64_t n, int64_t lda,
                              const double* tau, double* Q, int64_t ldq, int64_t qcols,
                              hipStream_t st) {
  (void)A; (void)lda;
  const int64_t kmax = m < n ? m : n;
  hipLaunchKernelGGL(fill_kernel, grid2(m, qcols), dim3(256), 0, st, Q, m, ldq, 0.0, 1);
  const double* Vh = tau + n;
  for (int64_t k = kmax - 1; k >= 0; --k)
    hipLaunchKernelGGL(hh_apply_q, dim3((unsigned)qcols), dim3(256), 0, st, Vh, m, m, k, tau, Q,
                       ldq);
  return hipGetLastError();
}

hipError_t convert(const void* src, T ts, int64_t M, int64_t N, int64_t lds, void* dst, T td,
                   int64_t ldd, double scale, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (ts == T::F32 && td == T::BF16) {
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)((M / 4 + 256) / 256), (unsigned)N),
                       dim3(256), 0, st, (const float*)src, M, lds, (uint16_t*)dst, ldd,
                       (float)scale);
  } else if (ts == T::F64 && td == T::BF16) {
    hipLaunchKernelGGL(f64_to_bf16_kernel, grid2(M, N), dim3(256), 0, st, (const double*)src, M,
                       lds, (uint16_t*)dst, ldd, scale);
  } else if (ts == T::F64 && td == T::F32) {
    hipLaunchKernelGGL((convert_kernel<double, float>), grid2(M, N), dim3(256), 0, st,
                       (const double*)src, M, lds, (float*)dst, ldd, scale);
  } else if (ts == T::F32 && td == T::F64) {
    hipLaunchKernelGGL((convert_kernel<float, double>), grid2(M, N), dim3(256), 0, st,
                       (const float*)src, M, lds, (double*)dst, ldd, scale);
  } else if (ts == T::F64 && td == T::F64) {
    hipLaunchKernelGGL((convert_kernel<double, double>), grid2(M, N), dim3(256), 0, st,
                       (const double*)src, M, lds, (double*)dst, ldd, scale);
  } else if (ts == T::F32 && td == T::F32) {
    hipLaunchKernelGGL((convert_kernel<float, float>), grid2(M, N), dim3(256), 0, st,
                       (const float*)src, M, lds, (float*)dst, ldd, scale);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t convert_bf16_absmax(const float* src, int64_t M, int64_t N, int64_t lds, uint16_t* dst,
                              int64_t ldd, double* amax, hipStream_t st) {
  if ((M & 3) || (lds & 3) || (ldd & 3)) {
    hipError_t e = convert(src, T::F32, M, N, lds, dst, T::BF16, ldd, 1.0, st);
    return e != hipSuccess ? e : absmax(src, T::F32, M, N, lds, amax, st);
  }
  hipError_t e = hipMemsetAsync(amax, 0, sizeof(double), st);
  if (e != hipSuccess || M <= 0 || N <= 0) return e;
  const int64_t blocks = std::min<int64_t>(4096, N);
  hipLaunchKernelGGL(f32_to_bf16_amax_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, M, N,
                     lds, dst, ldd, (unsigned long long*)amax);
  return hipGetLastError();
}

hipError_t convert_bf16_transpose_absmax(const float* src, int64_t M, int64_t N, int64_t lds,
                                         uint16_t* dst, int64_t ldd, float* tr, int64_t ldt,
                                         double* amax, hipStream_t st) {
  if ((M & 3) || (N & 3) || (lds & 3) || (ldd & 3) || (ldt & 3)) return hipErrorInvalidValue;
  hipError_t e = hipMemsetAsync(amax, 0, sizeof(double), st);
  if (e != hipSuccess || M <= 0 || N <= 0) return e;
  const int64_t tiles = ((M + 63) / 64) * ((N + 63) / 64);
  hipLaunchKernelGGL(f32_bf16_transpose_kernel, dim3((unsigned)std::min<int64_t>(tiles, 4096)),
                     dim3(256), 0, st, src, M, N, lds, dst, ldd, tr, ldt,
                     (unsigned long long*)amax);
  return hipGetLastError();
}

hipError_t absmax(const void* src, T ts, int64_t M, int64_t N, int64_t lds, double* out,
                  hipStream_t st) {
  hipError_t e = hipMemsetAsync(out, 0, sizeof(double), st);
  if (e != hipSuccess || M <= 0 || N <= 0) return e;
  const int64_t bx = (M + AM_CH - 1) / AM_CH;
  const int64_t by = std::max<int64_t>(1, std::min<int64_t>(N, 2048 / std::min<int64_t>(bx, 2048)));
  const dim3 grid((unsigned)bx, (unsigned)by);
  if (ts == T::F64)
    hipLaunchKernelGGL(absmax_kernel<double>, grid, dim3(256), 0, st, (const double*)src, M, N,
                       lds, (unsigned long long*)out);
  else if (ts == T::F32)
    hipLaunchKernelGGL(absmax_kernel<float>, grid, dim3(256), 0, st, (const float*)src, M, N,
                       lds, (unsigned long long*)out);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t transpose_f64(const double* src, int64_t M, int64_t N, int64_t lds, double* dst,
                         int64_t ldd, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)((M + 31) / 32), (unsigned)((N + 31) / 32)),
                     dim3(256), 0, st, src, M, N, lds, dst, ldd);
  return hipGetLastError();
}

hipError_t fill_f64(double* p, int64_t M, int64_t N, int64_t ld, double v, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, grid2(M, N), dim3(256), 0, st, p, M, ld, v, 0);
  return hipGetLastError();
}

hipError_t set_identity(double* p, int64_t M, int64_t N, int64_t ld, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, grid2(M, N), dim3(256), 0, st, p, M, ld, 0.0, 1);
  return hipGetLastError();
}

}  // namespace rna
