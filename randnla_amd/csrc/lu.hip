// Full-pivot LU of a dense m x n FP64 matrix, returning the unit lower
// factor L (m x min(m, n)), the reference's Stabilizer
// (src/lora_helpers.rs:144-146: X.clone().full_piv_lu().l(), nalgebra 0.33.0
// FullPivLU::new).  Step k:
//   pivot   = the first maximum of |x| over the trailing block X[k:, k:],
//             scanned column by column (ties keep the earlier column, then
//             the earlier row); a zero maximum ends the factorisation (the
//             remaining columns of L are unit vectors);
//   swaps   = columns k <-> cp and rows k <-> rp over the whole matrix (the
//             already computed part of L moves with its row);
//   column  = X[k+1:, k] *= 1 / pivot (nalgebra's gauss_step scales by the
//             reciprocal);
//   update  = X[k+1:, k+1:] -= X[k+1:, k] X[k, k+1:].
// Each step is two launches: one work-group reduces the partial maxima,
// swaps and scales; then a 2-D grid applies the rank-1 update and finds the
// next step's partial maxima in the same pass over the trailing block.  The
// matrix is HBM-resident; the steps are latency-bound for the sizes of the
// path (min(m, n) <= 320).
#include <limits.h>

#include "common.h"
#include "kernels.h"

namespace rna {
namespace {

constexpr int LU_ROWS = 256;   // rows per update work-group (one per thread)
constexpr int LU_COLS = 64;    // columns per update work-group
constexpr int LU_PIV = 1024;   // threads of the pivot work-group

// Candidate order: larger |x| first; equal |x| -> smaller key, where
// key = c * m + r is the column-major position (the first maximum).
__device__ __forceinline__ bool lu_better(double v, int64_t key, double bv, int64_t bkey) {
  return v > bv || (v == bv && key < bkey);
}

// Work-group reduction of (v, key); the result is valid in thread 0.
template <int NT>
__device__ __forceinline__ void lu_block_best(double& v, int64_t& key, double* sv, int64_t* sk) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(v, o, WAVE);
    const int64_t ok = __shfl_xor(key, o, WAVE);
    if (lu_better(ov, ok, v, key)) { v = ov; key = ok; }
  }
  constexpr int NW = NT / WAVE;
  const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  if (l == 0) { sv[w] = v; sk[w] = key; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < NW; ++i)
      if (lu_better(sv[i], sk[i], v, key)) { v = sv[i]; key = sk[i]; }
  }
}

// Rank-1 update of step k (rows and columns > k) and the partial maxima of
// the updated block, one (value, key) per work-group; k < 0 only searches
// the whole matrix.  Thread t of block (bx, by) owns row k+1 + 256 bx + t and
// columns k+1 + 64 by .. +63, walked in increasing order (so its first
// maximum is the column-major first).
__global__ __launch_bounds__(LU_ROWS) void lu_update_kernel(double* X, int64_t m, int64_t n,
                                                            int64_t ld, int64_t k,
                                                            const int* done, double* pv,
                                                            int64_t* pk) {
  if (*done) return;
  __shared__ double sv[LU_ROWS / WAVE];
  __shared__ int64_t sk[LU_ROWS / WAVE];
  const int64_t r = k + 1 + (int64_t)blockIdx.x * LU_ROWS + threadIdx.x;
  const int64_t cb = k + 1 + (int64_t)blockIdx.y * LU_COLS;
  const int64_t ce = cb + LU_COLS < n ? cb + LU_COLS : n;
  double best = -1.0;
  int64_t bkey = LLONG_MAX;
  if (r < m) {
    const double lr = k >= 0 ? X[r + k * ld] : 0.0;
    for (int64_t c = cb; c < ce; ++c) {
      double x = X[r + c * ld];
      if (k >= 0) {
        x = x - lr * X[k + c * ld];
        X[r + c * ld] = x;
      }
      const double a = fabs(x);
      if (a > best) { best = a; bkey = c * m + r; }
    }
  }
  lu_block_best<LU_ROWS>(best, bkey, sv, sk);
  if (threadIdx.x == 0) {
    const int64_t p = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    pv[p] = best;
    pk[p] = bkey;
  }
}

// Pivot of step k from np partial maxima: swaps, then the multipliers.
// A zero (or absent) maximum sets *done and leaves X unchanged.
__global__ __launch_bounds__(LU_PIV) void lu_pivot_kernel(double* X, int64_t m, int64_t n,
                                                          int64_t ld, int64_t k, int64_t np,
                                                          const double* pv, const int64_t* pk,
                                                          int* done) {
  if (*done) return;
  __shared__ double sv[LU_PIV / WAVE];
  __shared__ int64_t sk[LU_PIV / WAVE];
  __shared__ double s_best;
  __shared__ int64_t s_key;
  double best = -1.0;
  int64_t bkey = LLONG_MAX;
  for (int64_t i = threadIdx.x; i < np; i += LU_PIV)
    if (lu_better(pv[i], pk[i], best, bkey)) { best = pv[i]; bkey = pk[i]; }
  lu_block_best<LU_PIV>(best, bkey, sv, sk);
  if (threadIdx.x == 0) { s_best = best; s_key = bkey; }
  __syncthreads();
  if (!(s_best > 0.0)) {
    if (threadIdx.x == 0) *done = 1;
    return;
  }
  const int64_t cp = s_key / m, rp = s_key % m;
  const double piv = X[rp + cp * ld];
  __syncthreads();
  if (cp != k)
    for (int64_t r = threadIdx.x; r < m; r += LU_PIV) {
      const double a = X[r + k * ld], b = X[r + cp * ld];
      X[r + k * ld] = b;
      X[r + cp * ld] = a;
    }
  __syncthreads();
  if (rp != k)
    for (int64_t c = threadIdx.x; c < n; c += LU_PIV) {
      const double a = X[k + c * ld], b = X[rp + c * ld];
      X[k + c * ld] = b;
      X[rp + c * ld] = a;
    }
  __syncthreads();
  const double inv = 1.0 / piv;
  for (int64_t r = k + 1 + threadIdx.x; r < m; r += LU_PIV) X[r + k * ld] *= inv;
}

// L (m x kl): strictly lower part of the factored X, unit diagonal.
__global__ void lu_extract_l_kernel(const double* X, int64_t m, int64_t kl, int64_t ld, double* L,
                                    int64_t ldl) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y;
  if (r >= m || c >= kl) return;
  L[r + c * ldl] = r > c ? X[r + c * ld] : (r == c ? 1.0 : 0.0);
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace

int64_t full_piv_lu_work(int64_t m, int64_t n) {
  // done flag (padded to 16 bytes) + partial values + partial keys
  return 2 + 2 * cdiv(m, LU_ROWS) * cdiv(n, LU_COLS);
}

hipError_t full_piv_lu_l(double* X, int64_t m, int64_t n, int64_t ld, double* L, int64_t ldl,
                         double* work, hipStream_t st) {
  if (m <= 0 || n <= 0) return hipSuccess;
  const int64_t kmin = m < n ? m : n;
  const int64_t maxp = cdiv(m, LU_ROWS) * cdiv(n, LU_COLS);
  int* done = (int*)work;
  double* pv = work + 2;
  int64_t* pk = (int64_t*)(pv + maxp);
  hipError_t e = hipMemsetAsync(done, 0, sizeof(int), st);
  if (e != hipSuccess) return e;
  // step k's pivot reads the partials of the update (or search) before it
  for (int64_t k = 0; k < kmin; ++k) {
    const int64_t kk = k - 1;   // update of step k - 1 (none before step 0)
    const int64_t gx = cdiv(m - kk - 1, LU_ROWS), gy = cdiv(n - kk - 1, LU_COLS);
    hipLaunchKernelGGL(lu_update_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(LU_ROWS), 0, st,
                       X, m, n, ld, kk, (const int*)done, pv, pk);
    hipLaunchKernelGGL(lu_pivot_kernel, dim3(1), dim3(LU_PIV), 0, st, X, m, n, ld, k, gx * gy,
                       (const double*)pv, (const int64_t*)pk, done);
  }
  // the last step's update (it changes no entry of L, but keeps X = L U)
  if (kmin < m && kmin < n)
    hipLaunchKernelGGL(lu_update_kernel,
                       dim3((unsigned)cdiv(m - kmin, LU_ROWS), (unsigned)cdiv(n - kmin, LU_COLS)),
                       dim3(LU_ROWS), 0, st, X, m, n, ld, kmin - 1, (const int*)done, pv, pk);
  hipLaunchKernelGGL(lu_extract_l_kernel, dim3((unsigned)cdiv(m, 256), (unsigned)kmin), dim3(256),
                     0, st, (const double*)X, m, kmin, ld, L, ldl);
  return hipGetLastError();
}

}  // namespace rna
