// Memory-bound vector kernels of the iterative least-squares solvers
// (GEMV with A and A^T, triangular matrix-vector products, dots, axpby).
//
// A is column-major FP64.  y = A x streams A once with one thread per row
// (each column read is a coalesced 2 KB line per 256 rows); y = A^T x uses
// one work-group per column and a block reduction.  Both are HBM-bound:
// bytes per call = 8 m n.
#include "common.h"
#include "kernels.h"

namespace rna {
namespace {

constexpr int TPB = 256;

__global__ __launch_bounds__(TPB) void gemv_n_kernel(const double* __restrict__ A, int64_t m,
                                                     int64_t n, int64_t lda,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y, double alpha,
                                                     double beta) {
  __shared__ double xs[1024];
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int64_t c0 = 0; c0 < n; c0 += 1024) {
    const int64_t cn = min((int64_t)1024, n - c0);
    __syncthreads();
    for (int64_t c = threadIdx.x; c < cn; c += TPB) xs[c] = x[c0 + c];
    __syncthreads();
    if (i < m) {
      const double* a = A + c0 * lda + i;
      int64_t c = 0;
      for (; c + 3 < cn; c += 4) {
        s0 += a[c * lda] * xs[c];
        s1 += a[(c + 1) * lda] * xs[c + 1];
        s2 += a[(c + 2) * lda] * xs[c + 2];
        s3 += a[(c + 3) * lda] * xs[c + 3];
      }
      for (; c < cn; ++c) s0 += a[c * lda] * xs[c];
    }
  }
  if (i < m) {
    const double r = alpha * ((s0 + s1) + (s2 + s3));
    y[i] = (beta == 0.0) ? r : r + beta * y[i];
  }
}

// Partial dot products of columns with x over a row chunk; part[z * n + c].
__global__ __launch_bounds__(TPB) void gemv_t_kernel(const double* __restrict__ A, int64_t m,
                                                     int64_t n, int64_t lda,
                                                     const double* __restrict__ x, int64_t chunk,
                                                     double* __restrict__ part) {
  __shared__ double sh[TPB / 64];
  const int64_t c = blockIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * chunk, r1 = min(m, r0 + chunk);
  const double* a = A + c * lda;
  double s = 0.0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += TPB) s += a[i] * x[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.y * n + c] = s;
}

__global__ void gemv_t_finish(const double* __restrict__ part, int nz, int64_t n,
                              double* __restrict__ y, double alpha, double beta) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double s = 0.0;
  for (int z = 0; z < nz; ++z) s += part[z * n + c];
  y[c] = (beta == 0.0) ? alpha * s : alpha * s + beta * y[c];
}

// y = R x or R^T x, R upper triangular (one work-group per output row
// block of 64; the product is tiny next to the GEMVs with A).
__global__ __launch_bounds__(TPB) void trmv_kernel(const double* __restrict__ R, int64_t n,
                                                   int64_t ldr, int trans,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ y) {
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * (TPB / 64) + wv;   // output row
  if (i >= n) return;
  double s = 0.0;
  if (!trans) {
    for (int64_t c = i + ln; c < n; c += 64) s += R[c * ldr + i] * x[c];
  } else {
    for (int64_t c = ln; c <= i; c += 64) s += R[i * ldr + c] * x[c];
  }
  s = wave_sum(s);
  if (ln == 0) y[i] = s;
}

__global__ __launch_bounds__(TPB) void dot_kernel(const double* __restrict__ x,
                                                  const double* __restrict__ y, int64_t n,
                                                  double* __restrict__ part) {
  __shared__ double sh[TPB / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    s += x[i] * y[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void dot_finish(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) *out = s;
}

__global__ void axpby_kernel(int64_t n, double a, const double* __restrict__ x, double b,
                             double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

}  // namespace

hipError_t gemv_n(const double* A, int64_t m, int64_t n, int64_t lda, const double* x, double* y,
                  double alpha, double beta, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(gemv_n_kernel, dim3((unsigned)((m + TPB - 1) / TPB)), dim3(TPB), 0, st, A, m,
                     n, lda, x, y, alpha, beta);
  return hipGetLastError();
}

// work needs 64 * n doubles.
hipError_t gemv_t(const double* A, int64_t m, int64_t n, int64_t lda, const double* x, double* y,
                  double alpha, double beta, double* work, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int nz = (int)((1024 + n - 1) / n);
  if (nz > 64) nz = 64;
  if (nz < 1) nz = 1;
  int64_t chunk = (m + nz - 1) / nz;
  if (chunk < 1) chunk = 1;
  nz = (int)((m + chunk - 1) / chunk);
  if (nz < 1) nz = 1;
  hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)n, (unsigned)nz), dim3(TPB), 0, st, A, m, n,
                     lda, x, chunk, work);
  hipLaunchKernelGGL(gemv_t_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, work, nz,
                     n, y, alpha, beta);
  return hipGetLastError();
}

hipError_t trmv_upper(const double* R, int64_t n, int64_t ldr, bool trans, const double* x,
                      double* y, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(trmv_kernel, dim3((unsigned)((n + 3) / 4)), dim3(TPB), 0, st, R, n, ldr,
                     trans ? 1 : 0, x, y);
  return hipGetLastError();
}

// work needs 256 doubles.
hipError_t dot(const double* x, const double* y, int64_t n, double* out, double* work,
               hipStream_t st) {
  int nb = (int)((n + TPB - 1) / TPB);
  if (nb > 256) nb = 256;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(TPB), 0, st, x, y, n, work);
  hipLaunchKernelGGL(dot_finish, dim3(1), dim3(256), 0, st, work, nb, out);
  return hipGetLastError();
}

hipError_t axpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, x,
                     b, y);
  return hipGetLastError();
}

}  // namespace rna
