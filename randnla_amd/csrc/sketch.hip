// Sketch generation and application.
//
// Every sketch entry is a pure function of (stream, entry index), so tiles
// are produced independently, in registers, by whichever work-group needs
// them.  Dense sketches fused into a product are handled by the GEMM kernels
// (generated operand, gemm_f64.hip); this file holds the stand-alone panel
// generator, the sparse-sign application S * A, and the test-matrix
// generator.
#include <stdint.h>

#include "common.h"
#include "kernels.h"

namespace rna {
namespace {

template <typename TO>
__device__ __forceinline__ void put(TO* p, double v) { *p = (TO)v; }
template <>
__device__ __forceinline__ void put<uint16_t>(uint16_t* p, double v) {
  *p = (uint16_t)bf16_bits((float)v);
}

// Entry pairs (2p, 2p+1) share one generator block.
template <typename TO>
__global__ void panel_kernel(int dist, Stream s, int64_t rows, int64_t cols, TO* out, int64_t ld,
                             double scale) {
  const int64_t total = rows * cols;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (dist == 2) {
    // Rademacher: thread p covers the 128 entries of block p.
    const int64_t e0 = p * 128;
    if (e0 >= total) return;
    const u64x2 b = block64(s, (uint64_t)p);
    for (int q = 0; q < 128 && e0 + q < total; ++q) {
      const int64_t e = e0 + q;
      const uint64_t w = (q < 64) ? b.v[0] : b.v[1];
      const double v = ((w >> (q & 63)) & 1ull) ? 1.0 : -1.0;
      put(out + (e / rows) * ld + (e % rows), scale * v);
    }
    return;
  }
  const int64_t e0 = 2 * p;
  if (e0 >= total) return;
  double z0, z1;
  if (dist == 0) {
    gauss_pair(s, (uint64_t)p, z0, z1);
  } else {
    const u64x2 b = block64(s, (uint64_t)p);
    z0 = 2.0 * unit53(b.v[0]) - 1.0;
    z1 = 2.0 * unit53(b.v[1]) - 1.0;
  }
  put(out + (e0 / rows) * ld + (e0 % rows), scale * z0);
  if (e0 + 1 < total) put(out + ((e0 + 1) / rows) * ld + ((e0 + 1) % rows), scale * z1);
}

__global__ void words_kernel(Stream s, uint64_t first, int64_t n, uint64_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = word64(s, first + (uint64_t)i);
}

// Sparse-sign S (d x m) times A (m x n): work-group (column block cb,
// row chunk z) accumulates S[:, chunk] * A[chunk, cols] in LDS with FP64
// atomics, then writes its partial d x CB block.
constexpr int SS_THR = 256, SS_THR_R = 512;
constexpr int SS_LDS_DOUBLES = 16384;   // 128 KB

template <typename TA>
__global__ __launch_bounds__(SS_THR) void sparse_sign_kernel(Stream s, int64_t d, int zeta,
                                                             const TA* __restrict__ A, int64_t m,
                                                             int64_t n, int64_t lda, int cb,
                                                             int64_t chunk, double* part) {
  extern __shared__ double acc[];           // d x cb, column-major
  const int64_t c0 = (int64_t)blockIdx.x * cb;
  const int ncol = (int)min((int64_t)cb, n - c0);
  const int64_t j0 = (int64_t)blockIdx.y * chunk;
  const int64_t j1 = min(m, j0 + chunk);
  for (int64_t i = threadIdx.x; i < d * cb; i += SS_THR) acc[i] = 0.0;
  __syncthreads();
  for (int64_t j = j0 + threadIdx.x; j < j1; j += SS_THR) {
    int rows[SPARSE_MAX_ZETA];
    float sg[SPARSE_MAX_ZETA];
    const int z = sparse_sign_column(s, (uint64_t)j, (int)d, zeta, rows, sg);
    const double sc = 1.0 / sqrt((double)z);
    for (int c = 0; c < ncol; ++c) {
      const double a = (double)A[(c0 + c) * lda + j] * sc;
      for (int q = 0; q < z; ++q) atomicAdd(&acc[(int64_t)c * d + rows[q]], a * (double)sg[q]);
    }
  }
  __syncthreads();
  // part layout: [chunk z][column][row], ld = d, stride d * n per chunk.
  double* out = part + (int64_t)blockIdx.y * d * n;
  for (int64_t i = threadIdx.x; i < d * ncol; i += SS_THR) {
    const int64_t r = i % d, c = i / d;
    out[(c0 + c) * d + r] = acc[c * d + r];
  }
}

// The same with the column's rows and signs in registers (zeta <= ZMAX).
template <typename TA, int ZMAX>
__global__ __launch_bounds__(SS_THR_R) void sparse_sign_reg_kernel(Stream s, int64_t d, int zeta,
                                                                   const TA* __restrict__ A,
                                                                   int64_t m, int64_t n,
                                                                   int64_t lda, int cb,
                                                                   int64_t chunk, double* part) {
  extern __shared__ double acc[];           // d x cb, column-major
  const int64_t c0 = (int64_t)blockIdx.x * cb;
  const int ncol = (int)min((int64_t)cb, n - c0);
  const int64_t j0 = (int64_t)blockIdx.y * chunk;
  const int64_t j1 = min(m, j0 + chunk);
  for (int64_t i = threadIdx.x; i < d * cb; i += SS_THR_R) acc[i] = 0.0;
  __syncthreads();
  for (int64_t j = j0 + threadIdx.x; j < j1; j += SS_THR_R) {
    int rows[ZMAX];
    float sg[ZMAX];
    const int z = sparse_sign_column_fixed<ZMAX>(s, (uint64_t)j, (int)d, zeta, rows, sg);
    const double sc = 1.0 / sqrt((double)z);
    for (int c = 0; c < ncol; ++c) {
      const double a = (double)A[(c0 + c) * lda + j] * sc;
      double* ac = acc + (int64_t)c * d;
#pragma unroll
      for (int q = 0; q < ZMAX; ++q)
        if (q < z) atomicAdd(ac + rows[q], a * (double)sg[q]);
    }
  }
  __syncthreads();
  double* out = part + (int64_t)blockIdx.y * d * n;
  for (int64_t i = threadIdx.x; i < d * ncol; i += SS_THR_R) {
    const int64_t r = i % d, c = i / d;
    out[(c0 + c) * d + r] = acc[c * d + r];
  }
}

template <int ZMAX>
__global__ void sparse_sign_t_kernel(Stream s, int64_t d, int zeta, int64_t m,
                                     const double* __restrict__ y, double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  int rows[ZMAX];
  float sg[ZMAX];
  const int z = sparse_sign_column_fixed<ZMAX>(s, (uint64_t)j, (int)d, zeta, rows, sg);
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < ZMAX; ++q) {   // unconditional loads (q >= z reads y[rows[0]], weight 0)
    const double v = y[rows[q < z ? q : 0]];
    acc += (q < z ? (double)sg[q] : 0.0) * v;
  }
  out[j] = acc / sqrt((double)z);
}

// ---- Hadamard test matrix -------------------------------------------------
__device__ __forceinline__ double hsign(uint64_t r, uint64_t c) {
  return (__popcll(r & c) & 1) ? -1.0 : 1.0;
}
__device__ __forceinline__ double rsign(const Stream& s, uint64_t e) {
  const uint64_t w = word64(s, e >> 6);
  return ((w >> (e & 63)) & 1ull) ? 1.0 : -1.0;
}

__device__ __forceinline__ Stream substream(Stream s, uint64_t tag) {
  s.ctr[1] += tag;
  return s;
}

__global__ void hm_scatter(int64_t m, int64_t n, const double* sigma, Stream s, uint64_t a,
                           uint64_t b, double* work) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // singular index
  const int64_t c = blockIdx.y;                                        // column of A
  if (j >= n) return;
  const Stream sc = substream(s, 2);
  const uint64_t row = (a * (uint64_t)j + b) & (uint64_t)(m - 1);
  const double v = sigma[j] * hsign((uint64_t)c, (uint64_t)j) * rsign(sc, (uint64_t)j);
  work[c * m + (int64_t)row] = v;
}

__global__ void hm_butterfly(int64_t m, int64_t n, int64_t h, double* work) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // pair index < m/2
  const int64_t c = blockIdx.y;
  if (q >= m / 2) return;
  const int64_t r = ((q / h) * 2 * h) + (q % h);
  double* col = work + c * m;
  const double x = col[r], y = col[r + h];
  col[r] = x + y;
  col[r + h] = x - y;
}

// In-LDS butterflies for h = 1 .. 2^(LOGB-1) on blocks of 2^LOGB rows.
constexpr int HM_LOGB = 11;   // 2048 doubles = 16 KB
__global__ __launch_bounds__(256) void hm_local(int64_t m, int lg, double* work) {
  __shared__ double x[1 << HM_LOGB];
  const int B = 1 << lg;
  const int64_t blk = blockIdx.x, c = blockIdx.y;
  double* col = work + c * m + blk * B;
  for (int i = threadIdx.x; i < B; i += 256) x[i] = col[i];
  __syncthreads();
  for (int h = 1; h < B; h <<= 1) {
    for (int q = threadIdx.x; q < B / 2; q += 256) {
      const int r = ((q / h) * 2 * h) + (q % h);
      const double u = x[r], v = x[r + h];
      x[r] = u + v;
      x[r + h] = u - v;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < B; i += 256) col[i] = x[i];
}

template <typename TO>
__global__ void hm_finalize(int64_t m, int64_t n, const double* work, Stream s, double noise,
                            TO* out, int64_t ld) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y;
  if (r >= m) return;
  const Stream srow = substream(s, 1), snoise = substream(s, 4);
  const double scale = 1.0 / sqrt((double)m * (double)n);
  double v = rsign(srow, (uint64_t)r) * work[c * m + r] * scale;
  if (noise != 0.0) v += noise * gauss(snoise, (uint64_t)(r + c * m));
  put(out + c * ld + r, v);
}

}  // namespace

hipError_t sketch_panel(int dist, const Stream& s, int64_t rows, int64_t cols, void* out, T tout,
                        int64_t ld, double scale, hipStream_t st) {
  const int64_t total = rows * cols;
  if (total <= 0) return hipSuccess;
  const int64_t nthreads = (dist == 2) ? (total + 127) / 128 : (total + 1) / 2;
  const dim3 grid((unsigned)((nthreads + 255) / 256));
  switch (tout) {
    case T::F64:
      hipLaunchKernelGGL(panel_kernel<double>, grid, dim3(256), 0, st, dist, s, rows, cols,
                         (double*)out, ld, scale);
      break;
    case T::F32:
      hipLaunchKernelGGL(panel_kernel<float>, grid, dim3(256), 0, st, dist, s, rows, cols,
                         (float*)out, ld, scale);
      break;
    case T::BF16:
      hipLaunchKernelGGL(panel_kernel<uint16_t>, grid, dim3(256), 0, st, dist, s, rows, cols,
                         (uint16_t*)out, ld, scale);
      break;
  }
  return hipGetLastError();
}

hipError_t stream_words(const Stream& s, uint64_t first, int64_t n, uint64_t* out,
                        hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(words_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, first,
                     n, out);
  return hipGetLastError();
}

// zeta == 8 (every column has exactly 8 rows: d >= 8) with CB columns per
// work-group known at compile time: the row's CB values of A are loaded
// first (their latency overlaps the counter-based row generation), then the
// 8 x CB LDS atomics run without per-row branches.
template <typename TA, int CB>
__global__ __launch_bounds__(SS_THR_R) void sparse_sign_z8_kernel(Stream s, int64_t d,
                                                                  const TA* __restrict__ A,
                                                                  int64_t m, int64_t n,
                                                                  int64_t lda, int64_t chunk,
                                                                  double* part) {
  extern __shared__ double acc[];           // d x CB, column-major
  const int64_t c0 = (int64_t)blockIdx.x * CB;
  const int ncol = (int)min((int64_t)CB, n - c0);
  const int64_t j0 = (int64_t)blockIdx.y * chunk;
  const int64_t j1 = min(m, j0 + chunk);
  for (int64_t i = threadIdx.x; i < d * CB; i += SS_THR_R) acc[i] = 0.0;
  __syncthreads();
  const double sc = 0.35355339059327373;     // 1 / sqrt(8)
  const TA* Ac = A + c0 * lda;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += SS_THR_R) {
    double a[CB];
    // unconditional loads (a column past ncol reads column 0 and is scaled
    // to zero), issued back to back
#pragma unroll
    for (int c = 0; c < CB; ++c) a[c] = (double)Ac[(c < ncol ? c : 0) * lda + j];
#pragma unroll
    for (int c = 0; c < CB; ++c) a[c] *= c < ncol ? sc : 0.0;
    int rows[8];
    float sg[8];
    (void)sparse_sign_column_fixed<8>(s, (uint64_t)j, (int)d, 8, rows, sg);
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      double* ac = acc + c * d;
#pragma unroll
      for (int q = 0; q < 8; ++q) atomicAdd(ac + rows[q], a[c] * (double)sg[q]);
    }
  }
  __syncthreads();
  double* out = part + (int64_t)blockIdx.y * d * n;
  for (int64_t i = threadIdx.x; i < d * ncol; i += SS_THR_R) {
    const int64_t r = i % d, c = i / d;
    out[(c0 + c) * d + r] = acc[c * d + r];
  }
}

template <typename TA>
bool launch_z8(int cb, dim3 grid, size_t lds, hipStream_t st, const Stream& s, int64_t d,
               const TA* A, int64_t m, int64_t n, int64_t lda, int64_t chunk, double* work) {
  switch (cb) {
    case 16: hipLaunchKernelGGL((sparse_sign_z8_kernel<TA, 16>), grid, dim3(SS_THR_R), lds, st, s, d, A, m, n, lda, chunk, work); return true;
    case 8: hipLaunchKernelGGL((sparse_sign_z8_kernel<TA, 8>), grid, dim3(SS_THR_R), lds, st, s, d, A, m, n, lda, chunk, work); return true;
    case 4: hipLaunchKernelGGL((sparse_sign_z8_kernel<TA, 4>), grid, dim3(SS_THR_R), lds, st, s, d, A, m, n, lda, chunk, work); return true;
    default: return false;
  }
}

// Work-group shape of the sparse-sign product; at most max_chunks row chunks
// (each writes a d x n partial slab).
static void ss_shape(int64_t d, int64_t m, int64_t n, int& cb, int64_t& nchunks, int64_t& chunk,
                     int64_t max_chunks = INT64_MAX) {
  cb = (int)(SS_LDS_DOUBLES / d);
  if (cb > 16) cb = 16;
  if (cb > n) cb = (int)n;
  if (cb < 1) cb = 1;
  const int64_t cblocks = (n + cb - 1) / cb;
  // Enough work-groups to cover the chip; each chunk at least 4096 rows.
  nchunks = (512 + cblocks - 1) / cblocks;
  const int64_t maxchunks = (m + 4095) / 4096;
  if (nchunks > maxchunks) nchunks = maxchunks;
  if (nchunks > max_chunks) nchunks = max_chunks;
  if (nchunks < 1) nchunks = 1;
  chunk = (m + nchunks - 1) / nchunks;
  nchunks = (m + chunk - 1) / chunk;
}

int64_t sparse_sign_work(int64_t d, int64_t m, int64_t n) {
  int cb;
  int64_t nchunks, chunk;
  ss_shape(d, m, n, cb, nchunks, chunk);
  return nchunks * d * n;
}

hipError_t sparse_sign_apply(const Stream& s, int64_t d, int zeta, const void* A, T ta, int64_t m,
                             int64_t n, int64_t lda, double* out, int64_t ldo, double* work,
                             int64_t work_elems, hipStream_t st) {
  if (d <= 0 || n <= 0) return hipSuccess;
  if (d > SS_LDS_DOUBLES || zeta < 1 || zeta > SPARSE_MAX_ZETA) return hipErrorInvalidValue;
  // A workspace smaller than the default shape needs means fewer row chunks
  // (at least one slab is needed).
  if (work_elems < d * n) return hipErrorInvalidValue;
  int cb;
  int64_t nchunks, chunk;
  ss_shape(d, m, n, cb, nchunks, chunk, work_elems / (d * n));
  if (work_elems < nchunks * d * n) return hipErrorInvalidValue;
  static bool attr = false;
  if (!attr) {
    const void* ks[] = {(const void*)sparse_sign_kernel<double>,
                        (const void*)sparse_sign_kernel<float>,
                        (const void*)sparse_sign_reg_kernel<double, 8>,
                        (const void*)sparse_sign_reg_kernel<float, 8>,
                        (const void*)sparse_sign_reg_kernel<double, 16>,
                        (const void*)sparse_sign_reg_kernel<float, 16>};
    for (const void* k : ks)
      (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, SS_LDS_DOUBLES * 8);
    attr = true;
  }
  const dim3 grid((unsigned)((n + cb - 1) / cb), (unsigned)nchunks);
  const size_t lds = (size_t)d * cb * sizeof(double);
  if (ta != T::F64 && ta != T::F32) return hipErrorInvalidValue;
  const bool f64 = ta == T::F64;
  // zeta == 8 with a full column block width: the specialised kernel
  if (zeta == 8 && d >= 8 && (cb == 4 || cb == 8 || cb == 16)) {
    static bool attr_z8 = false;
    if (!attr_z8) {
      const void* kz[] = {(const void*)sparse_sign_z8_kernel<double, 16>,
                          (const void*)sparse_sign_z8_kernel<double, 8>,
                          (const void*)sparse_sign_z8_kernel<double, 4>,
                          (const void*)sparse_sign_z8_kernel<float, 16>,
                          (const void*)sparse_sign_z8_kernel<float, 8>,
                          (const void*)sparse_sign_z8_kernel<float, 4>};
      for (const void* k : kz)
        (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, SS_LDS_DOUBLES * 8);
      attr_z8 = true;
    }
    if (f64) launch_z8<double>(cb, grid, lds, st, s, d, (const double*)A, m, n, lda, chunk, work);
    else launch_z8<float>(cb, grid, lds, st, s, d, (const float*)A, m, n, lda, chunk, work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return reduce_partials_f64(work, d, n, d, d * n, (int)nchunks, out, ldo, 1.0, st);
  }
  // rows and signs in registers for zeta <= 16 (the common 1..16 range)
  if (zeta <= 8) {
    if (f64)
      hipLaunchKernelGGL((sparse_sign_reg_kernel<double, 8>), grid, dim3(SS_THR_R), lds, st, s, d,
                         zeta, (const double*)A, m, n, lda, cb, chunk, work);
    else
      hipLaunchKernelGGL((sparse_sign_reg_kernel<float, 8>), grid, dim3(SS_THR_R), lds, st, s, d,
                         zeta, (const float*)A, m, n, lda, cb, chunk, work);
  } else if (zeta <= 16) {
    if (f64)
      hipLaunchKernelGGL((sparse_sign_reg_kernel<double, 16>), grid, dim3(SS_THR_R), lds, st, s,
                         d, zeta, (const double*)A, m, n, lda, cb, chunk, work);
    else
      hipLaunchKernelGGL((sparse_sign_reg_kernel<float, 16>), grid, dim3(SS_THR_R), lds, st, s,
                         d, zeta, (const float*)A, m, n, lda, cb, chunk, work);
  } else if (f64) {
    hipLaunchKernelGGL(sparse_sign_kernel<double>, grid, dim3(SS_THR), lds, st, s, d, zeta,
                       (const double*)A, m, n, lda, cb, chunk, work);
  } else {
    hipLaunchKernelGGL(sparse_sign_kernel<float>, grid, dim3(SS_THR), lds, st, s, d, zeta,
                       (const float*)A, m, n, lda, cb, chunk, work);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return reduce_partials_f64(work, d, n, d, d * n, (int)nchunks, out, ldo, 1.0, st);
}

hipError_t sparse_sign_transpose(const Stream& s, int64_t d, int zeta, int64_t m,
                                 const double* y, double* out, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  const dim3 grid((unsigned)((m + 255) / 256));
  if (zeta <= 8)
    hipLaunchKernelGGL((sparse_sign_t_kernel<8>), grid, dim3(256), 0, st, s, d, zeta, m, y, out);
  else if (zeta <= 16)
    hipLaunchKernelGGL((sparse_sign_t_kernel<16>), grid, dim3(256), 0, st, s, d, zeta, m, y, out);
  else
    hipLaunchKernelGGL((sparse_sign_t_kernel<SPARSE_MAX_ZETA>), grid, dim3(256), 0, st, s, d, zeta,
                       m, y, out);
  return hipGetLastError();
}

hipError_t hadamard_test_matrix(int64_t m, int64_t n, const double* sigma_dev, double noise,
                                const Stream& s, double* work, void* out, T tout, int64_t ld,
                                hipStream_t st) {
  if ((m & (m - 1)) || (n & (n - 1)) || n > m || n <= 0) return hipErrorInvalidValue;
  // Column map J(j) = (a j + b) mod m, a odd: the first two words of sub-stream 3.
  uint64_t a = 1, b = 0;
  {
    Stream sj = s;
    sj.ctr[1] += 3;
    const u64x2 w = block64(sj, 0);
    a = (w.v[0] | 1ull) & (uint64_t)(m - 1);
    if (a == 0) a = 1;
    b = w.v[1] & (uint64_t)(m - 1);
  }
  hipError_t e = hipMemsetAsync(work, 0, sizeof(double) * m * n, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(hm_scatter, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, st,
                     m, n, sigma_dev, s, a, b, work);
  int lg = 0;
  while ((1ll << lg) < m) ++lg;
  const int lb = lg < HM_LOGB ? lg : HM_LOGB;
  if (lb > 0)
    hipLaunchKernelGGL(hm_local, dim3((unsigned)(m >> lb), (unsigned)n), dim3(256), 0, st, m, lb,
                       work);
  for (int64_t h = (int64_t)1 << lb; h < m; h <<= 1)
    hipLaunchKernelGGL(hm_butterfly, dim3((unsigned)((m / 2 + 255) / 256), (unsigned)n),
                       dim3(256), 0, st, m, n, h, work);
  const dim3 g((unsigned)((m + 255) / 256), (unsigned)n);
  if (tout == T::F64)
    hipLaunchKernelGGL(hm_finalize<double>, g, dim3(256), 0, st, m, n, work, s, noise,
                       (double*)out, ld);
  else if (tout == T::F32)
    hipLaunchKernelGGL(hm_finalize<float>, g, dim3(256), 0, st, m, n, work, s, noise,
                       (float*)out, ld);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace rna
