// FP64 GEMM on gfx950 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Work-group tile 128 x 64, K step 16, 256 threads = 4 waves in a 2 x 2
// arrangement; each wave owns a 64 x 32 block = 4 x 2 MFMA tiles (32 FP64
// accumulators per lane).  Operands are staged through LDS in the layout of
// their memory (double buffered): an M-contiguous A as [k][row] planes, a
// K-contiguous A (op(A) = A^T) and B as [row][k] strips of LDK doubles, so
// that every store writes whole 16-byte pairs along its contiguous
// dimension.  LDK = 18 spreads the 16 rows of an MFMA operand read over
// distinct bank pairs.
//
// Operands may be FP32 (promoted exactly on load), which lets the FP64 path
// and the FP64 orthonormalisation read the FP32 data of the mixed path
// without a conversion pass.
#include "common.h"
#include "kernels.h"

namespace rna {
namespace {

constexpr int BM = 128, BN = 64, BK = 16;
constexpr int PADM = BM + 16;   // [k][row] plane stride: 128 B mod 256 B
constexpr int LDK = BK + 2;     // [row][k] strip stride
constexpr int NTHR = 256;

template <typename TI>
__device__ __forceinline__ double ld_as_f64(const TI* p) { return (double)(*p); }
template <>
__device__ __forceinline__ double ld_as_f64<uint16_t>(const uint16_t* p) {
  return (double)bf16_to_f32(*p);
}

// Operand element: from memory (G == 0) or generated from the stream
// (1 Gaussian, 2 Uniform(-1,1), 3 Rademacher) at entry index e.
template <int G, typename TI>
__device__ __forceinline__ double fetch(const TI* p, int64_t off, const Stream& s, uint64_t e) {
  if constexpr (G == 0) return ld_as_f64(p + off);
  else if constexpr (G == 1) return gauss(s, e);
  else if constexpr (G == 2) return uniform_pm1(s, e);
  else return rademacher(s, e);
}

template <typename TO>
__device__ __forceinline__ void st_from_f64(TO* p, double v) { *p = (TO)v; }
template <>
__device__ __forceinline__ void st_from_f64<uint16_t>(uint16_t* p, double v) {
  *p = (uint16_t)bf16_bits((float)v);
}

template <bool AT, typename TA, typename TB, typename TC, int GA, int GB>
__global__ __launch_bounds__(NTHR) void gemm_f64_kernel(
    int64_t M, int64_t N, int64_t K, const TA* __restrict__ A, int64_t lda,
    const TB* __restrict__ B, int64_t ldb, TC* __restrict__ C, int64_t ldc,
    int64_t cstride, double alpha, double beta, int64_t kchunk, Stream gs, int tri) {
  __shared__ double As[2][AT ? BM * LDK : BK * PADM];
  __shared__ double Bs[2][BN * LDK];

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t m0 = (int64_t)blockIdx.x * BM;
  const int64_t n0 = (int64_t)blockIdx.y * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  int64_t kend = min(K, kbeg + kchunk);
  // tri 1: only the upper triangle of C is wanted; tiles strictly below the
  // diagonal skip the product and store zeros.  tri 2: B is upper
  // triangular, so rows k >= n0 + BN of B are zero and are not read.
  if (tri == 1 && m0 >= n0 + BN) kend = kbeg;
  if (tri == 2) kend = min(kend, n0 + BN);
  C += (int64_t)blockIdx.z * cstride;

  // Per-thread load assignments.
  // A, !AT (M-contiguous): 2 consecutive rows x 4 k.  AT (K-contiguous):
  // 2 consecutive k x 4 rows.  B (K-contiguous): 2 consecutive k x 2 cols.
  double ra[8], rb[4];

  auto load = [&](int64_t k0) {
    if (!AT) {
      const int mm = (t & 63) * 2;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int kk = (t >> 6) + 4 * i;
        const int64_t k = k0 + kk;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int64_t m = m0 + mm + e;
          ra[2 * i + e] = (m < M && k < kend)
                                 ? fetch<GA>(A, k * lda + m, gs, (uint64_t)(m + k * M))
                                 : 0.0;
        }
      }
    } else {
      const int kk = (t & 7) * 2;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t m = m0 + (t >> 3) + 32 * i;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int64_t k = k0 + kk + e;
          ra[2 * i + e] = (m < M && k < kend)
                                 ? fetch<GA>(A, m * lda + k, gs, (uint64_t)(k + m * K))
                                 : 0.0;
        }
      }
    }
    {
      const int kk = (t & 7) * 2;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int64_t n = n0 + (t >> 3) + 32 * i;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int64_t k = k0 + kk + e;
          rb[2 * i + e] = (n < N && k < kend)
                                 ? fetch<GB>(B, n * ldb + k, gs, (uint64_t)(k + n * K))
                                 : 0.0;
        }
      }
    }
  };
  auto store = [&](int buf) {
    if (!AT) {
      const int mm = (t & 63) * 2;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int kk = (t >> 6) + 4 * i;
        *(double2*)&As[buf][kk * PADM + mm] = double2{ra[2 * i], ra[2 * i + 1]};
      }
    } else {
      const int kk = (t & 7) * 2;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int mm = (t >> 3) + 32 * i;
        *(double2*)&As[buf][mm * LDK + kk] = double2{ra[2 * i], ra[2 * i + 1]};
      }
    }
    const int kk = (t & 7) * 2;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int nn = (t >> 3) + 32 * i;
      *(double2*)&Bs[buf][nn * LDK + kk] = double2{rb[2 * i], rb[2 * i + 1]};
    }
  };

  f64x4 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  const int64_t nk = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;
  if (nk > 0) {
    load(kbeg);
    store(0);
  }
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int buf = (int)(kt & 1);
    if (kt + 1 < nk) load(kbeg + (kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      double a[4], b[2];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = wm * 64 + i * 16 + (lane & 15);
        a[i] = AT ? As[buf][row * LDK + kr] : As[buf][kr * PADM + row];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[buf][(wn * 32 + j * 16 + (lane & 15)) * LDK + kr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }

  // Epilogue: accumulator v of a lane holds row 4*v + lane/16, column
  // lane % 16 of its 16 x 16 tile (v_mfma_f64_16x16x4_f64 D layout).
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t n = n0 + wn * 32 + j * 16 + (lane & 15);
      if (n >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t m = m0 + wm * 64 + i * 16 + 4 * v + (lane >> 4);
        if (m >= M) continue;
        double r = alpha * acc[i][j][v];
        TC* p = C + n * ldc + m;
        if (beta != 0.0) r += beta * ld_as_f64(p);
        st_from_f64(p, r);
      }
    }
}

template <bool AT, typename TA, typename TB, typename TC, int GA = 0, int GB = 0>
hipError_t launch4(const GemmArgs& g, hipStream_t st, const Stream& gs = Stream{}) {
  const int64_t mt = (g.M + BM - 1) / BM, nt = (g.N + BN - 1) / BN;
  int splits = g.splits < 1 ? 1 : g.splits;
  int64_t kchunk = (g.K + splits - 1) / splits;
  kchunk = ((kchunk + BK - 1) / BK) * BK;
  if (kchunk == 0) kchunk = BK;
  dim3 grid((unsigned)mt, (unsigned)nt, (unsigned)splits);
  hipLaunchKernelGGL((gemm_f64_kernel<AT, TA, TB, TC, GA, GB>), grid, dim3(NTHR), 0, st, g.M,
                     g.N, g.K, (const TA*)g.A, g.lda, (const TB*)g.B, g.ldb, (TC*)g.C, g.ldc,
                     g.cstride, g.alpha, g.beta, kchunk, gs, g.tri);
  return hipGetLastError();
}

template <bool AT, typename TA, typename TB>
hipError_t launch3(const GemmArgs& g, hipStream_t st) {
  switch (g.tc) {
    case T::F64: return launch4<AT, TA, TB, double>(g, st);
    case T::F32: return launch4<AT, TA, TB, float>(g, st);
    case T::BF16: return launch4<AT, TA, TB, uint16_t>(g, st);
  }
  return hipErrorInvalidValue;
}

template <bool AT, typename TA>
hipError_t launch2(const GemmArgs& g, hipStream_t st) {
  switch (g.tb) {
    case T::F64: return launch3<AT, TA, double>(g, st);
    case T::F32: return launch3<AT, TA, float>(g, st);
    default: return hipErrorInvalidValue;
  }
}

template <bool AT>
hipError_t launch1(const GemmArgs& g, hipStream_t st) {
  switch (g.ta) {
    case T::F64: return launch2<AT, double>(g, st);
    case T::F32: return launch2<AT, float>(g, st);
    default: return hipErrorInvalidValue;
  }
}

template <int G, typename TUser>
hipError_t launch_gen(int side, const GemmArgs& g, const Stream& s, hipStream_t st) {
  if (side == 0) return launch4<false, double, TUser, double, G, 0>(g, st, s);
  return launch4<false, TUser, double, double, 0, G>(g, st, s);
}

}  // namespace

int64_t dense_sketch_work(int side, int64_t m, int64_t n, int64_t d) {
  if (side != 0) return 0;
  const int64_t tiles = ((d + BM - 1) / BM) * ((n + BN - 1) / BN);
  int64_t splits = (512 + tiles - 1) / tiles;
  const int64_t maxs = (m + 1023) / 1024;
  if (splits > maxs) splits = maxs;
  if (splits < 1) splits = 1;
  return splits > 1 ? splits * d * n : 0;
}

hipError_t dense_sketch_apply(int side, int dist, const Stream& s, const void* A, T ta, int64_t m,
                              int64_t n, int64_t lda, int64_t d, double* out, int64_t ldo,
                              double* work, int64_t work_elems, hipStream_t st) {
  if (dist < 0 || dist > 2) return hipErrorInvalidValue;
  if (ta != T::F64 && ta != T::F32) return hipErrorInvalidValue;
  GemmArgs g;
  if (side == 0) {            // out (d x n) = S (d x m) * A (m x n)
    g.M = d; g.N = n; g.K = m;
    g.A = nullptr; g.lda = d; g.ta = T::F64; g.a_trans = false;
    g.B = A; g.ldb = lda; g.tb = ta;
    // Split-K partial slabs in `work`; a workspace smaller than the default
    // shape needs means fewer splits (one split writes out directly).
    const int64_t need = dense_sketch_work(side, m, n, d);
    int64_t splits = need > 0 ? need / (d * n) : 1;
    if (work_elems < need) splits = work_elems / (d * n);
    if (splits > 1) {
      g.splits = (int)splits;
      g.C = work; g.ldc = d; g.cstride = d * n;
    } else {
      g.C = out; g.ldc = ldo;
    }
  } else {                    // out (m x d) = A (m x n) * Omega (n x d)
    g.M = m; g.N = d; g.K = n;
    g.A = A; g.lda = lda; g.ta = ta; g.a_trans = false;
    g.B = nullptr; g.ldb = n; g.tb = T::F64;
    g.C = out; g.ldc = ldo;
  }
  g.tc = T::F64;
  hipError_t e;
  const bool f32 = (ta == T::F32);
  switch (dist) {
    case 0: e = f32 ? launch_gen<1, float>(side, g, s, st) : launch_gen<1, double>(side, g, s, st); break;
    case 1: e = f32 ? launch_gen<2, float>(side, g, s, st) : launch_gen<2, double>(side, g, s, st); break;
    default: e = f32 ? launch_gen<3, float>(side, g, s, st) : launch_gen<3, double>(side, g, s, st); break;
  }
  if (e != hipSuccess || g.splits <= 1) return e;
  return reduce_partials_f64(work, d, n, d, d * n, g.splits, out, ldo, 1.0, st);
}

hipError_t dense_sketch_transpose(int dist, const Stream& s, int64_t d, int64_t m,
                                  const double* y, double* out, hipStream_t st) {
  // out (m x 1) = S^T y, S (d x m) generated: op(A) = A^T with A = S (K = d, M = m).
  GemmArgs g;
  g.M = m; g.N = 1; g.K = d;
  g.A = nullptr; g.lda = d; g.ta = T::F64; g.a_trans = true;
  g.B = y; g.ldb = d; g.tb = T::F64;
  g.C = out; g.ldc = m; g.tc = T::F64;
  switch (dist) {
    case 0: return launch4<true, double, double, double, 1, 0>(g, st, s);
    case 1: return launch4<true, double, double, double, 2, 0>(g, st, s);
    case 2: return launch4<true, double, double, double, 3, 0>(g, st, s);
  }
  return hipErrorInvalidValue;
}

hipError_t gemm_f64(const GemmArgs& g, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  if (g.splits > 1 && g.beta != 0.0) return hipErrorInvalidValue;
  if (g.K <= 0) {
    // C = beta * C
    GemmArgs h = g;
    h.K = 0;
    return g.a_trans ? launch1<true>(h, st) : launch1<false>(h, st);
  }
  return g.a_trans ? launch1<true>(g, st) : launch1<false>(g, st);
}

}  // namespace rna
