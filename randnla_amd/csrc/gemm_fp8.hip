// FP8-e4m3 GEMM with per-128-block scales on gfx950 matrix cores.
//
// C (FP32, M x N) = alpha * sum_k  sA(i, k) A8[i, k] * sB(k, j) B8[k, j]
// where the FP8 operands are stored K-contiguous (row i of A8 and column j of
// B8 are contiguous in k) and the scales are powers of two (E8M0 bytes,
// 2^(e - 127)) constant over 128-long K blocks:
//   sA(i, k) = SA[(i / 128) * sa_r + (k / 128) * sa_k]
//   sB(k, j) = SB[j * sb_j + (k / 128)]
// The products run on v_mfma_scale_f32_32x32x64_f8f6f4, which multiplies each
// lane's 32 K-values of A and B by the lane's E8M0 scales in hardware, so the
// per-block scaling costs nothing in the inner loop.
//
// Work-group tile 128 x 320, 512 threads (8 waves as 4 x 2, a wave owns
// 1 x 5 tiles of 32 x 32), K step 128 (one scale block), operands staged in
// LDS as [row][128 bytes] with 16-byte chunks XOR-swizzled by row, double
// buffered (2 x 56 KB).  Split-K over grid.z writes partial slabs.
//
// Quantisation kernels make the operands: A (FP32, column-major) -> a
// row-major and a column-major FP8 copy with one scale per 128 x 128 block
// (the two copies serve A X and A^T X with K-contiguous loads); a panel X
// (FP32, K x N column-major) -> FP8 with one scale per column and 128-K block.
#include "common.h"
#include "kernels.h"

namespace rna {
namespace {

constexpr int Q_NTHR = 512;
constexpr int Q_BM = 128;
constexpr int Q_NT = 5;
constexpr int Q_BN = 64 * Q_NT;
constexpr int Q_BK = 128;
constexpr int Q_ROWB = 128;                       // bytes per LDS row
constexpr int Q_ABYTES = Q_BM * Q_ROWB, Q_BBYTES = Q_BN * Q_ROWB;
constexpr int Q_STAGE = Q_ABYTES + Q_BBYTES;      // 56 KB
constexpr int FP8_MAX = 448;                      // largest finite e4m3 value

typedef int v8i __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int qswz(int r, int c) { return (c ^ (r & 7)) * 16; }

// E8M0 exponent e (scale 2^(e-127)) so that absmax / scale <= 448.
__device__ __forceinline__ int e8m0_for(float amax) {
  if (!(amax > 0.f)) return 127;
  int e = (int)ceilf(log2f(amax / (float)FP8_MAX));
  // guard against rounding of the division and log2: the smallest e with
  // amax * 2^-e <= 448 (ldexpf is exact here)
  while (ldexpf(amax, -e) > (float)FP8_MAX) ++e;
  while (e > -127 && ldexpf(amax, 1 - e) <= (float)FP8_MAX) --e;
  return max(-127, min(127, e)) + 127;
}

__device__ __forceinline__ uint32_t pack4_fp8(float a, float b, float c, float d) {
  int w = 0;
  w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, w, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
  return (uint32_t)w;
}

// D: every tile is full (M, N multiples of the tile, K of 128, 16-byte
// aligned rows): both operands go from global memory straight into LDS
// (global_load_lds_dwordx4, lane l of a wave lands at base + 16 l), with no
// register staging and no bounds tests.  Instruction j of a tile fills rows
// 8j..8j+7; lane l takes row 8j + l/8 and the logical chunk whose swizzled
// slot is l % 8.
template <bool D>
__global__ __launch_bounds__(Q_NTHR, 1) void gemm_fp8_kernel(
    int64_t M, int64_t N, int64_t K, const uint8_t* __restrict__ A, int64_t lda,
    const uint8_t* __restrict__ SA, int64_t sa_r, int64_t sa_k, const uint8_t* __restrict__ B,
    int64_t ldb, const uint8_t* __restrict__ SB, int64_t sb_j, float* __restrict__ C, int64_t ldc,
    int64_t cstride, float alpha, int64_t kchunk) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 3, wn = wave >> 2;
  const int64_t m0 = (int64_t)blockIdx.x * Q_BM, n0 = (int64_t)blockIdx.y * Q_BN;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = min(K, kbeg + kchunk);
  C += (int64_t)blockIdx.z * cstride;
  constexpr int ACH = Q_BM * 8 / Q_NTHR;   // 2 chunks of 16 B per thread
  constexpr int BCH = Q_BN * 8 / Q_NTHR;   // 5
  u32x4v ra[ACH], rb[BCH];
  auto load_lds = [&](int64_t k0, int stage) {
    char* sA = smem + stage * Q_STAGE;
    char* sB = sA + Q_ABYTES;
#pragma unroll
    for (int i = 0; i < ACH; ++i) {
      const int j = i * (Q_NTHR / 64) + wave;
      const int r = 8 * j + (lane >> 3), c = (lane & 7) ^ (r & 7);
      __builtin_amdgcn_global_load_lds((const void*)(A + (m0 + r) * lda + k0 + 16 * c),
                                       (__attribute__((address_space(3))) void*)(sA + j * 1024),
                                       16, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < BCH; ++i) {
      const int j = i * (Q_NTHR / 64) + wave;
      const int r = 8 * j + (lane >> 3), c = (lane & 7) ^ (r & 7);
      __builtin_amdgcn_global_load_lds((const void*)(B + (n0 + r) * ldb + k0 + 16 * c),
                                       (__attribute__((address_space(3))) void*)(sB + j * 1024),
                                       16, 0, 0);
    }
  };
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < ACH; ++i) {
      const int idx = t + Q_NTHR * i, r = idx >> 3, c = idx & 7;
      const int64_t m = m0 + r, k = k0 + 16 * c;
      ra[i] = (m < M && k < kend) ? *(const u32x4v*)(A + m * lda + k) : u32x4v{0, 0, 0, 0};
    }
#pragma unroll
    for (int i = 0; i < BCH; ++i) {
      const int idx = t + Q_NTHR * i, r = idx >> 3, c = idx & 7;
      const int64_t n = n0 + r, k = k0 + 16 * c;
      rb[i] = (n < N && k < kend) ? *(const u32x4v*)(B + n * ldb + k) : u32x4v{0, 0, 0, 0};
    }
  };
  auto store = [&](int stage) {
    char* sA = smem + stage * Q_STAGE;
    char* sB = sA + Q_ABYTES;
#pragma unroll
    for (int i = 0; i < ACH; ++i) {
      const int idx = t + Q_NTHR * i, r = idx >> 3, c = idx & 7;
      *(u32x4v*)(sA + r * Q_ROWB + qswz(r, c)) = ra[i];
    }
#pragma unroll
    for (int i = 0; i < BCH; ++i) {
      const int idx = t + Q_NTHR * i, r = idx >> 3, c = idx & 7;
      *(u32x4v*)(sB + r * Q_ROWB + qswz(r, c)) = rb[i];
    }
  };
  f32x16 acc[1][Q_NT];
#pragma unroll
  for (int i = 0; i < 1; ++i)
#pragma unroll
    for (int j = 0; j < Q_NT; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

  const int lr = lane & 31, lk = lane >> 5;
  // scale rows/columns owned by this lane
  const int64_t arow = min(M - 1, m0 + wm * 32 + lr);
  int bcol[Q_NT];
#pragma unroll
  for (int ni = 0; ni < Q_NT; ++ni) bcol[ni] = (int)min(N - 1, n0 + wn * 32 * Q_NT + ni * 32 + lr);

  const int64_t nk = (kend > kbeg) ? (kend - kbeg + Q_BK - 1) / Q_BK : 0;
  if (nk > 0) {
    if constexpr (D) {
      load_lds(kbeg, 0);
      __builtin_amdgcn_s_waitcnt(0);
    } else {
      load(kbeg);
      store(0);
    }
  }
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int stage = (int)(kt & 1);
    const int64_t k0 = kbeg + kt * Q_BK;
    if (kt + 1 < nk) {
      if constexpr (D) load_lds(k0 + Q_BK, stage ^ 1);
      else load(k0 + Q_BK);
    }
    const int64_t kb = k0 / Q_BK;
    const int sa = SA[(arow / Q_BK) * sa_r + kb * sa_k];
    int sb[Q_NT];
#pragma unroll
    for (int ni = 0; ni < Q_NT; ++ni) sb[ni] = SB[(int64_t)bcol[ni] * sb_j + kb];
    const char* sA = smem + stage * Q_STAGE;
    const char* sB = sA + Q_ABYTES;
#pragma unroll
    for (int s = 0; s < Q_BK / 64; ++s) {
      // lane holds K = 64 s + 32 lk .. +31 of its row: chunks c, c + 1
      const int c = 4 * s + 2 * lk;
      const int arw = wm * 32 + lr;
      const u32x4v x0 = *(const u32x4v*)(sA + arw * Q_ROWB + qswz(arw, c));
      const u32x4v x1 = *(const u32x4v*)(sA + arw * Q_ROWB + qswz(arw, c + 1));
      const v8i a = v8i{(int)x0[0], (int)x0[1], (int)x0[2], (int)x0[3],
                        (int)x1[0], (int)x1[1], (int)x1[2], (int)x1[3]};
#pragma unroll
      for (int ni = 0; ni < Q_NT; ++ni) {
        const int row = wn * 32 * Q_NT + ni * 32 + lr;
        const u32x4v y0 = *(const u32x4v*)(sB + row * Q_ROWB + qswz(row, c));
        const u32x4v y1 = *(const u32x4v*)(sB + row * Q_ROWB + qswz(row, c + 1));
        const v8i b = v8i{(int)y0[0], (int)y0[1], (int)y0[2], (int)y0[3],
                          (int)y1[0], (int)y1[1], (int)y1[2], (int)y1[3]};
        acc[0][ni] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc[0][ni], 0, 0, 0, sa,
                                                                     0, sb[ni]);
      }
    }
    if constexpr (D) __builtin_amdgcn_s_waitcnt(0);
    else if (kt + 1 < nk) store(stage ^ 1);
    __syncthreads();
  }
  // Epilogue: lane holds rows 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32.
  {
    const int mi = 0;
#pragma unroll
    for (int ni = 0; ni < Q_NT; ++ni) {
      const int64_t n = n0 + wn * 32 * Q_NT + ni * 32 + lr;
      if (n >= N) continue;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int64_t m = m0 + wm * 32 + 8 * g + 4 * lk;
        float* p = C + n * ldc + m;
        const f32x4 v = {alpha * acc[mi][ni][4 * g], alpha * acc[mi][ni][4 * g + 1],
                         alpha * acc[mi][ni][4 * g + 2], alpha * acc[mi][ni][4 * g + 3]};
        if (m + 3 < M && ((ldc & 3) == 0)) {
          *(f32x4*)p = v;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (m + e < M) p[e] = v[e];
        }
      }
    }
  }
}

// One work-group per 128 x 128 block of A (FP32, column-major, m x n, lda):
// block absmax -> E8M0 scale into S[(i/128) * nb + j/128]; the scaled block
// is written column-major into Acm (ld ldcm; zeros past m and n up to the
// 128-multiples).  The row-major copy is made from Acm by the transpose
// kernel below.
__global__ __launch_bounds__(256) void quantize_a_kernel(const float* __restrict__ A, int64_t m,
                                                         int64_t n, int64_t lda,
                                                         uint8_t* __restrict__ Acm, int64_t ldcm,
                                                         uint8_t* __restrict__ S, int64_t nb) {
  __shared__ float red[4];
  const int t = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * 128, j0 = (int64_t)blockIdx.y * 128;
  float amax = 0.f;
  for (int e = t; e < 128 * 128; e += 256) {
    const int r = e & 127, c = e >> 7;
    const int64_t i = i0 + r, j = j0 + c;
    if (i < m && j < n) amax = fmaxf(amax, fabsf(A[j * lda + i]));
  }
  amax = wave_max(amax);
  if ((t & 63) == 0) red[t >> 6] = amax;
  __syncthreads();
  const float bm = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  const int eb = e8m0_for(bm);
  const float inv = ldexpf(1.f, 127 - eb);
  if (t == 0) S[(i0 / 128) * nb + j0 / 128] = (uint8_t)eb;
  // thread: 4 consecutive rows of one column -> one packed word
  for (int e = t; e < 32 * 128; e += 256) {
    const int rq = e & 31, c = e >> 5;
    const int64_t j = j0 + c;
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t i = i0 + 4 * rq + q;
      v[q] = (i < m && j < n) ? A[j * lda + i] * inv : 0.f;
    }
    *(uint32_t*)(Acm + j * ldcm + i0 + 4 * rq) = pack4_fp8(v[0], v[1], v[2], v[3]);
  }
}

// Row-major copy (mp rows of np bytes, ld ldr) of the column-major FP8 copy
// (np columns of mp bytes, ld ldcm) through LDS, 64 x 64 byte tiles.
__global__ __launch_bounds__(256) void quantize_a_rowmajor_kernel(
    const uint8_t* __restrict__ Acm, int64_t ldcm, int64_t mp, int64_t np,
    uint8_t* __restrict__ Arm, int64_t ldr) {
  __shared__ uint8_t T[64][65];
  const int t = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * 64, j0 = (int64_t)blockIdx.y * 64;
  for (int e = t; e < 64 * 64; e += 256) {
    const int r = e & 63, c = e >> 6;
    T[c][r] = (i0 + r < mp && j0 + c < np) ? Acm[(j0 + c) * ldcm + i0 + r] : 0;
  }
  __syncthreads();
  for (int e = t; e < 64 * 64; e += 256) {
    const int c = e & 63, r = e >> 6;
    if (i0 + r < mp && j0 + c < np) Arm[(i0 + r) * ldr + j0 + c] = T[c][r];
  }
}

// Panel X (FP32, K x N, ldx) -> X8 (column-major, ld ldq, K padded to a
// multiple of 128 with zeros) and S[j * kb + k / 128].  One work-group of 128
// threads per (128-K block, column).
__global__ __launch_bounds__(128) void quantize_panel_kernel(const float* __restrict__ X,
                                                             int64_t K, int64_t ldx,
                                                             uint8_t* __restrict__ X8, int64_t ldq,
                                                             uint8_t* __restrict__ S, int64_t kb) {
  __shared__ float red[2];
  const int t = threadIdx.x;
  const int64_t k0 = (int64_t)blockIdx.x * 128, j = blockIdx.y;
  const int64_t k = k0 + t;
  const float x = k < K ? X[j * ldx + k] : 0.f;
  float amax = wave_max(fabsf(x));
  if ((t & 63) == 0) red[t >> 6] = amax;
  __syncthreads();
  const int eb = e8m0_for(fmaxf(red[0], red[1]));
  const float inv = ldexpf(1.f, 127 - eb);
  if (t == 0) S[j * kb + blockIdx.x] = (uint8_t)eb;
  // 4 values per packed word: threads 0..31 write one word each
  const float v = x * inv;
  const float v1 = __shfl_down(v, 1, 64), v2 = __shfl_down(v, 2, 64), v3 = __shfl_down(v, 3, 64);
  if ((t & 3) == 0) *(uint32_t*)(X8 + j * ldq + k) = pack4_fp8(v, v1, v2, v3);
}

}  // namespace

hipError_t quantize_fp8_a(const float* A, int64_t m, int64_t n, int64_t lda, uint8_t* Arm,
                          int64_t ldr, uint8_t* Acm, int64_t ldcm, uint8_t* S, hipStream_t st) {
  const int64_t mb = (m + 127) / 128, nb = (n + 127) / 128;
  hipLaunchKernelGGL(quantize_a_kernel, dim3((unsigned)mb, (unsigned)nb), dim3(256), 0, st, A, m,
                     n, lda, Acm, ldcm, S, nb);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t mp = mb * 128, np = nb * 128;
  hipLaunchKernelGGL(quantize_a_rowmajor_kernel, dim3((unsigned)(mp / 64), (unsigned)(np / 64)),
                     dim3(256), 0, st, Acm, ldcm, mp, np, Arm, ldr);
  return hipGetLastError();
}

hipError_t quantize_fp8_panel(const float* X, int64_t K, int64_t N, int64_t ldx, uint8_t* X8,
                              int64_t ldq, uint8_t* S, hipStream_t st) {
  const int64_t kb = (K + 127) / 128;
  hipLaunchKernelGGL(quantize_panel_kernel, dim3((unsigned)kb, (unsigned)N), dim3(128), 0, st, X,
                     K, ldx, X8, ldq, S, kb);
  return hipGetLastError();
}

hipError_t gemm_fp8(const Fp8GemmArgs& g, hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)gemm_fp8_kernel<false>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 2 * Q_STAGE);
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)gemm_fp8_kernel<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 2 * Q_STAGE);
    if (e != hipSuccess) return e;
    attr = true;
  }
  if ((g.lda % 16) || (g.ldb % 16)) return hipErrorInvalidValue;
  const int64_t mt = (g.M + Q_BM - 1) / Q_BM, nt = (g.N + Q_BN - 1) / Q_BN;
  const int splits = g.splits < 1 ? 1 : g.splits;
  int64_t kchunk = (g.K + splits - 1) / splits;
  kchunk = ((kchunk + Q_BK - 1) / Q_BK) * Q_BK;
  if (kchunk == 0) kchunk = Q_BK;
  const bool full = g.M % Q_BM == 0 && g.N % Q_BN == 0 && g.K % Q_BK == 0 &&
                    ((uintptr_t)g.A & 15) == 0 && ((uintptr_t)g.B & 15) == 0;
  if (full)
    hipLaunchKernelGGL(gemm_fp8_kernel<true>, dim3((unsigned)mt, (unsigned)nt, (unsigned)splits),
                       dim3(Q_NTHR), 2 * Q_STAGE, st, g.M, g.N, g.K, g.A, g.lda, g.SA, g.sa_r,
                       g.sa_k, g.B, g.ldb, g.SB, g.sb_j, g.C, g.ldc, g.cstride, g.alpha, kchunk);
  else
    hipLaunchKernelGGL(gemm_fp8_kernel<false>, dim3((unsigned)mt, (unsigned)nt, (unsigned)splits),
                       dim3(Q_NTHR), 2 * Q_STAGE, st, g.M, g.N, g.K, g.A, g.lda, g.SA, g.sa_r,
                       g.sa_k, g.B, g.ldb, g.SB, g.sb_j, g.C, g.ldc, g.cstride, g.alpha, kchunk);
  return hipGetLastError();
}

}  // namespace rna
