// Low-precision GEMMs on gfx950 matrix cores.
//
//   mode 0  BF16 x BF16 -> FP32     v_mfma_f32_32x32x16_bf16, K step 64
//   mode 1  FP32 x FP32 split into FP16 hi/lo pairs (x*s = hi + lo), three
//           v_mfma_f32_32x32x16_f16 per tile (hi*hi + hi*lo + lo*hi),
//           about 22 significant bits per product, K step 32
//   mode 2  op(A) split into FP16 hi/lo, B rounded once to FP16: two
//           products per tile (hi*b + lo*b).  A keeps ~22 bits; B's rounding
//           (2^-12 relative) is a perturbation of the panel alone, which a
//           subspace iteration absorbs (it changes the iterate, not the
//           operator), K step 32
//
// Work-group tile 256 x (64*NT), 512 threads = 8 waves arranged 4 (M) x 2 (N);
// a wave owns 64 x (32*NT) = 2 x NT MFMA tiles (NT <= 5, i.e. up to 160 FP32
// accumulators per lane).  With N = 320 one work-group spans all of N, so each
// element of op(A) is read from HBM exactly once per product.
//
// Operand tiles are staged through LDS as K-contiguous rows ([row][k]) in a
// double buffer.  Rows are XOR-swizzled in 16-byte chunks so that the 16-byte
// fragment reads of a wave hit distinct banks.  An M-contiguous A (the
// product A * X with column-major A) is transposed on its way into LDS:
// each thread loads 8 rows x KQ columns and writes 8 row pieces.
#include "common.h"
#include "kernels.h"

namespace rna {
namespace {

constexpr int NTHR = 512;
constexpr int BM = 256;

template <int MODE>
struct Cfg {
  static constexpr int BK = (MODE == 0) ? 64 : 32;
  static constexpr int ROWB = BK * 2;          // bytes per LDS row (one plane)
  static constexpr int CH = ROWB / 16;         // 16-byte chunks per row
  static constexpr int SW = (MODE == 0) ? 1 : 2;
  static constexpr int PLA = (MODE == 0 || MODE >= 3) ? 1 : 2;   // LDS planes of A (hi, lo)
  static constexpr int PLB = (MODE == 1 || MODE == 4) ? 2 : 1;   // LDS planes of B
  static constexpr int KQ = BK / 16;           // k per thread for M-contig A
};

template <int MODE>
__device__ __forceinline__ int swz(int r, int c) {
  return (c ^ ((r >> Cfg<MODE>::SW) & (Cfg<MODE>::CH - 1))) * 16;
}

__device__ __forceinline__ void split_f16(float x, float s, uint32_t& hi, uint32_t& lo) {
  const float xs = x * s;
  const _Float16 h = (_Float16)xs;
  const float r = xs - (float)h;
  hi = (uint32_t)__builtin_bit_cast(uint16_t, h);
  lo = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)r);
}

template <int MODE>
__device__ __forceinline__ f32x16 mma(u32x4v a, u32x4v b, f32x16 c) {
  if constexpr (MODE == 0)
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

// 2^(14 - ceil(log2 x)): brings max |x| into [2^13, 2^14] (1 for x <= 0 or
// not finite); ceil(log2 x) from the exponent, exact for powers of two.
__device__ __forceinline__ float pow2_scale_dev(double x) {
  if (!(x > 0.0) || !isfinite(x)) return 1.f;
  int e;
  const double f = frexp(x, &e);             // x = f 2^e, f in [0.5, 1)
  if (f == 0.5) --e;
  return ldexpf(1.f, 14 - e);
}

template <int MODE, bool AT, int NT, bool D = false>
__global__ __launch_bounds__(NTHR, 1) void gemm_lp_kernel(
    int64_t M, int64_t N, int64_t K, const void* __restrict__ Av, int64_t lda,
    const void* __restrict__ Bv, int64_t ldb, float* __restrict__ C, int64_t ldc,
    int64_t cstride, float alpha, float sa, float sb, const double* __restrict__ amax,
    int64_t kchunk) {
  using CF = Cfg<MODE>;
  if (MODE != 0 && amax) {
    // power-of-two split scales from max|op(A)| and max|B| on the device
    sa = pow2_scale_dev(amax[0]);
    sb = pow2_scale_dev(amax[1]);
    alpha = alpha / (sa * sb);
  }
  constexpr int BK = CF::BK, ROWB = CF::ROWB, CH = CF::CH, PLA = CF::PLA, KQ = CF::KQ;
  constexpr int BN = 64 * NT;
  constexpr int A_BYTES = BM * ROWB, B_BYTES = BN * ROWB;
  constexpr int STAGE = PLA * A_BYTES + CF::PLB * B_BYTES;
  constexpr int BCH = BN * CH;                            // B chunks per tile
  constexpr int BIT = (BCH + NTHR - 1) / NTHR;            // B chunks per thread
  extern __shared__ __attribute__((aligned(16))) char smem[];

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 3, wn = wave >> 2;
  // Split-K order across the XCDs: work-groups are dispatched to the 8 XCDs
  // round-robin by linear id, so (M tile, split) is derived from that id such
  // that each XCD runs whole splits and a split's B rows live in one XCD's L2
  // (with the plain order every XCD would fetch every split's B rows).
  int64_t bx = blockIdx.x, bz = blockIdx.z;
  if (gridDim.y == 1 && gridDim.z > 1) {
    const unsigned S = gridDim.z, Mt = gridDim.x;
    const unsigned L = blockIdx.x + Mt * blockIdx.z, x8 = L & 7, r = L >> 3;
    if (8 % S == 0 && Mt % (8 / S) == 0) {
      bz = x8 % S;
      bx = r * (8 / S) + x8 / S;
    } else if (S % 8 == 0) {
      bz = x8 + 8 * (r / Mt);
      bx = r % Mt;
    }
  }
  const int64_t m0 = bx * BM;
  const int64_t n0 = (int64_t)blockIdx.y * BN;
  const int64_t kbeg = bz * kchunk;
  const int64_t kend = min(K, kbeg + kchunk);
  C += bz * cstride;

  // ---- register staging of the next tile ----
  // MODE 0: raw BF16 chunks.  MODE 1: raw FP32 (8 per chunk).
  constexpr int AREG = (MODE == 0) ? 4 : 4;   // 16-byte units for A
  u32x4v ra[MODE == 0 ? AREG : 2 * AREG];
  u32x4v rb[MODE == 0 ? BIT : 2 * BIT];

  auto load = [&](int64_t k0) {
    if constexpr (MODE == 0) {
      const uint16_t* A = (const uint16_t*)Av;
      if constexpr (AT) {
        if constexpr (!D) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int idx = t + NTHR * i, r = idx / CH, c = idx % CH;
            const int64_t m = m0 + r, k = k0 + 8 * c;
            ra[i] = (m < M && k < kend) ? *(const u32x4v*)(A + m * lda + k) : u32x4v{0, 0, 0, 0};
          }
        }
      } else {
        const int rb8 = (t & 31) * 8, kb = (t >> 5) * KQ;
#pragma unroll
        for (int j = 0; j < KQ; ++j) {
          const int64_t k = k0 + kb + j, m = m0 + rb8;
          if constexpr (D) ra[j] = *(const u32x4v*)(A + k * lda + m);
          else ra[j] = (m < M && k < kend) ? *(const u32x4v*)(A + k * lda + m) : u32x4v{0, 0, 0, 0};
        }
      }
      if constexpr (!D) {
        const uint16_t* B = (const uint16_t*)Bv;
#pragma unroll
        for (int i = 0; i < BIT; ++i) {
          const int idx = t + NTHR * i;
          const int r = idx / CH, c = idx % CH;
          const int64_t n = n0 + r, k = k0 + 8 * c;
          rb[i] = (idx < BCH && n < N && k < kend) ? *(const u32x4v*)(B + n * ldb + k)
                                                   : u32x4v{0, 0, 0, 0};
        }
      }
    } else {
      const float* A = (const float*)Av;
      if constexpr (AT) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int idx = t + NTHR * i, r = idx / CH, c = idx % CH;
          const int64_t m = m0 + r, k = k0 + 8 * c;
          const bool ok = (m < M && k < kend);
          ra[2 * i] = ok ? *(const u32x4v*)(A + m * lda + k) : u32x4v{0, 0, 0, 0};
          ra[2 * i + 1] = ok ? *(const u32x4v*)(A + m * lda + k + 4) : u32x4v{0, 0, 0, 0};
        }
      } else {
        const int rb8 = (t & 31) * 8, kb = (t >> 5) * KQ;
#pragma unroll
        for (int j = 0; j < KQ; ++j) {
          const int64_t k = k0 + kb + j, m = m0 + rb8;
          const bool ok = (m < M && k < kend);
          ra[2 * j] = ok ? *(const u32x4v*)(A + k * lda + m) : u32x4v{0, 0, 0, 0};
          ra[2 * j + 1] = ok ? *(const u32x4v*)(A + k * lda + m + 4) : u32x4v{0, 0, 0, 0};
        }
      }
      if constexpr (!D) {
        const float* B = (const float*)Bv;
#pragma unroll
        for (int i = 0; i < BIT; ++i) {
          const int idx = t + NTHR * i;
          const int r = idx / CH, c = idx % CH;
          const int64_t n = n0 + r, k = k0 + 8 * c;
          const bool ok = (idx < BCH && n < N && k < kend);
          rb[2 * i] = ok ? *(const u32x4v*)(B + n * ldb + k) : u32x4v{0, 0, 0, 0};
          rb[2 * i + 1] = ok ? *(const u32x4v*)(B + n * ldb + k + 4) : u32x4v{0, 0, 0, 0};
        }
      }
    }
  };

  // D (every tile full; B in 16-bit form: BF16, or the FP16 panel of mode 2):
  // B's tile goes from global memory straight into LDS
  // (global_load_lds_dwordx4: lane l of a wave lands at base + 16 l).
  // Instruction j of a tile fills RPI = 64 / CH rows from row RPI j; lane l
  // takes row RPI j + l / CH and the logical chunk whose swizzled slot is
  // l % CH, which is the layout the register path stores.
  auto load_b_lds = [&](int64_t k0, int stage) {
    if constexpr (D) {
      constexpr int RPI = 64 / CH, NINS = BN * CH / 64;
      if constexpr (AT && MODE == 0) {
        // A^T's rows are K-contiguous too: the same direct path (BM rows).
        const uint16_t* A = (const uint16_t*)Av;
        char* sA = smem + stage * STAGE;
#pragma unroll
        for (int i = 0; i < BM * CH / NTHR; ++i) {
          const int j = i * (NTHR / 64) + wave;
          const int r = 8 * j + (lane >> 3), p = lane & 7;
          const int c = p ^ ((r >> CF::SW) & (CH - 1));
          const uint16_t* gp = A + (m0 + r) * lda + k0 + 8 * c;
          __builtin_amdgcn_global_load_lds((const void*)gp,
                                           (__attribute__((address_space(3))) void*)(sA + j * 1024),
                                           16, 0, 0);
        }
      }
      char* sB = smem + stage * STAGE + PLA * A_BYTES;
#pragma unroll
      for (int pl = 0; pl < CF::PLB; ++pl) {   // mode 1: hi plane, then lo plane (Bv + N ldb)
        const uint16_t* B = (const uint16_t*)Bv + pl * N * ldb;
#pragma unroll
        for (int i = 0; i < BIT; ++i) {
          const int j = i * (NTHR / 64) + wave;
          if (NINS % (NTHR / 64) == 0 || j < NINS) {
            const int r = RPI * j + lane / CH, p = lane % CH;
            const int c = p ^ ((r >> CF::SW) & (CH - 1));
            const uint16_t* gp = B + (n0 + r) * ldb + k0 + 8 * c;
            __builtin_amdgcn_global_load_lds(
                (const void*)gp,
                (__attribute__((address_space(3))) void*)(sB + pl * B_BYTES + j * 1024), 16, 0,
                0);
          }
        }
      }
    }
  };

  auto store = [&](int stage) {
    char* sA = smem + stage * STAGE;
    char* sB = sA + PLA * A_BYTES;
    if constexpr (MODE == 0) {
      if constexpr (AT) {
        if constexpr (!D) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int idx = t + NTHR * i, r = idx / CH, c = idx % CH;
            *(u32x4v*)(sA + r * ROWB + swz<MODE>(r, c)) = ra[i];
          }
        }
      } else {
        // 8 rows x 4 k -> per row one 8-byte piece (k .. k+3).
        const int rb8 = (t & 31) * 8, kb = (t >> 5) * KQ;
        const int c = kb >> 3, off = (kb & 7) * 2;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const int w = r >> 1, sh = (r & 1) * 16;
          const uint32_t e0 = (ra[0][w] >> sh) & 0xffffu, e1 = (ra[1][w] >> sh) & 0xffffu;
          const uint32_t e2 = (ra[2][w] >> sh) & 0xffffu, e3 = (ra[3][w] >> sh) & 0xffffu;
          const int row = rb8 + r;
          *(u32x2v*)(sA + row * ROWB + swz<MODE>(row, c) + off) = u32x2v{e0 | (e1 << 16), e2 | (e3 << 16)};
        }
      }
      if constexpr (!D) {
#pragma unroll
        for (int i = 0; i < BIT; ++i) {
          const int idx = t + NTHR * i;
          if (idx < BCH) {
            const int r = idx / CH, c = idx % CH;
            *(u32x4v*)(sB + r * ROWB + swz<MODE>(r, c)) = rb[i];
          }
        }
      }
    } else {
      if constexpr (AT) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int idx = t + NTHR * i, r = idx / CH, c = idx % CH;
          uint32_t h[8], l[8];
#pragma unroll
          for (int e = 0; e < 8; ++e)
            split_f16(__uint_as_float(ra[2 * i + (e >> 2)][e & 3]), sa, h[e], l[e]);
          const int o = r * ROWB + swz<MODE>(r, c);
          *(u32x4v*)(sA + o) = u32x4v{h[0] | (h[1] << 16), h[2] | (h[3] << 16),
                                      h[4] | (h[5] << 16), h[6] | (h[7] << 16)};
          if constexpr (MODE < 3)
            *(u32x4v*)(sA + A_BYTES + o) = u32x4v{l[0] | (l[1] << 16), l[2] | (l[3] << 16),
                                                  l[4] | (l[5] << 16), l[6] | (l[7] << 16)};
        }
      } else {
        // 8 rows x 2 k -> per row one 4-byte piece (k, k+1) per plane.
        const int rb8 = (t & 31) * 8, kb = (t >> 5) * KQ;
        const int c = kb >> 3, off = (kb & 7) * 2;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          uint32_t h0, l0, h1, l1;
          split_f16(__uint_as_float(ra[r >> 2][r & 3]), sa, h0, l0);        // column k
          split_f16(__uint_as_float(ra[2 + (r >> 2)][r & 3]), sa, h1, l1);  // column k+1
          const int row = rb8 + r;
          const int o = row * ROWB + swz<MODE>(row, c) + off;
          *(uint32_t*)(sA + o) = h0 | (h1 << 16);
          if constexpr (MODE < 3) *(uint32_t*)(sA + A_BYTES + o) = l0 | (l1 << 16);
        }
      }
#pragma unroll
      for (int i = 0; i < BIT; ++i) {
        const int idx = t + NTHR * i;
        if (!D && idx < BCH) {
          const int r = idx / CH, c = idx % CH;
          uint32_t h[8], l[8];
          const int o = r * ROWB + swz<MODE>(r, c);
          if constexpr (MODE == 1 || MODE == 4) {
#pragma unroll
            for (int e = 0; e < 8; ++e)
              split_f16(__uint_as_float(rb[2 * i + (e >> 2)][e & 3]), sb, h[e], l[e]);
            *(u32x4v*)(sB + B_BYTES + o) = u32x4v{l[0] | (l[1] << 16), l[2] | (l[3] << 16),
                                                  l[4] | (l[5] << 16), l[6] | (l[7] << 16)};
          } else {
#pragma unroll
            for (int e = 0; e < 8; ++e)
              h[e] = f16_bits(__uint_as_float(rb[2 * i + (e >> 2)][e & 3]) * sb);
          }
          *(u32x4v*)(sB + o) = u32x4v{h[0] | (h[1] << 16), h[2] | (h[3] << 16),
                                      h[4] | (h[5] << 16), h[6] | (h[7] << 16)};
        }
      }
    }
  };

  f32x16 acc[2][NT];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

  const int64_t nk = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;
  if (nk > 0) {
    load(kbeg);
    load_b_lds(kbeg, 0);
    store(0);
  }
  if constexpr (D) __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  const int lr = lane & 31, lk = lane >> 5;
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int stage = (int)(kt & 1);
    if (kt + 1 < nk) load(kbeg + (kt + 1) * BK);
    const char* sA = smem + stage * STAGE;
    const char* sB = sA + PLA * A_BYTES;
#pragma unroll
    for (int s = 0; s < BK / 16; ++s) {
      const int c = 2 * s + lk;
      u32x4v a[2], al[2];
#pragma unroll
      for (int mi = 0; mi < 2; ++mi) {
        const int row = wm * 64 + mi * 32 + lr;
        a[mi] = *(const u32x4v*)(sA + row * ROWB + swz<MODE>(row, c));
        if constexpr (MODE == 1 || MODE == 2)
          al[mi] = *(const u32x4v*)(sA + A_BYTES + row * ROWB + swz<MODE>(row, c));
      }
#pragma unroll
      for (int ni = 0; ni < NT; ++ni) {
        const int row = wn * 32 * NT + ni * 32 + lr;
        const u32x4v b = *(const u32x4v*)(sB + row * ROWB + swz<MODE>(row, c));
        if constexpr (MODE == 1) {
          const u32x4v bl = *(const u32x4v*)(sB + B_BYTES + row * ROWB + swz<MODE>(row, c));
#pragma unroll
          for (int mi = 0; mi < 2; ++mi) {
            acc[mi][ni] = mma<MODE>(a[mi], bl, acc[mi][ni]);
            acc[mi][ni] = mma<MODE>(al[mi], b, acc[mi][ni]);
            acc[mi][ni] = mma<MODE>(a[mi], b, acc[mi][ni]);
          }
        } else if constexpr (MODE == 2) {
#pragma unroll
          for (int mi = 0; mi < 2; ++mi) {
            acc[mi][ni] = mma<MODE>(al[mi], b, acc[mi][ni]);
            acc[mi][ni] = mma<MODE>(a[mi], b, acc[mi][ni]);
          }
        } else if constexpr (MODE == 4) {
          const u32x4v bl = *(const u32x4v*)(sB + B_BYTES + row * ROWB + swz<MODE>(row, c));
#pragma unroll
          for (int mi = 0; mi < 2; ++mi) {
            acc[mi][ni] = mma<MODE>(a[mi], bl, acc[mi][ni]);
            acc[mi][ni] = mma<MODE>(a[mi], b, acc[mi][ni]);
          }
        } else {
#pragma unroll
          for (int mi = 0; mi < 2; ++mi) acc[mi][ni] = mma<MODE>(a[mi], b, acc[mi][ni]);
        }
      }
      // the next step's direct-to-LDS loads go out once this wave's first
      // MFMAs are issued (the register loads, needed sooner, at the top)
      if (s == 0 && kt + 1 < nk) load_b_lds(kbeg + (kt + 1) * BK, stage ^ 1);
    }
    if (kt + 1 < nk) store(stage ^ 1);
    if constexpr (D) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }

  // Epilogue: lane holds rows 8*(v/4) + 4*(lane/32) + v%4, column lane%32.
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < NT; ++ni) {
      const int64_t n = n0 + wn * 32 * NT + ni * 32 + lr;
      if (n >= N) continue;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int64_t m = m0 + wm * 64 + mi * 32 + 8 * g + 4 * lk;
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

// The FP16 panel of modes 2 and 1: hi[n ldbh + k] = fp16(B[n ldb + k] sb)
// (mode 2: the rounded panel) and, for mode 1 (SPLIT), the low part of the
// same split as split_f16 in the plane after it (hi + N ldbh); sb from max|B|
// exactly as the GEMM kernel derives it (or g.sb without amax).
template <bool SPLIT>
__global__ void panel_f16_kernel(const float* __restrict__ B, int64_t K, int64_t N, int64_t ldb,
                                 uint16_t* __restrict__ Bh, int64_t ldbh,
                                 const double* __restrict__ amax, float sb0) {
  const float sb = amax ? pow2_scale_dev(amax[1]) : sb0;
  const int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
  const int64_t n = blockIdx.y;
  if (k >= K) return;                      // K % 8 == 0
  const u32x4v x0 = *(const u32x4v*)(B + n * ldb + k);
  const u32x4v x1 = *(const u32x4v*)(B + n * ldb + k + 4);
  uint32_t h[8], l[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float x = __uint_as_float(e < 4 ? x0[e] : x1[e - 4]);
    if constexpr (SPLIT) split_f16(x, sb, h[e], l[e]);
    else h[e] = f16_bits(x * sb);
  }
  *(u32x4v*)(Bh + n * ldbh + k) = u32x4v{h[0] | (h[1] << 16), h[2] | (h[3] << 16),
                                         h[4] | (h[5] << 16), h[6] | (h[7] << 16)};
  if constexpr (SPLIT)
    *(u32x4v*)(Bh + (N + n) * ldbh + k) = u32x4v{l[0] | (l[1] << 16), l[2] | (l[3] << 16),
                                                 l[4] | (l[5] << 16), l[6] | (l[7] << 16)};
}

hipError_t panel_f16(const LpGemmArgs& g, hipStream_t st) {
  dim3 grid((unsigned)((g.K / 8 + 255) / 256), (unsigned)g.N);
  if (g.mode == 1 || g.mode == 4)
    hipLaunchKernelGGL(panel_f16_kernel<true>, grid, dim3(256), 0, st, (const float*)g.B, g.K,
                       g.N, g.ldb, g.bh, g.ldbh, g.amax, g.sb);
  else
    hipLaunchKernelGGL(panel_f16_kernel<false>, grid, dim3(256), 0, st, (const float*)g.B, g.K,
                       g.N, g.ldb, g.bh, g.ldbh, g.amax, g.sb);
  return hipGetLastError();
}

template <int MODE, bool AT, int NT, bool D>
hipError_t launch_d(const LpGemmArgs& g, hipStream_t st) {
  using CF = Cfg<MODE>;
  constexpr int BN = 64 * NT;
  constexpr int STAGE = (CF::PLA * BM + CF::PLB * BN) * CF::ROWB;
  static bool attr_set = false;
  auto kern = gemm_lp_kernel<MODE, AT, NT, D>;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 2 * STAGE);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  const int64_t mt = (g.M + BM - 1) / BM, nt = (g.N + BN - 1) / BN;
  int splits = g.splits < 1 ? 1 : g.splits;
  int64_t kchunk = (g.K + splits - 1) / splits;
  kchunk = ((kchunk + CF::BK - 1) / CF::BK) * CF::BK;
  if (kchunk == 0) kchunk = CF::BK;
  dim3 grid((unsigned)mt, (unsigned)nt, (unsigned)splits);
  hipLaunchKernelGGL(kern, grid, dim3(NTHR), 2 * STAGE, st, g.M, g.N, g.K, g.A, g.lda, g.B,
                     g.ldb, g.C, g.ldc, g.cstride, g.alpha, g.sa, g.sb, g.amax, kchunk);
  return hipGetLastError();
}

// Full tiles everywhere (M, N multiples of the tile, every K tile complete)
// take the variant that loads B's tiles (and A^T's, for the transposed
// product) straight into LDS.
template <int MODE, bool AT, int NT>
hipError_t launch(const LpGemmArgs& g, hipStream_t st) {
  using CF = Cfg<MODE>;
  constexpr int BN = 64 * NT;
  {
    const int splits = g.splits < 1 ? 1 : g.splits;
    int64_t kchunk = (g.K + splits - 1) / splits;
    kchunk = ((kchunk + CF::BK - 1) / CF::BK) * CF::BK;
    const bool full = g.M % BM == 0 && g.N % BN == 0 && g.K % CF::BK == 0 && kchunk > 0;
    if constexpr (MODE != 0) {
      // the FP16 panel (mode 1: hi and lo planes): made once here, then
      // loaded straight into LDS
      if (!(AT && full && g.bh && g.ldbh >= g.K && g.ldbh % 8 == 0))
        return launch_d<MODE, AT, NT, false>(g, st);
      hipError_t e = panel_f16(g, st);
      if (e != hipSuccess) return e;
      LpGemmArgs h = g;
      h.B = g.bh;
      h.ldb = g.ldbh;
      return launch_d<MODE, AT, NT, true>(h, st);
    } else {
      const bool d = full && g.ldb % 8 == 0 && ((uintptr_t)g.B & 15) == 0 &&
                     (!AT || (g.lda % 8 == 0 && ((uintptr_t)g.A & 15) == 0));
      return d ? launch_d<MODE, AT, NT, true>(g, st) : launch_d<MODE, AT, NT, false>(g, st);
    }
  }
}

template <int MODE, bool AT>
hipError_t launch_nt(const LpGemmArgs& g, hipStream_t st) {
  const int64_t nt = (g.N + 63) / 64;
  switch (nt) {
    case 1: return launch<MODE, AT, 1>(g, st);
    case 2: return launch<MODE, AT, 2>(g, st);
    case 3: return launch<MODE, AT, 3>(g, st);
    case 4: return launch<MODE, AT, 4>(g, st);
    default: return launch<MODE, AT, 5>(g, st);
  }
}

// ---- partial-sum reduction ----------------------------------------------
template <typename TD>
__global__ void reduce_kernel(const float* __restrict__ src, int64_t M, int64_t N, int64_t lds,
                              int64_t cstride, int splits, TD* __restrict__ dst, int64_t ldd,
                              double alpha) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= M) return;
  // Eight loads in flight, summed in split order (a loop of one load and one
  // add waits for each load in turn).
  const float* p = src + j * lds + i;
  double s = 0.0;
  int z = 0;
  for (; z + 8 <= splits; z += 8) {
    float x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = p[(int64_t)(z + u) * cstride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += (double)x[u];
  }
  for (; z < splits; ++z) s += (double)p[(int64_t)z * cstride];
  dst[j * ldd + i] = (TD)(alpha * s);
}

__global__ void reduce_f64_kernel(const double* __restrict__ src, int64_t M, int64_t N,
                                  int64_t lds, int64_t cstride, int splits,
                                  double* __restrict__ dst, int64_t ldd, double alpha) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= M) return;
  // Eight loads in flight, summed in split order.
  const double* p = src + j * lds + i;
  double s = 0.0;
  int z = 0;
  for (; z + 8 <= splits; z += 8) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = p[(int64_t)(z + u) * cstride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += x[u];
  }
  for (; z < splits; ++z) s += p[(int64_t)z * cstride];
  dst[j * ldd + i] = alpha * s;
}

}  // namespace

hipError_t gemm_lp(const LpGemmArgs& g, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  // Vector loads need 8-element alignment of the leading dimensions and of
  // the contiguous extents; callers pad their buffers accordingly.
  if (g.K % 8 != 0 || g.lda % 8 != 0 || g.ldb % 8 != 0) return hipErrorInvalidValue;
  if (!g.a_trans && g.M % 8 != 0) return hipErrorInvalidValue;
  if (g.mode == 0)
    return g.a_trans ? launch_nt<0, true>(g, st) : launch_nt<0, false>(g, st);
  if (g.mode == 1)
    return g.a_trans ? launch_nt<1, true>(g, st) : launch_nt<1, false>(g, st);
  if (g.mode == 2)
    return g.a_trans ? launch_nt<2, true>(g, st) : launch_nt<2, false>(g, st);
  if (g.mode == 3)
    return g.a_trans ? launch_nt<3, true>(g, st) : launch_nt<3, false>(g, st);
  if (g.mode == 4)
    return g.a_trans ? launch_nt<4, true>(g, st) : launch_nt<4, false>(g, st);
  return hipErrorInvalidValue;
}

hipError_t reduce_partials(const float* src, int64_t M, int64_t N, int64_t lds, int64_t cstride,
                           int splits, void* dst, T tdst, int64_t ldd, double alpha,
                           hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  dim3 grid((unsigned)((M + 255) / 256), (unsigned)N);
  if (tdst == T::F64)
    hipLaunchKernelGGL(reduce_kernel<double>, grid, dim3(256), 0, st, src, M, N, lds, cstride,
                       splits, (double*)dst, ldd, alpha);
  else if (tdst == T::F32)
    hipLaunchKernelGGL(reduce_kernel<float>, grid, dim3(256), 0, st, src, M, N, lds, cstride,
                       splits, (float*)dst, ldd, alpha);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t reduce_partials_f64(const double* src, int64_t M, int64_t N, int64_t lds,
                               int64_t cstride, int splits, double* dst, int64_t ldd,
                               double alpha, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  dim3 grid((unsigned)((M + 255) / 256), (unsigned)N);
  hipLaunchKernelGGL(reduce_f64_kernel, grid, dim3(256), 0, st, src, M, N, lds, cstride,
                     splits, dst, ldd, alpha);
  return hipGetLastError();
}

}  // namespace rna
