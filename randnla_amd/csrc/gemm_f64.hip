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
#include <algorithm>

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

// Register-stage element: the raw memory value, converted when the stage is
// stored to LDS (so that a stage's loads issue back to back, each from a
// clamped address, with no branch or conversion waiting on any of them), or
// the generated double of a generated operand.
template <int G, typename TI> struct Raw { using type = double; };
template <typename TI> struct Raw<0, TI> { using type = TI; };
template <typename TI>
__device__ __forceinline__ double to_f64(TI v) { return (double)v; }
template <>
__device__ __forceinline__ double to_f64<uint16_t>(uint16_t v) { return (double)bf16_to_f32(v); }
// (out-of-range elements load p[0]; the stage's mask zeroes them at the store)
template <int G, typename TI>
__device__ __forceinline__ typename Raw<G, TI>::type fetch_raw(const TI* p, int64_t off, bool ok,
                                                               const Stream& s, uint64_t e) {
  if constexpr (G == 0) return p[ok ? off : 0];
  else return ok ? fetch<G>(p, off, s, e) : 0.0;
}

template <typename TO>
__device__ __forceinline__ void st_from_f64(TO* p, double v) { *p = (TO)v; }
template <>
__device__ __forceinline__ void st_from_f64<uint16_t>(uint16_t* p, double v) {
  *p = (uint16_t)bf16_bits((float)v);
}

template <bool AT, typename TA, typename TB, typename TC, int GA, int GB>
__global__ __launch_bounds__(NTHR, 2) void gemm_f64_kernel(
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
  // Two register stages: the tiles of step k + 2 are loaded while step k is
  // multiplied and step k + 1's are stored to LDS (the K step's compute is
  // shorter than the memory latency).
  using RA = typename Raw<GA, TA>::type;
  using RB = typename Raw<GB, TB>::type;
  RA ra0[8], ra1[8];
  RB rb0[4], rb1[4];
  uint32_t mk0 = 0, mk1 = 0;   // bit i (A) / 8 + i (B): element in range

  auto load = [&](int64_t k0, RA (&ra)[8], RB (&rb)[4], uint32_t& mk) {
    mk = 0;
    if (!AT) {
      const int mm = (t & 63) * 2;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int kk = (t >> 6) + 4 * i;
        const int64_t k = k0 + kk;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int64_t m = m0 + mm + e;
          const bool ok = m < M && k < kend;
          ra[2 * i + e] = fetch_raw<GA>(A, k * lda + m, ok, gs, (uint64_t)(m + k * M));
          mk |= (uint32_t)ok << (2 * i + e);
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
          const bool ok = m < M && k < kend;
          ra[2 * i + e] = fetch_raw<GA>(A, m * lda + k, ok, gs, (uint64_t)(k + m * K));
          mk |= (uint32_t)ok << (2 * i + e);
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
          const bool ok = n < N && k < kend;
          rb[2 * i + e] = fetch_raw<GB>(B, n * ldb + k, ok, gs, (uint64_t)(k + n * K));
          mk |= (uint32_t)ok << (8 + 2 * i + e);
        }
      }
    }
  };
  auto store = [&](int buf, const RA (&ra)[8], const RB (&rb)[4], uint32_t mk) {
    auto va = [&](int i) { return (mk >> i) & 1 ? to_f64(ra[i]) : 0.0; };
    auto vb = [&](int i) { return (mk >> (8 + i)) & 1 ? to_f64(rb[i]) : 0.0; };
    if (!AT) {
      const int mm = (t & 63) * 2;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int kk = (t >> 6) + 4 * i;
        *(double2*)&As[buf][kk * PADM + mm] = double2{va(2 * i), va(2 * i + 1)};
      }
    } else {
      const int kk = (t & 7) * 2;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int mm = (t >> 3) + 32 * i;
        *(double2*)&As[buf][mm * LDK + kk] = double2{va(2 * i), va(2 * i + 1)};
      }
    }
    const int kk = (t & 7) * 2;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int nn = (t >> 3) + 32 * i;
      *(double2*)&Bs[buf][nn * LDK + kk] = double2{vb(2 * i), vb(2 * i + 1)};
    }
  };

  f64x4 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  const int64_t nk = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;
  auto compute = [&](int buf) {
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
  };
  // Barrier that leaves the next stage's global loads in flight
  // (__syncthreads waits for every outstanding load): only this wave's LDS
  // accesses must be complete.
  auto lds_barrier = []() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); };
  // Loads past the K range are issued all the same (clamped to the operand's
  // first element, masked to zero): unconditional loads keep the waits
  // before each stage's store exact (the other stage stays in flight).
  if constexpr (GA != 0 || GB != 0) {
    // generated operands: no load latency to hide, one register stage
    if (nk > 0) {
      load(kbeg, ra0, rb0, mk0);
      store(0, ra0, rb0, mk0);
    }
    __syncthreads();
    for (int64_t kt = 0; kt < nk; ++kt) {
      if (kt + 1 < nk) load(kbeg + (kt + 1) * BK, ra0, rb0, mk0);
      compute((int)(kt & 1));
      if (kt + 1 < nk) store((int)((kt + 1) & 1), ra0, rb0, mk0);
      __syncthreads();
    }
  } else {
  if (nk > 0) {
    load(kbeg, ra0, rb0, mk0);
    store(0, ra0, rb0, mk0);
  }
  load(kbeg + BK, ra1, rb1, mk1);
  lds_barrier();
  for (int64_t kt = 0; kt < nk; kt += 2) {
    // even step: LDS buffer 0 holds step kt, stage 1 step kt + 1
    load(kbeg + (kt + 2) * BK, ra0, rb0, mk0);
    compute(0);
    if (kt + 1 < nk) store(1, ra1, rb1, mk1);
    lds_barrier();
    if (kt + 1 >= nk) break;
    // odd step: LDS buffer 1 holds step kt + 1, stage 0 step kt + 2
    load(kbeg + (kt + 3) * BK, ra1, rb1, mk1);
    compute(1);
    if (kt + 2 < nk) store(0, ra0, rb0, mk0);
    lds_barrier();
  }
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

// ---- Gram of a tall panel: G = Y^T Y (upper triangle, FP64) ---------------
// Y is m x l column-major (FP32 or FP64), l <= 320.  The upper 16 x 16 tiles
// of G (a <= b, numbered column by column: i = b (b + 1) / 2 + a) are cut into
// four groups; work-group (group, split) accumulates its group's tiles over
// one K-range of rows with v_mfma_f64_16x16x4f64, every wave holding up to
// seven tiles in registers.  The rows' slab (128 bytes of every column the
// group touches) is staged in LDS as [column][k] with a stride that puts the
// 64 lanes of an operand read on distinct banks, double buffered, one
// barrier per K step.  Every needed Y column is read once per group from
// global memory (no tile re-reads), and the four groups of a split run on one
// XCD (work-group id = 32 r + 8 group + x, split = 8 r + x), so each row of Y
// comes from HBM once.  The splits' partial tiles are summed in split order
// (syrk_reduce_kernel), giving the same G on every run.
constexpr int SY_THR = 512;
constexpr int SY_GROUPS = 4;
constexpr int SY_TPW = 7;        // tiles per wave: ceil(210 / 4 / 8)
constexpr int SY_MAXL = 320;
template <typename TA> struct SyCfg;
template <> struct SyCfg<float> { static constexpr int BK = 32, LDK = 36; };
template <> struct SyCfg<double> { static constexpr int BK = 16, LDK = 18; };

__device__ __forceinline__ int tile_col(int i) {   // b of upper tile i
  int b = (int)((sqrtf(8.0f * (float)i + 1.0f) - 1.0f) * 0.5f);
  while (b * (b + 1) / 2 > i) --b;
  while ((b + 1) * (b + 2) / 2 <= i) ++b;
  return b;
}

template <typename TA>
__global__ __launch_bounds__(SY_THR, 1) void syrk_kernel(const TA* __restrict__ Y, int64_t ldy,
                                                         int64_t m, int l, int64_t kchunk,
                                                         double* __restrict__ part) {
  using CF = SyCfg<TA>;
  constexpr int BK = CF::BK, LDK = CF::LDK;
  constexpr int CPC = BK * (int)sizeof(TA) / 16;   // 16-byte chunks per column and slab
  constexpr int EPC = 16 / (int)sizeof(TA);        // elements per chunk
  constexpr int SLAB = SY_MAXL * LDK;              // elements per buffer
  constexpr int NLD = SY_MAXL * CPC / SY_THR;      // chunks per thread (5)
  extern __shared__ __attribute__((aligned(16))) char smem[];
  TA* buf = (TA*)smem;
  const int t = threadIdx.x, ln = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int bid = blockIdx.x;
  const int grp = (bid >> 3) & (SY_GROUPS - 1);
  const int split = (bid >> 5) * 8 + (bid & 7);
  const int T = (l + 15) / 16, ntile = T * (T + 1) / 2;
  const int i0 = grp * ntile / SY_GROUPS, i1 = (grp + 1) * ntile / SY_GROUPS;
  // columns this group reads: those of tile columns b <= b(i1 - 1) (a <= b)
  const int ncol = i1 > i0 ? min(16 * (tile_col(i1 - 1) + 1), 16 * T) : 0;
  const int64_t kbeg = (int64_t)split * kchunk, kend = min(m, kbeg + kchunk);
  const int nsteps = kend > kbeg ? (int)((kend - kbeg + BK - 1) / BK) : 0;

  // this wave's tiles: i = i0 + wv + 8 j; LDS element offsets of their
  // operand reads (row block a, column block b)
  int offa[SY_TPW], offb[SY_TPW];
  bool on[SY_TPW];
#pragma unroll
  for (int j = 0; j < SY_TPW; ++j) {
    const int i = i0 + wv + 8 * j;
    on[j] = i < i1;
    const int b = on[j] ? tile_col(i) : 0, a = on[j] ? i - b * (b + 1) / 2 : 0;
    offa[j] = (16 * a + (ln & 15)) * LDK + (ln >> 4);
    offb[j] = (16 * b + (ln & 15)) * LDK + (ln >> 4);
  }
  f64x4 acc[SY_TPW];
#pragma unroll
  for (int j = 0; j < SY_TPW; ++j) acc[j] = f64x4{0.0, 0.0, 0.0, 0.0};

  // slab loads: chunk q = t + SY_THR u -> column q / CPC, 16-byte part q % CPC
  using V4 = __attribute__((ext_vector_type(4))) unsigned;
  V4 rg[NLD];
  auto load = [&](int step) {
    const int64_t k0 = kbeg + (int64_t)step * BK;
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
      const int q = t + SY_THR * u, col = q / CPC, pc = q % CPC;
      const int64_t k = k0 + pc * EPC;
      V4 v = {0u, 0u, 0u, 0u};
      if (col < ncol && col < l) {
        const TA* src = Y + (int64_t)col * ldy + k;
        if (k + EPC <= kend) {
          v = *(const V4*)src;
        } else {
          TA e[EPC];
#pragma unroll
          for (int x = 0; x < EPC; ++x) e[x] = k + x < kend ? src[x] : (TA)0;
          v = *(const V4*)e;
        }
      }
      rg[u] = v;
    }
  };
  auto store = [&](int b) {
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
      const int q = t + SY_THR * u, col = q / CPC, pc = q % CPC;
      if (col < ncol) *(V4*)(buf + b * SLAB + col * LDK + pc * EPC) = rg[u];
    }
  };

  if (nsteps > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int s = 0; s < nsteps; ++s) {
    if (s + 1 < nsteps) load(s + 1);
    const TA* bp = buf + (s & 1) * SLAB;
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
#pragma unroll
      for (int j = 0; j < SY_TPW; ++j) {      // unused slots compute tile (0, 0), unstored
        const double x = (double)bp[offa[j] + 4 * kk];
        const double y = (double)bp[offb[j] + 4 * kk];
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[j], 0, 0, 0);
      }
    }
    if (s + 1 < nsteps) store((s + 1) & 1);
    __syncthreads();
  }
  // partial tiles of this split: entry (row, col) at part[split][col * l + row]
  double* P = part + (int64_t)split * l * l;
#pragma unroll
  for (int j = 0; j < SY_TPW; ++j) {
    if (!on[j]) continue;
    const int i = i0 + wv + 8 * j, b = tile_col(i), a = i - b * (b + 1) / 2;
    const int col = 16 * b + (ln & 15);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 16 * a + 4 * v + (ln >> 4);
      if (row < l && col < l) P[(int64_t)col * l + row] = acc[j][v];
    }
  }
}

// G[c][r] (column-major, ldg) = sum over splits in order for r <= c, 0 below.
__global__ void syrk_reduce_kernel(const double* __restrict__ part, int l, int splits,
                                   double* __restrict__ G, int64_t ldg) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (r >= l) return;
  double s = 0.0;
  if (r <= c) {
    // Eight loads in flight, summed in split order.
    const double* p = part + (int64_t)c * l + r;
    int k = 0;
    for (; k + 8 <= splits; k += 8) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = p[(int64_t)(k + u) * l * l];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += x[u];
    }
    for (; k < splits; ++k) s += p[(int64_t)k * l * l];
  }
  G[(int64_t)c * ldg + r] = s;
}

void syrk_shape(int64_t m, int64_t l, T ty, int64_t* kchunk, int* splits) {
  const int bk = ty == T::F32 ? SyCfg<float>::BK : SyCfg<double>::BK;
  int64_t kc = (m + 63) / 64;                       // 64 splits for tall panels
  kc = std::max<int64_t>((kc + bk - 1) / bk * bk, 4 * bk);
  int sp = (int)((m + kc - 1) / kc);
  sp = (sp + 7) / 8 * 8;                            // whole XCD rounds
  *kchunk = kc;
  *splits = sp;
  (void)l;
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

bool syrk_supported(T ty, int64_t l, int64_t ldy, const void* Y) {
  const size_t es = ty == T::F32 ? 4 : 8;
  return (ty == T::F32 || ty == T::F64) && l > 0 && l <= SY_MAXL && ((ldy * es) % 16) == 0 &&
         ((uintptr_t)Y & 15) == 0;
}

size_t syrk_workspace(int64_t m, int64_t l, T ty) {
  int64_t kc;
  int sp;
  syrk_shape(m, l, ty, &kc, &sp);
  return (size_t)sp * l * l * sizeof(double);
}

hipError_t syrk_f64(const void* Y, T ty, int64_t m, int64_t l, int64_t ldy, double* G,
                    int64_t ldg, double* part, hipStream_t st) {
  if (!syrk_supported(ty, l, ldy, Y) || m <= 0) return hipErrorInvalidValue;
  int64_t kc;
  int sp;
  syrk_shape(m, l, ty, &kc, &sp);
  const size_t lds = (size_t)2 * SY_MAXL *
                     (ty == T::F32 ? SyCfg<float>::LDK * 4 : SyCfg<double>::LDK * 8);
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)syrk_kernel<float>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)syrk_kernel<double>,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(2 * SY_MAXL * SyCfg<double>::LDK * 8));
    if (e != hipSuccess) return e;
    attr = true;
  }
  const dim3 grid((unsigned)(sp * SY_GROUPS));
  if (ty == T::F32)
    hipLaunchKernelGGL(syrk_kernel<float>, grid, dim3(SY_THR), lds, st, (const float*)Y, ldy, m,
                       (int)l, kc, part);
  else
    hipLaunchKernelGGL(syrk_kernel<double>, grid, dim3(SY_THR), lds, st, (const double*)Y, ldy, m,
                       (int)l, kc, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(syrk_reduce_kernel, dim3((unsigned)((l + 255) / 256), (unsigned)l), dim3(256),
                     0, st, part, (int)l, sp, G, ldg);
  return hipGetLastError();
}

}  // namespace rna
