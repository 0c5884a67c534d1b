// Batched uneven overdetermined least squares in one persistent launch.
//
// Each problem i is min ||A_i x_i - b_i||_2 with A_i (m_i x n_i, FP32,
// column-major) and b_i (FP64), solved the way blendenpik_overdetermined
// (src/sketch_and_precondition.rs:26-59) solves one problem with its LSQR
// option: a Gaussian sketch S_i (d_i = min(m_i, floor(sf * n_i)) rows) gives
// S_i [A_i | b_i]; R_i from the sketch's Gram matrix (FP64 Cholesky) is the
// right preconditioner; LSQR (src/solvers.rs:115-278) runs on A_i R_i^{-1}
// from z0 = Q^T S_i b_i = R_i^{-T} (S_i A_i)^T S_i b_i; x_i = R_i^{-1} z.
//
// One launch of a persistent kernel drains a work queue:
//   * sketch items (problem, 64-row block of S_i): S_i [A_i | b_i] on the
//     FP32 matrix cores (v_mfma_f32_16x16x4f32), S_i's Gaussian entries made
//     in registers from Philox4x32-10 and never stored;
//   * solve items (problem): Gram (v_mfma_f64_16x16x4f64), Cholesky of the
//     augmented Gram (its last column becomes z0), R^{-1}, and the LSQR loop
//     with one fused pass over A_i per iteration: t = A y - a u and
//     A^T t are both formed from one LDS-staged chunk of rows.
// Items are ordered largest problem first, each problem's sketch items just
// before its solve item; a solve item waits (bounded) for its problem's
// sketch items, which were popped earlier and never wait, so the queue cannot
// deadlock whatever the residency.  The zeroing of the queue counters and the
// launch are captured in a hipGraph and replayed; input pointers reach the
// kernel through a device-side argument block.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace rna {
namespace {

constexpr int BT = 256;                        // threads per work-group
constexpr int SK_ROWS = 64;                    // sketch rows per item (4 waves x 16)
constexpr int SK_KC = 32;                      // K rows of [A | b] staged per step
constexpr int SK_KS = SK_KC + 4;               // LDS stride (floats) of a staged column
constexpr int MAX_N = 256;
constexpr int MAX_NCB = (MAX_N + 1 + 15) / 16; // 17 column blocks of [A | b]
constexpr int NVEC = 8;                        // n-vectors (and scratch) in LDS
constexpr int VSTR = MAX_N + 8;                // LDS doubles per vector (holds n + 1)
constexpr int CHUNK_FLOATS = 2 * (4352 + MAX_N); // two LSQR chunk buffers
constexpr unsigned WAIT_SPINS = 1u << 22;

struct Prob {
  int m, n, d, lda;
  int nrb, ldsa;
  int64_t a_off, b_off, x_off;
  int64_t sid;      // sketch substream
  int64_t sa_off;   // floats: [S A | S b], d x (n+1), ld ldsa
  int64_t g_off;    // doubles: (n+1) x (n+1) augmented Gram, then R
  int64_t xi_off;   // doubles: n x n, X = R^{-1}
  int64_t u_off;    // doubles: m, LSQR u (unnormalised)
};

struct Info {
  int32_t status, istop, itn, d;
  double r2norm, arnorm, anorm, acond, xnorm;
};

struct Args {
  const Prob* probs;
  const int2* items;
  int n_items, nprob;
  int* ctl;         // [0] queue head, [1] error flag, [2 + p] sketch blocks done
  float* wsf;
  double* wsd;
  Info* info;
  const float* A;
  const double* b;
  double* x;
  uint32_t key0, key1;
  uint64_t ctr0, ctr1;
  double tol, conlim;
  int iter_lim, rshift;   // rshift: halvings of the LSQR rows per wave step
};

// ---- sketch entries --------------------------------------------------------
// Entry (r, k) of S_p, for k in the 4-group kq = k / 4, comes from Philox4x32
// block  blk = r * ceil(m/4) + kq  of the problem's substream id: 128-bit
// counter (ctr1:ctr0) + blk + ((id + 1) << 96).  The block's words w0..w3 give
// Box-Muller pairs (w0, w1) -> (g0, g1) and (w2, w3) -> (g2, g3) with
//   u = (2 (w >> 9) + 1) 2^-24  (in (0, 1)),  theta = (w' >> 8) 2^-24 turns,
//   g = sqrt(-2 ln u) (cos 2 pi theta, sin 2 pi theta),
// evaluated with the hardware log2 / sin / cos / sqrt in FP32.
__device__ __forceinline__ f32x4 gauss4_f32(uint32_t k0, uint32_t k1, uint64_t c0, uint64_t c1,
                                            int64_t sid, uint64_t blk) {
  const uint64_t lo = c0 + blk;
  const uint64_t hi = c1 + (lo < blk ? 1ull : 0ull) + ((uint64_t)(sid + 1) << 32);
  u32x4 c;
  c.v[0] = (uint32_t)lo; c.v[1] = (uint32_t)(lo >> 32);
  c.v[2] = (uint32_t)hi; c.v[3] = (uint32_t)(hi >> 32);
  const u32x4 w = philox4x32(c, k0, k1);
  const float u1 = (float)(2u * (w.v[0] >> 9) + 1u) * 0x1p-24f;
  const float t1 = (float)(w.v[1] >> 8) * 0x1p-24f;
  const float u2 = (float)(2u * (w.v[2] >> 9) + 1u) * 0x1p-24f;
  const float t2 = (float)(w.v[3] >> 8) * 0x1p-24f;
  const float m2ln2 = -1.3862943611198906f;   // -2 ln 2
  const float r1 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u1));
  const float r2 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u2));
  f32x4 g;
  g[0] = r1 * __builtin_amdgcn_cosf(t1);
  g[1] = r1 * __builtin_amdgcn_sinf(t1);
  g[2] = r2 * __builtin_amdgcn_cosf(t2);
  g[3] = r2 * __builtin_amdgcn_sinf(t2);
  return g;
}

// ---- sketch item: rows [64 rb, 64 rb + 64) of S_p [A_p | b_p] --------------
// Wave w owns 16 rows.  Per 16-wide K slice a lane makes the 4 entries of its
// row for K positions 4 (lane / 16) .. +3; MFMA step s then pairs K slot
// (lane / 16) with logical K = 4 (lane / 16) + s on both operands (a fixed
// permutation of the reduction order), so one Philox block feeds 4 MFMAs and
// one 16-byte LDS read gives the 4 matching B operands.
__device__ void sketch_item(const Args& a, const Prob& P, int p, int rb, float* lds) {
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int n = P.n, N = n + 1, m = P.m;
  const int ncb = (N + 15) >> 4;
  const uint64_t m4 = (uint64_t)((m + 3) >> 2);
  const int row = rb * SK_ROWS + w * 16 + (l & 15);
  const bool row_ok = row < P.d;
  f32x4 acc[MAX_NCB];
#pragma unroll
  for (int cb = 0; cb < MAX_NCB; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const float* Ap = a.A + P.a_off;
  const double* bp = a.b + P.b_off;
  for (int k0 = 0; k0 < m; k0 += SK_KC) {
    __syncthreads();
    for (int idx = t; idx < ncb * 16 * SK_KC; idx += BT) {
      const int c = idx / SK_KC, kk = idx % SK_KC, k = k0 + kk;
      float v = 0.f;
      if (k < m) {
        if (c < n) v = Ap[(int64_t)c * P.lda + k];
        else if (c == n) v = (float)bp[k];
      }
      lds[c * SK_KS + kk] = v;
    }
    __syncthreads();
#pragma unroll
    for (int sc = 0; sc < SK_KC / 16; ++sc) {
      f32x4 g = f32x4{0.f, 0.f, 0.f, 0.f};
      if (row_ok) {
        const uint64_t kq = (uint64_t)(k0 >> 2) + 4 * sc + (l >> 4);
        g = gauss4_f32(a.key0, a.key1, a.ctr0, a.ctr1, P.sid, (uint64_t)row * m4 + kq);
      }
#pragma unroll
      for (int cb = 0; cb < MAX_NCB; ++cb) {
        if (cb < ncb) {
          const f32x4 bv =
              *(const f32x4*)&lds[(cb * 16 + (l & 15)) * SK_KS + sc * 16 + 4 * (l >> 4)];
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(g[0], bv[0], acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(g[1], bv[1], acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(g[2], bv[2], acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(g[3], bv[3], acc[cb], 0, 0, 0);
        }
      }
    }
  }
  float* sa = a.wsf + P.sa_off;
#pragma unroll
  for (int cb = 0; cb < MAX_NCB; ++cb) {
    if (cb < ncb) {
      const int col = cb * 16 + (l & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = rb * SK_ROWS + w * 16 + 4 * (l >> 4) + v;
        if (r < P.d && col < N) sa[(int64_t)col * P.ldsa + r] = acc[cb][v];
      }
    }
  }
  __threadfence();
  __syncthreads();
  if (t == 0) atomicAdd(&a.ctl[2 + p], 1);
}

// ---- solve item -------------------------------------------------------------
struct SolveLds {
  double* z;      // LSQR iterate (preconditioned variable)
  double* v;
  double* w;
  double* y;      // X v
  double* s;      // A^T u accumulator
  double* row;    // a row of R / scratch
  double* tpart;  // partial row sums, BT entries
  double* uc;     // chunk of u', <= BT entries
  double* red;    // block reduction scratch (>= 8)
  float* chunk;   // two LSQR chunk buffers, CHUNK_FLOATS
};

__device__ __forceinline__ double bsum(double v, double* red) { return block_sum(v, red); }

// y = X v with X upper n x n (ld n): y_i = sum_{c >= i} X[i, c] v_c.
__device__ void x_apply(const double* X, int n, const double* v, double* y) {
  for (int i = threadIdx.x; i < n; i += BT) {
    double acc = 0.0;
    for (int c = i; c < n; ++c) acc += X[i + (int64_t)c * n] * v[c];
    y[i] = acc;
  }
}
// y = X^T v: y_c = sum_{i <= c} X[i, c] v_i.
__device__ void xt_apply(const double* X, int n, const double* v, double* y) {
  for (int c = threadIdx.x; c < n; c += BT) {
    double acc = 0.0;
    const double* col = X + (int64_t)c * n;
    for (int i = 0; i <= c; ++i) acc += col[i] * v[i];
    y[c] = acc;
  }
}

// One pass over A (m x n): u'_r = c1 (A y)_r + c2 src_r written to dst (may
// alias src), s = A^T u' into L.s; returns ||u'||^2 (same value in every
// thread).
//
// Chunks of cr rows (a power of two) are staged in LDS as [column][row]
// (stride cr + 1), double-buffered: while chunk i is processed from one
// buffer, chunk i + 1 (loaded into registers during chunk i - 1) is written
// to the other and chunk i + 2 is being loaded.  Per chunk:
//   t-phase: thread (r, g) = (t / G, t mod G), G = BT / cr, sums columns
//            c = g, g + G, ... of row r; the G partial sums of a row sit in
//            consecutive lanes of one wave and are added with xor shuffles;
//            lane g = 0 forms u'_r, writes it to dst and to LDS;
//   s-phase: thread c accumulates sum_r A[r, c] u'_r over the chunk.
// Two work-group barriers per chunk.
constexpr int KPF = 17;   // prefetched floats per thread (n * cr <= KPF * BT)

// Thread t loads row r = t mod cr, columns c0 + k G (c0 = t / cr).  Only
// lanes whose (row, column) lies inside A_i issue a load.
__device__ __forceinline__ void load_chunk(__amdgpu_buffer_rsrc_t rs, int m, int n, int lda,
                                           int lcr, int r0, float (&pf)[KPF]) {
  const int t = threadIdx.x, cr = 1 << lcr;
  const int r = t & (cr - 1), c0 = t >> lcr, G = BT >> lcr;
  const bool row_ok = r0 + r < m;
  const int voff = (c0 * lda + r0 + r) * 4;
  const unsigned sstep = (unsigned)(G * lda) * 4u;
#pragma unroll
  for (int k = 0; k < KPF; ++k) {
    float v = 0.f;
    // an issued load has c = c0 + k G < n, so its byte offset is below the
    // problem's extent (< 2^31 by the host check); the scalar offset is not
    // covered by the buffer's range check
    if (row_ok && c0 + k * G < n)
      v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, (int)(k * sstep), 0));
    pf[k] = v;
  }
}

__device__ __forceinline__ void stage_chunk(float* buf, int n, int lcr, const float (&pf)[KPF]) {
  const int t = threadIdx.x, cr = 1 << lcr;
  const int r = t & (cr - 1), c0 = t >> lcr, G = BT >> lcr, ld = cr + 1;
#pragma unroll
  for (int k = 0; k < KPF; ++k) {
    const int c = c0 + k * G;
    if (c < n) buf[c * ld + r] = pf[k];
  }
}

__device__ double fused_pass(const float* A, int m, int n, int lda, int lcr, const SolveLds& L,
                             double c1, const double* src, double c2, double* dst) {
  const int t = threadIdx.x;
  const int cr = 1 << lcr, lg = 8 - lcr, G = 1 << lg;   // BT = 256 = cr * G
  const int ld = cr + 1;
  const int tr = t >> lg, tg = t & (G - 1);              // t-phase: row, column group
  float* buf[2] = {L.chunk, L.chunk + n * ld};
  double* ucb[2] = {L.uc, L.tpart};                     // u' of a chunk (<= 256 rows each)
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void*)A, (short)0, ((n - 1) * lda + m) * 4, 0x00020000);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, q = 0.0;
  float pf[KPF];
  load_chunk(rs, m, n, lda, lcr, 0, pf);
  __syncthreads();                                       // L.y complete, buffers free
  stage_chunk(buf[0], n, lcr, pf);
  double ps = (tg == 0 && tr < m) ? src[tr] : 0.0;
  if (cr < m) load_chunk(rs, m, n, lda, lcr, cr, pf);
  int x = 0;
  for (int r0 = 0; r0 < m; r0 += cr, x ^= 1) {
    const int rows = min(cr, m - r0);
    __syncthreads();                                     // chunk r0 staged in buf[x]
    // t-phase
    const float* cb = buf[x];
    double a0 = 0.0, a1 = 0.0;
    int c = tg;
    for (; c + G < n; c += 2 * G) {
      a0 += (double)cb[c * ld + tr] * L.y[c];
      a1 += (double)cb[(c + G) * ld + tr] * L.y[c + G];
    }
    if (c < n) a0 += (double)cb[c * ld + tr] * L.y[c];
    double tsum = a0 + a1;
    for (int o = 1; o < G; o <<= 1) tsum += __shfl_xor(tsum, o, 64);
    if (tg == 0) {
      double uv = 0.0;
      if (tr < rows) {
        uv = c1 * tsum + c2 * ps;
        dst[r0 + tr] = uv;
        q += uv * uv;
      }
      ucb[x][tr] = uv;
    }
    // next chunk: stage the prefetched registers, start loading the one after
    const int rn = r0 + cr;
    if (rn < m) {
      stage_chunk(buf[x ^ 1], n, lcr, pf);
      ps = (tg == 0 && rn + tr < m) ? src[rn + tr] : 0.0;
      if (rn + cr < m) load_chunk(rs, m, n, lda, lcr, rn + cr, pf);
    }
    __syncthreads();                                     // u' of chunk r0 in ucb[x]
    // s-phase
    if (t < n) {
      const float* col = cb + t * ld;
      const double* uc = ucb[x];
      int r = 0;
      for (; r + 3 < rows; r += 4) {
        s0 += (double)col[r] * uc[r];
        s1 += (double)col[r + 1] * uc[r + 1];
        s2 += (double)col[r + 2] * uc[r + 2];
        s3 += (double)col[r + 3] * uc[r + 3];
      }
      for (; r < rows; ++r) s0 += (double)col[r] * uc[r];
    }
  }
  if (t < n) L.s[t] = (s0 + s1) + (s2 + s3);
  return bsum(q, L.red);
}

__device__ void write_fail(const Args& a, const Prob& P, int p, int status) {
  for (int i = threadIdx.x; i < P.n; i += BT) a.x[P.x_off + i] = 0.0;
  if (threadIdx.x == 0) {
    Info in{};
    in.status = status; in.d = P.d;
    a.info[p] = in;
  }
}

__device__ double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// LSQR scalar state (src/solvers.rs:115-278 with damp = 0).
struct LsqrSt {
  double alfa, ubeta, rhobar, phibar, anorm, acond, ddnorm, res2, xnorm, xxnorm, zz, cs2, sn2,
      arnorm, r2norm, bnorm;
  int itn, istop;
};

__device__ void solve_item(const Args& a, const Prob& P, int p, unsigned char* lds_raw) {
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int n = P.n, N = n + 1, m = P.m, d = P.d;
  __shared__ int s_flag;
  // Wait for the sketch items of this problem.
  if (t == 0) {
    unsigned spins = 0;
    int ok = 1;
    // relaxed polling (an acquire would invalidate the L2 on every poll);
    // the fence after the barrier below is the acquire
    while (__hip_atomic_load(&a.ctl[2 + p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < P.nrb) {
      if (++spins > WAIT_SPINS) { ok = 0; atomicExch(&a.ctl[1], 1); break; }
      __builtin_amdgcn_s_sleep(8);
    }
    s_flag = ok;
  }
  __syncthreads();
  __threadfence();
  if (!s_flag) { write_fail(a, P, p, 2); return; }

  SolveLds L;
  double* vb = (double*)lds_raw;
  L.z = vb; L.v = vb + VSTR; L.w = vb + 2 * VSTR; L.y = vb + 3 * VSTR; L.s = vb + 4 * VSTR;
  L.row = vb + 5 * VSTR; L.tpart = vb + 6 * VSTR; L.uc = vb + 7 * VSTR;
  L.red = vb + NVEC * VSTR;
  L.chunk = (float*)(vb + NVEC * VSTR + 16);

  // 1. Augmented Gram [S A | S b]^T [S A | S b] (upper tiles), FP64 MFMA.
  const float* SA = a.wsf + P.sa_off;
  double* Gm = a.wsd + P.g_off;
  const int nb = (N + 15) >> 4;
  for (int tile = w; tile < nb * (nb + 1) / 2; tile += BT / 64) {
    int bi = 0, rem = tile;
    while (rem >= nb - bi) { rem -= nb - bi; ++bi; }
    const int bj = bi + rem;
    const int ci = bi * 16 + (l & 15), cj = bj * 16 + (l & 15);
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < d; k0 += 4) {
      const int k = k0 + (l >> 4);
      const double av = (k < d && ci < N) ? (double)SA[(int64_t)ci * P.ldsa + k] : 0.0;
      const double bv = (k < d && cj < N) ? (double)SA[(int64_t)cj * P.ldsa + k] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = bi * 16 + 4 * v + (l >> 4), j = bj * 16 + (l & 15);
      if (i < N && j < N && i <= j) Gm[i + (int64_t)j * N] = acc[v];
    }
  }
  __syncthreads();
  // 2. Cholesky of the leading n x n block, right-looking; the trailing
  //    updates also run over column n, which ends as R^{-T} g = z0.
  for (int j = 0; j < n; ++j) {
    if (t == 0) {
      const double piv = Gm[j + (int64_t)j * N];
      s_flag = (piv > 0.0 && piv < INFINITY) ? 1 : 0;
      if (s_flag) Gm[j + (int64_t)j * N] = sqrt(piv);
    }
    __syncthreads();
    if (!s_flag) break;
    const double inv = 1.0 / Gm[j + (int64_t)j * N];
    for (int c = j + 1 + t; c < N; c += BT) {
      const double v = Gm[j + (int64_t)c * N] * inv;
      Gm[j + (int64_t)c * N] = v;
      L.row[c] = v;
    }
    __syncthreads();
    for (int c = j + 1 + w; c < N; c += BT / 64) {
      const double rc = L.row[c];
      const int iend = min(c, n - 1);
      for (int i = j + 1 + l; i <= iend; i += 64) Gm[i + (int64_t)c * N] -= L.row[i] * rc;
    }
    __syncthreads();
  }
  if (!s_flag) { write_fail(a, P, p, 1); return; }
  // 3. X = R^{-1} (upper), rows from the bottom; thread c owns column c.
  double* X = a.wsd + P.xi_off;
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + t; k < n; k += BT) L.row[k] = Gm[i + (int64_t)k * N];
    __syncthreads();
    const double rii = L.row[i];
    for (int c = i + t; c < n; c += BT) {
      double v;
      if (c == i) {
        v = 1.0 / rii;
      } else {
        double s = 0.0;
        const double* col = X + (int64_t)c * n;
        for (int k = i + 1; k <= c; ++k) s += L.row[k] * col[k];
        v = -s / rii;
      }
      X[i + (int64_t)c * n] = v;
    }
    __syncthreads();
  }
  // 4. LSQR on M = A X from z0 (src/solvers.rs:115-278 with x0, damp = 0).
  const float* Ap = a.A + P.a_off;
  const double* bp = a.b + P.b_off;
  double* u = a.wsd + P.u_off;
  for (int i = t; i < n; i += BT) L.z[i] = Gm[i + (int64_t)n * N];
  double bb = 0.0;
  for (int r = t; r < m; r += BT) bb += bp[r] * bp[r];
  const double bnorm = sqrt(bsum(bb, L.red));
  __syncthreads();
  x_apply(X, n, L.z, L.y);
  // Chunk rows: the largest power of two (16..256) with n cr within the
  // prefetch registers and two n (cr + 1) buffers within LDS, halved rshift
  // times (the tuning knob).
  int lcr = 4;
  while (lcr < 8 && n * (2 << lcr) <= KPF * BT && 2 * n * ((2 << lcr) + 1) <= CHUNK_FLOATS) ++lcr;
  lcr = max(4, lcr - a.rshift);
  double beta2 = fused_pass(Ap, m, n, P.lda, lcr, L, -1.0, bp, 1.0, u);
  double beta = sqrt(beta2);
  const double ibeta = beta > 0.0 ? 1.0 / beta : 0.0;
  for (int i = t; i < n; i += BT) L.row[i] = L.s[i] * ibeta;   // A^T u
  __syncthreads();
  xt_apply(X, n, L.row, L.v);
  double vv = 0.0;
  for (int i = t; i < n; i += BT) vv += L.v[i] * L.v[i];
  double alfa = sqrt(bsum(vv, L.red));
  {
    const double ia = alfa > 0.0 ? 1.0 / alfa : 0.0;
    for (int i = t; i < n; i += BT) { L.v[i] *= ia; L.w[i] = L.v[i]; }
  }
  // The scalar state lives in LDS so that nothing is held in registers
  // across the pass over A; every thread evaluates the recurrences on the
  // same values and thread 0 stores them.
  __shared__ LsqrSt S;
  if (t == 0) {
    LsqrSt st{};
    st.alfa = alfa; st.ubeta = beta;          // u holds beta * u_normalised
    st.rhobar = alfa; st.phibar = beta; st.r2norm = beta; st.bnorm = bnorm;
    st.cs2 = -1.0; st.arnorm = alfa * beta;
    st.istop = 0; st.itn = 0;
    S = st;
  }
  const double eps = 2.220446049250313e-16;
  const double ctol = a.conlim > 0.0 ? 1.0 / a.conlim : 0.0;
  const double atol = a.tol, btol = a.tol;
  for (;;) {
    __syncthreads();
    if (S.arnorm == 0.0 || S.istop != 0 || S.itn >= a.iter_lim) break;
    x_apply(X, n, L.v, L.y);
    const double c2 = S.ubeta > 0.0 ? -S.alfa / S.ubeta : 0.0;
    beta2 = fused_pass(Ap, m, n, P.lda, lcr, L, 1.0, u, c2, u);
    LsqrSt st = S;
    st.itn += 1;
    beta = sqrt(beta2);
    st.ubeta = beta;
    double ww = 0.0;
    for (int i = t; i < n; i += BT) ww += L.w[i] * L.w[i];
    if (beta > 0.0) {
      st.anorm = sqrt(st.anorm * st.anorm + st.alfa * st.alfa + beta * beta);
      const double ib = 1.0 / beta;
      for (int i = t; i < n; i += BT) L.row[i] = L.s[i] * ib;
      __syncthreads();
      xt_apply(X, n, L.row, L.tpart);   // X^T A^T u into tpart (n <= BT)
      double q = 0.0;
      for (int i = t; i < n; i += BT) {
        const double vn = L.tpart[i] - beta * L.v[i];
        L.v[i] = vn;
        q += vn * vn;
      }
      st.alfa = sqrt(bsum(q, L.red));
      const double ia = st.alfa > 0.0 ? 1.0 / st.alfa : 0.0;
      for (int i = t; i < n; i += BT) L.v[i] *= ia;
    }
    ww = bsum(ww, L.red);
    const double alfa_n = st.alfa;
    // damp = 0: rhobar1 = |rhobar|, psi = 0
    const double rhobar1 = sqrt(st.rhobar * st.rhobar);
    const double cs1 = st.rhobar / rhobar1;
    st.phibar *= cs1;
    double cs, sn, rho;
    {
      const double aa = rhobar1, b2 = beta;
      if (b2 == 0.0) { cs = sgn(aa); sn = 0.0; rho = fabs(aa); }
      else if (aa == 0.0) { cs = 0.0; sn = sgn(b2); rho = fabs(b2); }
      else if (fabs(b2) > fabs(aa)) {
        const double tau = aa / b2;
        sn = sgn(b2) / sqrt(1.0 + tau * tau);
        cs = sn * tau;
        rho = b2 / sn;
      } else {
        const double tau = b2 / aa;
        cs = sgn(aa) / sqrt(1.0 + tau * tau);
        sn = cs * tau;
        rho = aa / cs;
      }
    }
    const double theta = sn * alfa_n;
    st.rhobar = -cs * alfa_n;
    const double phi = cs * st.phibar;
    st.phibar *= sn;
    const double tau = sn * phi;
    const double t1 = phi / rho, t2 = -theta / rho;
    for (int i = t; i < n; i += BT) {
      const double wi = L.w[i];
      L.z[i] += t1 * wi;
      L.w[i] = L.v[i] + t2 * wi;
    }
    st.ddnorm += ww / (rho * rho);
    const double delta = st.sn2 * rho;
    const double gambar = -st.cs2 * rho;
    const double rhs = phi - delta * st.zz;
    const double zbar = rhs / gambar;
    st.xnorm = sqrt(st.xxnorm + zbar * zbar);
    const double gamma = sqrt(gambar * gambar + theta * theta);
    st.cs2 = gambar / gamma;
    st.sn2 = theta / gamma;
    st.zz = rhs / gamma;
    st.xxnorm += st.zz * st.zz;
    st.acond = st.anorm * sqrt(st.ddnorm);
    const double rnorm = sqrt(st.phibar * st.phibar + st.res2);
    st.arnorm = alfa_n * fabs(tau);
    st.r2norm = rnorm;
    const double test1 = rnorm / st.bnorm;
    const double test2 = st.arnorm / (st.anorm * rnorm + eps);
    const double test3 = 1.0 / (st.acond + eps);
    const double tt1 = test1 / (1.0 + st.anorm * st.xnorm / st.bnorm);
    const double rtol = atol + btol * (st.anorm * st.xnorm / st.bnorm);
    int istop = 0;
    if (st.itn >= a.iter_lim) istop = 7;
    if (1.0 + test3 <= 1.0) istop = 6;
    if (1.0 + test2 <= 1.0) istop = 5;
    if (1.0 + tt1 <= 1.0) istop = 4;
    if (test3 <= ctol) istop = 3;
    if (test2 <= atol) istop = 2;
    if (test1 <= rtol) istop = 1;
    st.istop = istop;
    __syncthreads();                 // every thread has read S
    if (t == 0) S = st;
  }
  // 5. x = X z
  x_apply(X, n, L.z, L.y);
  __syncthreads();
  for (int i = t; i < n; i += BT) a.x[P.x_off + i] = L.y[i];
  if (t == 0) {
    Info in;
    in.status = 0; in.istop = S.istop; in.itn = S.itn; in.d = d;
    in.r2norm = S.r2norm; in.arnorm = S.arnorm; in.anorm = S.anorm; in.acond = S.acond;
    in.xnorm = S.xnorm;
    a.info[p] = in;
  }
}

// Wave-uniform copies (SGPRs) of values that are the same in every lane.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ Prob uniform(Prob P) {
  P.m = uni(P.m); P.n = uni(P.n); P.d = uni(P.d); P.lda = uni(P.lda);
  P.nrb = uni(P.nrb); P.ldsa = uni(P.ldsa);
  P.a_off = uni(P.a_off); P.b_off = uni(P.b_off); P.x_off = uni(P.x_off); P.sid = uni(P.sid);
  P.sa_off = uni(P.sa_off); P.g_off = uni(P.g_off); P.xi_off = uni(P.xi_off);
  P.u_off = uni(P.u_off);
  return P;
}

__global__ __launch_bounds__(BT, 2) void batch_ls_kernel(const Args* __restrict__ ap) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ int s_item;
  const Args a = *ap;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(&a.ctl[0], 1);
    __syncthreads();
    // the item index is the same in every lane: readfirstlane makes it (and
    // every per-problem value derived from it) scalar, so problem sizes,
    // pointers and buffer descriptors stay in SGPRs
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= a.n_items) break;
    const int2 item = a.items[it];
    const int pi = uni(item.x), rb = uni(item.y);
    const Prob P = uniform(a.probs[pi]);
    if (rb >= 0) sketch_item(a, P, pi, rb, (float*)lds);
    else solve_item(a, P, pi, lds);
  }
}

__global__ void zero_ctl_kernel(int* ctl, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ctl[i] = 0;
}

__global__ void set_io_kernel(Args* a, const float* A, const double* b, double* x) {
  a->A = A; a->b = b; a->x = x;
}

int64_t roundup(int64_t x, int64_t k) { return (x + k - 1) / k * k; }

}  // namespace

struct BatchPlan {
  int64_t nprob = 0;
  std::vector<Prob> probs;
  std::vector<int2> items;
  Args host_args{};
  Args* d_args = nullptr;
  Prob* d_probs = nullptr;
  int2* d_items = nullptr;
  int* d_ctl = nullptr;
  float* wsf = nullptr;
  double* wsd = nullptr;
  Info* d_info = nullptr;
  size_t sketch_lds = 0;
  int rshift = 0;          // LSQR rows-per-wave-step halvings
  int wgs_per_cu = 2;
  int cus = 0;
  double best_ms = -1.0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap_stream = nullptr;
};

namespace {

size_t solve_lds() { return (size_t)(NVEC * VSTR + 16) * 8 + (size_t)CHUNK_FLOATS * 4; }
size_t plan_lds(const BatchPlan* pl) { return std::max(pl->sketch_lds, solve_lds()); }

hipError_t destroy_graph(BatchPlan* pl) {
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  pl->exec = nullptr; pl->graph = nullptr;
  return hipSuccess;
}

// Capture [zero counters, persistent kernel] for the plan's current tile
// configuration.
hipError_t build_graph(BatchPlan* pl, hipStream_t st) {
  destroy_graph(pl);
  pl->host_args.rshift = pl->rshift;
  hipError_t e = hipMemcpyAsync(pl->d_args, &pl->host_args, sizeof(Args), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  const size_t lds = plan_lds(pl);
  e = hipFuncSetAttribute((const void*)batch_ls_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lds);
  if (e != hipSuccess) return e;
  int occ = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_ls_kernel, BT, lds);
  if (e != hipSuccess) return e;
  if (occ < 1) return hipErrorInvalidConfiguration;
  const int per_cu = std::min(pl->wgs_per_cu, occ);
  const int64_t want = (int64_t)per_cu * pl->cus;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)pl->items.size()));
  hipStream_t cs = pl->cap_stream;
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const int nctl = (int)(2 + pl->nprob);
  hipLaunchKernelGGL(zero_ctl_kernel, dim3((unsigned)std::min(64, (nctl + 255) / 256)), dim3(256), 0,
                     cs, pl->d_ctl, nctl);
  hipLaunchKernelGGL(batch_ls_kernel, dim3(grid), dim3(BT), lds, cs, (const Args*)pl->d_args);
  hipError_t le = hipGetLastError();
  e = hipStreamEndCapture(cs, &pl->graph);
  if (le != hipSuccess) return le;
  if (e != hipSuccess) return e;
  return hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0);
}

}  // namespace

hipError_t batch_plan_create(const BatchProbDesc* desc, int64_t count, double sf, double tol,
                             int64_t iter_lim, const Stream& s, hipStream_t st, BatchPlan** out) {
  *out = nullptr;
  BatchPlan* pl = new BatchPlan();
  pl->nprob = count;
  pl->probs.resize(count);
  int64_t fo = 0, dof = 0;
  int max_ncb = 1;
  for (int64_t i = 0; i < count; ++i) {
    const BatchProbDesc& q = desc[i];
    Prob& P = pl->probs[i];
    P.m = (int)q.m; P.n = (int)q.n; P.lda = (int)q.lda;
    const double dd = std::floor(sf * (double)q.n);
    P.d = (int)(dd > (double)q.m ? q.m : (int64_t)dd);
    P.nrb = (P.d + SK_ROWS - 1) / SK_ROWS;
    P.ldsa = (int)roundup(P.d, 64);
    P.a_off = q.a_off; P.b_off = q.b_off; P.x_off = q.x_off; P.sid = q.stream_id;
    const int64_t N = q.n + 1;
    P.sa_off = fo; fo += (int64_t)P.ldsa * N;
    P.g_off = dof; dof += roundup(N * N, 32);
    P.xi_off = dof; dof += roundup(q.n * q.n, 32);
    P.u_off = dof; dof += roundup(q.m, 32);
    max_ncb = std::max<int>(max_ncb, (int)((N + 15) / 16));
  }
  pl->sketch_lds = (size_t)max_ncb * 16 * SK_KS * 4;
  // Largest problems first: sketch items, then solve items.
  std::vector<int64_t> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
    const double cx = (double)pl->probs[x].m * pl->probs[x].n;
    const double cy = (double)pl->probs[y].m * pl->probs[y].n;
    return cx > cy;
  });
  // Longest solve first: each problem's sketch items, then its solve item, in
  // decreasing size, so the largest solves start at once instead of after
  // every sketch.  A solve item only waits for sketch items queued before it.
  for (int64_t i : order) {
    for (int rb = 0; rb < pl->probs[i].nrb; ++rb) pl->items.push_back(int2{(int)i, rb});
    pl->items.push_back(int2{(int)i, -1});
  }
  hipError_t e = hipSuccess;
  int dev = 0;
  (void)hipGetDevice(&dev);
  e = hipDeviceGetAttribute(&pl->cus, hipDeviceAttributeMultiprocessorCount, dev);
  auto alloc = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 256));
  };
  alloc((void**)&pl->d_args, sizeof(Args));
  alloc((void**)&pl->d_probs, sizeof(Prob) * count);
  alloc((void**)&pl->d_items, sizeof(int2) * pl->items.size());
  alloc((void**)&pl->d_ctl, sizeof(int) * (2 + count));
  alloc((void**)&pl->wsf, sizeof(float) * fo);
  alloc((void**)&pl->wsd, sizeof(double) * dof);
  alloc((void**)&pl->d_info, sizeof(Info) * count);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pl->d_probs, pl->probs.data(), sizeof(Prob) * count, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pl->d_items, pl->items.data(), sizeof(int2) * pl->items.size(),
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&pl->cap_stream, hipStreamNonBlocking);
  Args& a = pl->host_args;
  a.probs = pl->d_probs; a.items = pl->d_items;
  a.n_items = (int)pl->items.size(); a.nprob = (int)count;
  a.ctl = pl->d_ctl; a.wsf = pl->wsf; a.wsd = pl->wsd; a.info = pl->d_info;
  a.A = nullptr; a.b = nullptr; a.x = nullptr;
  a.key0 = (uint32_t)s.key[0]; a.key1 = (uint32_t)s.key[1];
  a.ctr0 = s.ctr[0]; a.ctr1 = s.ctr[1];
  a.tol = tol; a.conlim = 1e8;
  a.iter_lim = (int)iter_lim;
  a.rshift = pl->rshift;
  if (e == hipSuccess) e = build_graph(pl, st);
  if (e != hipSuccess) { batch_plan_destroy(pl); return e; }
  *out = pl;
  return hipSuccess;
}

hipError_t batch_plan_execute(BatchPlan* pl, const float* A, const double* b, double* x,
                              hipStream_t st) {
  hipLaunchKernelGGL(set_io_kernel, dim3(1), dim3(1), 0, st, pl->d_args, A, b, x);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipGraphLaunch(pl->exec, st);
}

hipError_t batch_plan_autotune(BatchPlan* pl, const float* A, const double* b, double* x,
                               hipStream_t st) {
  // Candidates: LSQR rows per wave step (as derived from n, or halved) x
  // work-groups per CU (up to the occupancy limit).
  const int shifts[] = {0, 1};
  const int per_cu[] = {1, 2, 3, 4};
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) return e;
  int best_s = pl->rshift, best_w = pl->wgs_per_cu;
  float best = -1.f;
  const size_t lds = plan_lds(pl);
  (void)hipFuncSetAttribute((const void*)batch_ls_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
  int occ = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_ls_kernel, BT, lds);
  for (int sh : shifts) {
    if (e != hipSuccess) break;
    for (int wpc : per_cu) {
      if (wpc > occ) continue;
      pl->rshift = sh; pl->wgs_per_cu = wpc;
      e = build_graph(pl, st);
      if (e != hipSuccess) break;
      e = batch_plan_execute(pl, A, b, x, st);   // warm-up
      if (e == hipSuccess) e = hipEventRecord(e0, st);
      for (int r = 0; r < 2 && e == hipSuccess; ++r) e = batch_plan_execute(pl, A, b, x, st);
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) break;
      ms *= 0.5f;
      if (best < 0.f || ms < best) { best = ms; best_s = sh; best_w = wpc; }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return e;
  pl->rshift = best_s; pl->wgs_per_cu = best_w; pl->best_ms = best;
  e = build_graph(pl, st);
  if (e != hipSuccess) return e;
  return batch_plan_execute(pl, A, b, x, st);
}

hipError_t batch_plan_info(BatchPlan* pl, void* info, bool info_on_device, int* err_flag,
                           hipStream_t st) {
  hipError_t e = hipSuccess;
  if (info)
    e = hipMemcpyAsync(info, pl->d_info, sizeof(Info) * pl->nprob,
                       info_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && err_flag)
    e = hipMemcpyAsync(err_flag, pl->d_ctl + 1, sizeof(int), hipMemcpyDeviceToHost, st);
  return e;
}

hipError_t batch_plan_sketch(BatchPlan* pl, float* out, hipStream_t st) {
  // [S_i A_i | S_i b_i] of every problem, d_i x (n_i + 1) column-major, packed
  // in problem order.
  int64_t o = 0;
  for (const Prob& P : pl->probs) {
    hipError_t e = hipMemcpy2DAsync(out + o, (size_t)P.d * 4, pl->wsf + P.sa_off, (size_t)P.ldsa * 4,
                                    (size_t)P.d * 4, (size_t)(P.n + 1), hipMemcpyDefault, st);
    if (e != hipSuccess) return e;
    o += (int64_t)P.d * (P.n + 1);
  }
  return hipSuccess;
}

void batch_plan_config(const BatchPlan* pl, int* rshift, int* wgs_per_cu, double* best_ms) {
  if (rshift) *rshift = pl->rshift;
  if (wgs_per_cu) *wgs_per_cu = pl->wgs_per_cu;
  if (best_ms) *best_ms = pl->best_ms;
}

int64_t batch_plan_count(const BatchPlan* pl) { return (int64_t)pl->nprob; }

int64_t batch_plan_sketch_size(const BatchPlan* pl) {
  int64_t o = 0;
  for (const Prob& P : pl->probs) o += (int64_t)P.d * (P.n + 1);
  return o;
}

void batch_plan_destroy(BatchPlan* pl) {
  if (!pl) return;
  destroy_graph(pl);
  if (pl->cap_stream) (void)hipStreamDestroy(pl->cap_stream);
  void* ptrs[] = {pl->d_args, pl->d_probs, pl->d_items, pl->d_ctl, pl->wsf, pl->wsd, pl->d_info};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete pl;
}

}  // namespace rna
