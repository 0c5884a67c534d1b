// Counter-based generators and the sketch-entry transforms built on them.
//
// Host and device share this header, so the library, its host-side helpers and
// the GPU kernels all produce the same integer word stream.
//
// Block functions follow the vendored rust-random123 crate:
//   Philox{2x32,2x64,4x32,4x64}-10  rust-random123/src/philox.rs:137-237
//   ThreeFry2x64-20                 rust-random123/src/threefry.rs:30-93
// Counter stepping follows `next()` in the same files (128-bit / 64-bit
// little-endian counter with carry, philox.rs:87-100, threefry.rs:13-21).
// Seeding follows rand_core 0.6.4 `SeedableRng::seed_from_u64` (PCG32
// expansion of the u64 into the seed bytes, read little-endian into the key,
// philox.rs:101-105, threefry.rs:23-27).
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define RNA_HD __host__ __device__ __forceinline__
#else
#define RNA_HD static inline
#endif

namespace rna {

// ---------------------------------------------------------------- constants
constexpr uint32_t PHILOX_M2X32 = 0xd256d193u;
constexpr uint32_t PHILOX_M4X32_0 = 0xD2511F53u;
constexpr uint32_t PHILOX_M4X32_1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W32_0 = 0x9E3779B9u;
constexpr uint32_t PHILOX_W32_1 = 0xBB67AE85u;
constexpr uint64_t PHILOX_M2X64 = 0xD2B74407B1CE6E93ull;
constexpr uint64_t PHILOX_M4X64_0 = 0xD2E7470EE14C6C93ull;
constexpr uint64_t PHILOX_M4X64_1 = 0xCA5A826395121157ull;
constexpr uint64_t PHILOX_W64_0 = 0x9E3779B97F4A7C15ull;
constexpr uint64_t PHILOX_W64_1 = 0xBB67AE8584CAA73Bull;
constexpr uint64_t SKEIN_PARITY = 0x1BD11BDAA9FC1A22ull;

struct u32x4 { uint32_t v[4]; };
struct u32x2 { uint32_t v[2]; };
struct u64x2 { uint64_t v[2]; };
struct u64x4 { uint64_t v[4]; };

RNA_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// ------------------------------------------------------------ Philox family
// Ten rounds; the key is bumped by the Weyl constants before rounds 1..9.
RNA_HD u32x4 philox4x32(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += PHILOX_W32_0; k1 += PHILOX_W32_1; }
    const uint64_t p0 = (uint64_t)PHILOX_M4X32_0 * c.v[0];
    const uint64_t p1 = (uint64_t)PHILOX_M4X32_1 * c.v[2];
    u32x4 o;
    o.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
    o.v[1] = (uint32_t)p1;
    o.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
    o.v[3] = (uint32_t)p0;
    c = o;
  }
  return c;
}

RNA_HD u32x2 philox2x32(u32x2 c, uint32_t k) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) k += PHILOX_W32_0;
    const uint64_t p = (uint64_t)PHILOX_M2X32 * c.v[0];
    u32x2 o;
    o.v[0] = (uint32_t)(p >> 32) ^ k ^ c.v[1];
    o.v[1] = (uint32_t)p;
    c = o;
  }
  return c;
}

RNA_HD u64x2 philox2x64(u64x2 c, uint64_t k) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) k += PHILOX_W64_0;
    u64x2 o;
    o.v[0] = mulhi64(PHILOX_M2X64, c.v[0]) ^ k ^ c.v[1];
    o.v[1] = PHILOX_M2X64 * c.v[0];
    c = o;
  }
  return c;
}

RNA_HD u64x4 philox4x64(u64x4 c, uint64_t k0, uint64_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += PHILOX_W64_0; k1 += PHILOX_W64_1; }
    u64x4 o;
    o.v[0] = mulhi64(PHILOX_M4X64_1, c.v[2]) ^ c.v[1] ^ k0;
    o.v[1] = PHILOX_M4X64_1 * c.v[2];
    o.v[2] = mulhi64(PHILOX_M4X64_0, c.v[0]) ^ c.v[3] ^ k1;
    o.v[3] = PHILOX_M4X64_0 * c.v[0];
    c = o;
  }
  return c;
}

// ---------------------------------------------------------- ThreeFry2x64-20
RNA_HD uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

RNA_HD u64x2 threefry2x64(u64x2 c, uint64_t k0, uint64_t k1) {
  // Rotation schedule: the first and second halves of each 8-round cycle.
  constexpr int R[8] = {16, 42, 12, 31, 16, 32, 24, 21};
  const uint64_t ks[3] = {k0, k1, SKEIN_PARITY ^ k0 ^ k1};
  uint64_t x0 = c.v[0] + ks[0];
  uint64_t x1 = c.v[1] + ks[1];
#pragma unroll
  for (int g = 0; g < 5; ++g) {        // five groups of four rounds
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      x0 += x1;
      x1 = rotl64(x1, R[(g & 1) * 4 + j]);
      x1 ^= x0;
    }
    const int s = g + 1;               // key injection after each group
    x0 += ks[s % 3];
    x1 += ks[(s + 1) % 3] + (uint64_t)s;
  }
  u64x2 o; o.v[0] = x0; o.v[1] = x1;
  return o;
}

// ------------------------------------------------------ rand_core seeding
// PCG32 step + XSH-RR output, as used by rand_core 0.6 `seed_from_u64`.
RNA_HD uint32_t pcg32_next(uint64_t& state) {
  state = state * 6364136223846793005ull + 11634580027462260723ull;
  const uint64_t s = state;
  const uint32_t xs = (uint32_t)(((s >> 18) ^ s) >> 27);
  const uint32_t rot = (uint32_t)(s >> 59);
  return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

// Fill `nwords` little-endian u32 seed words from a u64 seed.
RNA_HD void seed_words_from_u64(uint64_t seed, uint32_t* words, int nwords) {
  uint64_t st = seed;
  for (int i = 0; i < nwords; ++i) words[i] = pcg32_next(st);
}

// --------------------------------------------------------- generator handle
enum Gen : int { GEN_THREEFRY2X64 = 0, GEN_PHILOX4X32 = 1 };

// A stream is a generator, a key and a 128-bit starting counter.  Block `b`
// of the stream is the block function at counter ctr_base + b (128-bit add).
struct Stream {
  int gen;
  uint64_t key[2];      // Philox4x32 uses the low 32 bits of key[0], key[1]
  uint64_t ctr[2];      // starting counter: low and high 64 bits
};

// Two u64 words of block `b`.  For Philox4x32 the four u32 words are paired
// low word first, matching rand_core's BlockRng `next_u64` on an aligned index.
RNA_HD u64x2 block64(const Stream& s, uint64_t b) {
  uint64_t lo = s.ctr[0] + b;
  uint64_t hi = s.ctr[1] + (lo < b ? 1ull : 0ull);
  u64x2 r;
  if (s.gen == GEN_THREEFRY2X64) {
    u64x2 c; c.v[0] = lo; c.v[1] = hi;
    r = threefry2x64(c, s.key[0], s.key[1]);
  } else {
    u32x4 c;
    c.v[0] = (uint32_t)lo; c.v[1] = (uint32_t)(lo >> 32);
    c.v[2] = (uint32_t)hi; c.v[3] = (uint32_t)(hi >> 32);
    u32x4 o = philox4x32(c, (uint32_t)s.key[0], (uint32_t)s.key[1]);
    r.v[0] = (uint64_t)o.v[0] | ((uint64_t)o.v[1] << 32);
    r.v[1] = (uint64_t)o.v[2] | ((uint64_t)o.v[3] << 32);
  }
  return r;
}

// Word `w` of the u64 stream.
RNA_HD uint64_t word64(const Stream& s, uint64_t w) {
  u64x2 b = block64(s, w >> 1);
  return (w & 1) ? b.v[1] : b.v[0];
}

// Stream as the reference seeds it: `Rng::seed_from_u64(seed)`, counter 0.
RNA_HD Stream stream_from_seed(int gen, uint64_t seed) {
  Stream s;
  s.gen = gen;
  s.ctr[0] = 0; s.ctr[1] = 0;
  if (gen == GEN_THREEFRY2X64) {
    uint32_t w[4];
    seed_words_from_u64(seed, w, 4);
    s.key[0] = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    s.key[1] = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
  } else {
    uint32_t w[2];
    seed_words_from_u64(seed, w, 2);
    s.key[0] = w[0];
    s.key[1] = w[1];
  }
  return s;
}

// ------------------------------------------------------- entry transforms
// Every transform maps a (stream, entry index) pair to a value without any
// sequential state, so any tile of a sketch can be produced independently.
constexpr double TWO_POW_M53 = 1.1102230246251565e-16;  // 2^-53

RNA_HD double unit53(uint64_t w) { return (double)(w >> 11) * TWO_POW_M53; }

// Uniform(-1, 1): one u64 word per entry (word index == entry index).
RNA_HD double uniform_pm1(const Stream& s, uint64_t e) {
  return 2.0 * unit53(word64(s, e)) - 1.0;
}

// Rademacher: one bit per entry; bit (e & 63) of u64 word (e >> 6).
RNA_HD double rademacher(const Stream& s, uint64_t e) {
  const uint64_t w = word64(s, e >> 6);
  return ((w >> (e & 63)) & 1ull) ? 1.0 : -1.0;
}

// Standard normal by the Box-Muller transform (Box & Muller, 1958).  Entries
// 2p and 2p+1 share block p: u1 in (0,1] from word 0, u2 in [0,1) from word 1,
// z_{2p} = r cos(2 pi u2), z_{2p+1} = r sin(2 pi u2), r = sqrt(-2 ln u1).
RNA_HD void gauss_pair(const Stream& s, uint64_t p, double& z0, double& z1) {
  const u64x2 b = block64(s, p);
  const double u1 = (double)((b.v[0] >> 11) + 1) * TWO_POW_M53;
  const double u2 = unit53(b.v[1]);
  const double r = sqrt(-2.0 * log(u1));
  const double t = 6.283185307179586 * u2;
  z0 = r * cos(t);
  z1 = r * sin(t);
}

RNA_HD double gauss(const Stream& s, uint64_t e) {
  double z0, z1;
  gauss_pair(s, e >> 1, z0, z1);
  return (e & 1) ? z1 : z0;
}

// Sparse-sign column: `zeta` distinct rows in [0, d) with signs, for column
// j.  Rows come from the u32 words of blocks ctr_base + j*256 + t (t < 255):
// row = (word * d) >> 32, skipping rows already taken.  Signs are bit i of
// block j*256 + 255 (u32 word i/32).  Returns the number of rows produced
// (== zeta unless zeta > d).
constexpr int SPARSE_BLOCKS_PER_COL = 256;
constexpr int SPARSE_MAX_ZETA = 64;

RNA_HD int sparse_sign_column(const Stream& s, uint64_t j, int d, int zeta,
                              int* rows, float* signs) {
  if (zeta > d) zeta = d;
  const uint64_t base = j * (uint64_t)SPARSE_BLOCKS_PER_COL;
  int got = 0;
  for (int t = 0; t < SPARSE_BLOCKS_PER_COL - 1 && got < zeta; ++t) {
    const u64x2 b = block64(s, base + (uint64_t)t);
    const uint32_t w[4] = {(uint32_t)b.v[0], (uint32_t)(b.v[0] >> 32),
                           (uint32_t)b.v[1], (uint32_t)(b.v[1] >> 32)};
    for (int q = 0; q < 4 && got < zeta; ++q) {
      const int r = (int)(((uint64_t)w[q] * (uint64_t)d) >> 32);
      bool dup = false;
      for (int i = 0; i < got; ++i) dup |= (rows[i] == r);
      if (!dup) rows[got++] = r;
    }
  }
  // Exhausted budget (only reachable when zeta is close to d): take the
  // smallest unused rows so the column is always complete.
  for (int r = 0; got < zeta && r < d; ++r) {
    bool dup = false;
    for (int i = 0; i < got; ++i) dup |= (rows[i] == r);
    if (!dup) rows[got++] = r;
  }
  const u64x2 sb = block64(s, base + (uint64_t)(SPARSE_BLOCKS_PER_COL - 1));
  for (int i = 0; i < got; ++i) {
    const uint64_t word = (i < 64) ? sb.v[0] : sb.v[1];
    signs[i] = ((word >> (i & 63)) & 1ull) ? 1.0f : -1.0f;
  }
  return got;
}

// The same column (identical rows, order and signs) with rows and signs in
// arrays of compile-time size ZMAX >= zeta that are indexed only by
// compile-time constants, so a GPU thread keeps them in registers.
template <int ZMAX>
RNA_HD void sparse_sign_take(int (&rows)[ZMAX], int& got, int r) {
  bool dup = false;
#pragma unroll
  for (int i = 0; i < ZMAX; ++i) dup |= (i < got && rows[i] == r);
  if (!dup) {
#pragma unroll
    for (int i = 0; i < ZMAX; ++i) rows[i] = (i == got) ? r : rows[i];
    ++got;
  }
}

template <int ZMAX>
RNA_HD int sparse_sign_column_fixed(const Stream& s, uint64_t j, int d, int zeta,
                                    int (&rows)[ZMAX], float (&signs)[ZMAX]) {
  if (zeta > d) zeta = d;
  const uint64_t base = j * (uint64_t)SPARSE_BLOCKS_PER_COL;
  int got = 0;
#pragma unroll
  for (int i = 0; i < ZMAX; ++i) rows[i] = 0;
  for (int t = 0; t < SPARSE_BLOCKS_PER_COL - 1 && got < zeta; ++t) {
    const u64x2 b = block64(s, base + (uint64_t)t);
    const uint32_t w[4] = {(uint32_t)b.v[0], (uint32_t)(b.v[0] >> 32),
                           (uint32_t)b.v[1], (uint32_t)(b.v[1] >> 32)};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (got < zeta) sparse_sign_take<ZMAX>(rows, got, (int)(((uint64_t)w[q] * (uint64_t)d) >> 32));
  }
  for (int r = 0; got < zeta && r < d; ++r) sparse_sign_take<ZMAX>(rows, got, r);
  const u64x2 sb = block64(s, base + (uint64_t)(SPARSE_BLOCKS_PER_COL - 1));
#pragma unroll
  for (int i = 0; i < ZMAX; ++i) {
    const uint64_t word = (i < 64) ? sb.v[0] : sb.v[1];
    signs[i] = ((word >> (i & 63)) & 1ull) ? 1.0f : -1.0f;
  }
  return got;
}

}  // namespace rna
