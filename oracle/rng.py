"""Counter-based RNG oracle (test infrastructure only, see ``oracle/__init__``).

Restates rust-random123 (vendored in the reference) and rand_core 0.6.4's
``seed_from_u64``:

* ``philox4x32``     rust-random123/src/philox.rs:150-154 (round), :174-176
                     (key bump), :211-223 (10 rounds).  Pinned: KAT
                     philox.rs:268-279 / :330-345.
* ``philox2x32``     philox.rs:138-141, :164-166, :183-195. Pinned: KAT :242-253.
* ``philox2x64``     philox.rs:144-147, :169-171, :197-209. Pinned: KAT :255-266.
* ``philox4x64``     philox.rs:157-161, :179-181, :225-237. Pinned: KAT :281-292.
* ``threefry2x64``   rust-random123/src/threefry.rs:30-93.  Pinned: KAT
                     threefry.rs:97-127.
* ``seed_from_u64``  rand_core 0.6.4 (third-party, Cargo.lock:890-891; not in
                     the reference tree): PCG32 expansion, words read
                     little-endian into the key (philox.rs:101-105,
                     threefry.rs:23-27).  Pinned: the commented-out vector
                     threefry.rs:140-141 (seed 42 -> first u64
                     391376552519608501).
* ``word64`` stream  BlockRng64 order (threefry.rs:13-21: ctr[0]++ with
                     carry).  For Philox4x32 the pairing of u32 words into u64
                     (low word first) is the library's convention: the
                     reference has no test of it -> parity unpinned beyond the
                     block function itself.

Entry transforms (Gaussian by Box-Muller, Uniform(-1,1), Rademacher,
sparse-sign) are the library's own definitions: the reference draws values
through rand_distr 0.4.3 (ziggurat) / rand 0.8.5, third-party crates absent
from the tree with no value-level test (src/sketch.rs:113-126) -> parity
unpinned; pinned instead by distribution tests and host == device agreement.
"""
import numpy as np

M64 = (1 << 64) - 1
M32 = (1 << 32) - 1

PHILOX_M2X32 = 0xD256D193
PHILOX_M4X32_0 = 0xD2511F53
PHILOX_M4X32_1 = 0xCD9E8D57
PHILOX_W32_0 = 0x9E3779B9
PHILOX_W32_1 = 0xBB67AE85
PHILOX_M2X64 = 0xD2B74407B1CE6E93
PHILOX_M4X64_0 = 0xD2E7470EE14C6C93
PHILOX_M4X64_1 = 0xCA5A826395121157
PHILOX_W64_0 = 0x9E3779B97F4A7C15
PHILOX_W64_1 = 0xBB67AE8584CAA73B
SKEIN_PARITY = 0x1BD11BDAA9FC1A22

THREEFRY2X64 = 0
PHILOX4X32 = 1

_U32 = np.uint64(0xFFFFFFFF)


def philox4x32(ctr, key):
    """Vectorised Philox4x32-10.  ``ctr``: (..., 4) uint32-valued uint64 array
    (or list), ``key``: 2 ints.  philox.rs:150-154, :174-176, :211-223."""
    c = [np.asarray(x, dtype=np.uint64) for x in np.moveaxis(np.asarray(ctr, dtype=np.uint64), -1, 0)]
    k0, k1 = np.uint64(key[0] & M32), np.uint64(key[1] & M32)
    m0, m1 = np.uint64(PHILOX_M4X32_0), np.uint64(PHILOX_M4X32_1)
    for r in range(10):
        if r:
            k0 = np.uint64((int(k0) + PHILOX_W32_0) & M32)
            k1 = np.uint64((int(k1) + PHILOX_W32_1) & M32)
        p0 = m0 * c[0]            # < 2^64, exact
        p1 = m1 * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & _U32,
             (p0 >> np.uint64(32)) ^ c[3] ^ k1, p0 & _U32]
    return np.stack(c, axis=-1)


def philox2x32(ctr, key):
    """Philox2x32-10 on Python ints (KAT use).  philox.rs:138-141, :183-195."""
    c0, c1 = ctr
    k = key & M32
    for r in range(10):
        if r:
            k = (k + PHILOX_W32_0) & M32
        p = PHILOX_M2X32 * c0
        c0, c1 = ((p >> 32) ^ k ^ c1) & M32, p & M32
    return [c0, c1]


def philox2x64(ctr, key):
    """Philox2x64-10 on Python ints.  philox.rs:144-147, :197-209."""
    c0, c1 = ctr
    k = key & M64
    for r in range(10):
        if r:
            k = (k + PHILOX_W64_0) & M64
        p = PHILOX_M2X64 * c0
        c0, c1 = ((p >> 64) ^ k ^ c1) & M64, p & M64
    return [c0, c1]


def philox4x64(ctr, key):
    """Philox4x64-10 on Python ints.  philox.rs:157-161, :225-237."""
    c = list(ctr)
    k0, k1 = key[0] & M64, key[1] & M64
    for r in range(10):
        if r:
            k0 = (k0 + PHILOX_W64_0) & M64
            k1 = (k1 + PHILOX_W64_1) & M64
        p0 = PHILOX_M4X64_0 * c[0]
        p1 = PHILOX_M4X64_1 * c[2]
        c = [((p1 >> 64) ^ c[1] ^ k0) & M64, p1 & M64,
             ((p0 >> 64) ^ c[3] ^ k1) & M64, p0 & M64]
    return c


_ROT = (16, 42, 12, 31, 16, 32, 24, 21)


def _rotl(x, r):
    return (x << np.uint64(r)) | (x >> np.uint64(64 - r))


def threefry2x64(ctr, key):
    """Vectorised ThreeFry2x64-20.  ``ctr``: (..., 2) uint64.
    threefry.rs:30-93 (20 rounds as 5 groups of 4, key injection after each
    group, injection index s adds s to word 1)."""
    ctr = np.asarray(ctr, dtype=np.uint64)
    k0, k1 = key[0] & M64, key[1] & M64
    ks = [np.uint64(k0), np.uint64(k1), np.uint64(SKEIN_PARITY ^ k0 ^ k1)]
    with np.errstate(over="ignore"):
        x0 = ctr[..., 0] + ks[0]
        x1 = ctr[..., 1] + ks[1]
        for g in range(5):
            for j in range(4):
                x0 = x0 + x1
                x1 = _rotl(x1, _ROT[(g & 1) * 4 + j])
                x1 = x1 ^ x0
            s = g + 1
            x0 = x0 + ks[s % 3]
            x1 = x1 + ks[(s + 1) % 3] + np.uint64(s)
    return np.stack([x0, x1], axis=-1)


def pcg32_words(seed, n):
    """rand_core 0.6.4 seed_from_u64: PCG32 (XSH-RR) words."""
    st = seed & M64
    out = []
    for _ in range(n):
        st = (st * 6364136223846793005 + 11634580027462260723) & M64
        xs = (((st >> 18) ^ st) >> 27) & M32
        rot = st >> 59
        out.append(((xs >> rot) | (xs << ((32 - rot) & 31))) & M32)
    return out


class Stream:
    """Generator + key + 128-bit starting counter.  Block b of the stream is
    the block function at counter ctr + b (as the reference's ``next()``
    steps it: philox.rs:87-100, threefry.rs:13-21)."""

    def __init__(self, gen, key, ctr=(0, 0)):
        self.gen = gen
        self.key = (int(key[0]), int(key[1]))
        self.ctr = (int(ctr[0]), int(ctr[1]))

    @classmethod
    def from_seed(cls, gen, seed):
        """``Rng::seed_from_u64(seed)`` (threefry.rs:23-27, philox.rs:101-105)."""
        if gen == THREEFRY2X64:
            w = pcg32_words(seed, 4)
            return cls(gen, (w[0] | (w[1] << 32), w[2] | (w[3] << 32)))
        w = pcg32_words(seed, 2)
        return cls(gen, (w[0], w[1]))

    def blocks(self, b):
        """Two u64 words for each block index in ``b`` -> (len, 2) uint64."""
        b = np.asarray(b, dtype=np.uint64).reshape(-1)
        lo = b + np.uint64(self.ctr[0])
        carry = (lo < b).astype(np.uint64)
        hi = np.uint64(self.ctr[1]) + carry
        if self.gen == THREEFRY2X64:
            return threefry2x64(np.stack([lo, hi], axis=-1), self.key)
        c = np.stack([lo & _U32, lo >> np.uint64(32), hi & _U32, hi >> np.uint64(32)], axis=-1)
        o = philox4x32(c, self.key)
        return np.stack([o[:, 0] | (o[:, 1] << np.uint64(32)),
                         o[:, 2] | (o[:, 3] << np.uint64(32))], axis=-1)

    def words64(self, first, n):
        """u64 words [first, first+n) of the stream (BlockRng64 order)."""
        w = np.arange(first, first + n, dtype=np.uint64)
        blk = self.blocks(w >> np.uint64(1))
        return np.where((w & np.uint64(1)) == 1, blk[:, 1], blk[:, 0])


_TWO_M53 = 2.0 ** -53


def _unit53(w):
    return (w >> np.uint64(11)).astype(np.float64) * _TWO_M53


def gaussian(stream, count):
    """Entries 0..count-1: Box-Muller on block p = e // 2 (entries 2p, 2p+1)."""
    npairs = (count + 1) // 2
    blk = stream.blocks(np.arange(npairs, dtype=np.uint64))
    u1 = ((blk[:, 0] >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * _TWO_M53
    u2 = _unit53(blk[:, 1])
    r = np.sqrt(-2.0 * np.log(u1))
    t = 6.283185307179586 * u2
    z = np.empty(2 * npairs)
    z[0::2] = r * np.cos(t)
    z[1::2] = r * np.sin(t)
    return z[:count]


def uniform_pm1(stream, count):
    """Uniform(-1, 1): entry e = 2 * unit53(word e) - 1."""
    return 2.0 * _unit53(stream.words64(0, count)) - 1.0


def rademacher(stream, count):
    """+-1: bit (e & 63) of u64 word e >> 6."""
    e = np.arange(count, dtype=np.uint64)
    nw = (count + 63) // 64
    w = stream.words64(0, nw)
    bits = (w[e >> np.uint64(6)] >> (e & np.uint64(63))) & np.uint64(1)
    return np.where(bits == 1, 1.0, -1.0)


BLOCKS_PER_COL = 256


def sparse_sign_columns(stream, d, ncols, zeta):
    """Rows (ncols, zeta) int and signs (ncols, zeta) for a d x ncols
    sparse-sign operator (values +-1/sqrt(zeta)).  Column j draws rows from
    the u32 words of blocks j*256 + t: row = (w * d) >> 32, duplicates
    skipped; sign i is bit i of block j*256 + 255.  Pure-Python loop over
    columns: small cases only."""
    z = min(zeta, d)
    rows = np.zeros((ncols, z), dtype=np.int64)
    signs = np.zeros((ncols, z))
    for j in range(ncols):
        base = j * BLOCKS_PER_COL
        got = []
        t = 0
        while len(got) < z and t < BLOCKS_PER_COL - 1:
            b = stream.blocks([base + t])[0]
            ws = [int(b[0]) & M32, int(b[0]) >> 32, int(b[1]) & M32, int(b[1]) >> 32]
            for w in ws:
                if len(got) >= z:
                    break
                r = (w * d) >> 32
                if r not in got:
                    got.append(r)
            t += 1
        r = 0
        while len(got) < z:
            if r not in got:
                got.append(r)
            r += 1
        sb = stream.blocks([base + BLOCKS_PER_COL - 1])[0]
        rows[j] = got
        for i in range(z):
            word = int(sb[0]) if i < 64 else int(sb[1])
            signs[j, i] = 1.0 if (word >> (i & 63)) & 1 else -1.0
    return rows, signs


def sketching_operator(dist, rows, cols, stream, zeta=8):
    """Dense rows x cols sketch, column-major entry order e = i + j*rows.
    Follows src/sketch.rs:102-130 (column-major fill from one stream; zero
    dimension is an error) with the library's entry transforms."""
    if rows == 0 or cols == 0:
        raise ValueError("Rows and columns must be greater than 0")
    n = rows * cols
    if dist == 0:
        v = gaussian(stream, n)
    elif dist == 1:
        v = uniform_pm1(stream, n)
    elif dist == 2:
        v = rademacher(stream, n)
    elif dist == 3:
        r, s = sparse_sign_columns(stream, rows, cols, zeta)
        out = np.zeros((rows, cols))
        sc = 1.0 / np.sqrt(r.shape[1])
        for j in range(cols):
            out[r[j], j] = s[j] * sc
        return out
    else:
        raise ValueError("unknown distribution")
    return v.reshape(cols, rows).T.copy()


def batch_gauss_sketch(stream, sid, d, m):
    """S (d x m, float64) of the batched least-squares path for substream
    ``sid``: entry (r, k) comes from Philox4x32 block blk = r*ceil(m/4) + k//4
    at 128-bit counter ctr + blk + ((sid + 1) << 96); the block's words
    (w0, w1, w2, w3) give Box-Muller pairs (w0, w1) and (w2, w3) with
    u = (2 (w >> 9) + 1) 2^-24 and theta = (w' >> 8) 2^-24 turns:
    g = sqrt(-2 ln u) (cos 2 pi theta, sin 2 pi theta), here in float64 (the
    GPU evaluates the same formula in FP32).  This transform is the
    library's own; the reference draws no batched sketches."""
    m4 = (m + 3) // 4
    blk = np.arange(d * m4, dtype=np.uint64)
    lo = blk + np.uint64(stream.ctr[0])
    carry = (lo < blk).astype(np.uint64)
    add = np.uint64(((sid + 1) << 32) & M64)
    with np.errstate(over="ignore"):
        hi = np.uint64(stream.ctr[1]) + carry + add
    c = np.stack([lo & _U32, lo >> np.uint64(32), hi & _U32, hi >> np.uint64(32)], axis=-1)
    o = philox4x32(c, stream.key)
    two24 = 2.0 ** -24
    u1 = (2 * (o[:, 0] >> np.uint64(9)) + 1).astype(np.float64) * two24
    t1 = (o[:, 1] >> np.uint64(8)).astype(np.float64) * two24
    u2 = (2 * (o[:, 2] >> np.uint64(9)) + 1).astype(np.float64) * two24
    t2 = (o[:, 3] >> np.uint64(8)).astype(np.float64) * two24
    r1 = np.sqrt(-2.0 * np.log(u1))
    r2 = np.sqrt(-2.0 * np.log(u2))
    tp = 6.283185307179586
    g = np.stack([r1 * np.cos(tp * t1), r1 * np.sin(tp * t1),
                  r2 * np.cos(tp * t2), r2 * np.sin(tp * t2)], axis=-1)
    return g.reshape(d, 4 * m4)[:, :m]
