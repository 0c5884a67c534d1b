// C ABI (include/randnla.h) and host orchestration of the sketch-and-factor
// path.  Each entry point mirrors one function of the reference crate
// `randblas` 0.1.0; the citation is on the entry point.
//
// Host-side work here is orchestration only: argument checks, buffer
// staging, kernel launches, and reading back a few status words.  Every
// floating-point operation on matrix data runs in the gfx950 kernels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/randnla.h"
#include "iterative.h"
#include "kernels.h"
#include "rng.h"

using namespace rna;

// ------------------------------------------------------------------ context
struct TimedRange {
  hipEvent_t a, b;
};
struct TimerClass {
  std::vector<TimedRange> pending;
  double total_ms = 0.0;
  int64_t launches = 0;
};

struct randnla_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  bool timing = false;
  std::map<std::string, TimerClass> timers;
  // last-call diagnostics (read by tests / bench through randnla_diag)
  double diag[16] = {0};
};

namespace {

struct Fail {
  int code;
  std::string msg;
};

int set_err(randnla_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

#define HIPCK(x)                                                                     \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      throw Fail{RANDNLA_HIP_ERROR, std::string(#x) + ": " + hipGetErrorString(e_)}; \
  } while (0)

[[noreturn]] void fail(int code, const std::string& m) { throw Fail{code, m}; }

}  // namespace

// ---- stream-ordered workspace cache ----------------------------------------
// Blocks released on a stream are kept, keyed by (stream, rounded size), and
// handed out again to later requests of that size on the same stream: the
// stream orders the reuse after every earlier use, exactly as the runtime's
// stream-ordered pool would, without its per-call cost (which reaches
// hundreds of microseconds after a stream synchronisation).  At most
// WS_CACHE_MAX bytes are kept; randnla_destroy returns a context's blocks.
namespace {
constexpr size_t WS_CACHE_MAX = size_t(32) << 30;
struct WsCache {
  std::mutex mu;
  std::unordered_map<void*, size_t> live;                          // block -> bytes
  std::map<std::pair<hipStream_t, size_t>, std::vector<void*>> idle;
  size_t idle_bytes = 0;
};
WsCache& wsc() {
  static WsCache* c = new WsCache();   // never destroyed: blocks may outlive static teardown
  return *c;
}
size_t ws_round(size_t b) { return ((b ? b : 16) + 255) & ~size_t(255); }
}  // namespace

hipError_t rna::ws_alloc(void** p, size_t bytes, hipStream_t st) {
  bytes = ws_round(bytes);
  WsCache& c = wsc();
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.idle.find({st, bytes});
    if (it != c.idle.end() && !it->second.empty()) {
      *p = it->second.back();
      it->second.pop_back();
      c.idle_bytes -= bytes;
      c.live[*p] = bytes;
      return hipSuccess;
    }
  }
  hipError_t e = hipMallocAsync(p, bytes, st);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    ws_flush_all();                    // return every idle block, then retry once
    (void)hipStreamSynchronize(st);
    e = hipMallocAsync(p, bytes, st);
  }
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(c.mu);
  c.live[*p] = bytes;
  return hipSuccess;
}

void rna::ws_free(void* p, hipStream_t st) {
  if (!p) return;
  WsCache& c = wsc();
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.live.find(p);
    if (it != c.live.end()) {
      const size_t b = it->second;
      c.live.erase(it);
      if (c.idle_bytes + b <= WS_CACHE_MAX) {
        c.idle[{st, b}].push_back(p);
        c.idle_bytes += b;
        return;
      }
    }
  }
  (void)hipFreeAsync(p, st);
}

namespace {
void ws_flush_if(bool all, hipStream_t st) {
  WsCache& c = wsc();
  std::lock_guard<std::mutex> g(c.mu);
  for (auto it = c.idle.begin(); it != c.idle.end();) {
    if (all || it->first.first == st) {
      for (void* p : it->second) (void)hipFreeAsync(p, it->first.first);
      c.idle_bytes -= it->first.second * it->second.size();
      it = c.idle.erase(it);
    } else {
      ++it;
    }
  }
}
}  // namespace

void rna::ws_flush(hipStream_t st) { ws_flush_if(false, st); }
void rna::ws_flush_all() { ws_flush_if(true, nullptr); }

// ---- small device -> host reads through pinned staging ----------------------
namespace {
std::mutex stage_mu;
void* stage_buf = nullptr;
int live_contexts = 0;   // guarded by stage_mu; the last randnla_destroy frees stage_buf
hipError_t stage_get() {
  if (stage_buf) return hipSuccess;
  return hipHostMalloc(&stage_buf, D2H_STAGE_BYTES, hipHostMallocDefault);
}
}  // namespace

hipError_t rna::d2h_small(hipStream_t st, void* dst, const void* src, size_t bytes) {
  if (bytes > D2H_STAGE_BYTES) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> g(stage_mu);
  hipError_t e = stage_get();
  if (e == hipSuccess) e = hipMemcpyAsync(stage_buf, src, bytes, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) memcpy(dst, stage_buf, bytes);
  return e;
}

hipError_t rna::d2h_small_2d(hipStream_t st, void* dst, size_t dpitch, const void* src,
                             size_t spitch, size_t width, size_t height) {
  if (width * height > D2H_STAGE_BYTES) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> g(stage_mu);
  hipError_t e = stage_get();
  if (e == hipSuccess)
    e = hipMemcpy2DAsync(stage_buf, width, src, spitch, width, height, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess)
    for (size_t r = 0; r < height; ++r)
      memcpy((char*)dst + r * dpitch, (const char*)stage_buf + r * width, width);
  return e;
}

// ---- fixed scratch of the no-allocation entry point -------------------------
// 64 MiB of device memory per GPU, reserved with the code object (no runtime
// allocation), for the partial sums of randnla_sketch_apply_noalloc.  Host
// code never uses the array's name as an address: hipGetSymbolAddress gives
// each device's own copy, looked up once per device at context creation.
namespace {
constexpr int64_t NOALLOC_SCRATCH_DOUBLES = int64_t(1) << 23;
__device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_ptr[MAX_DEVICES] = {nullptr};

// The current device's copy of `scratch` (looked up on first use).
hipError_t scratch_lookup(int device, double** out) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratch_ptr[device] = (double*)p;
  }
  *out = scratch_ptr[device];
  return hipSuccess;
}
}  // namespace

namespace {

// RAII device allocation on the context stream (through the workspace cache).
struct DBuf {
  void* p = nullptr;
  hipStream_t st = nullptr;
  DBuf() = default;
  DBuf(size_t bytes, hipStream_t s) : st(s) { HIPCK(ws_alloc(&p, bytes, s)); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), st(o.st) { o.p = nullptr; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; st = o.st; o.p = nullptr; }
    return *this;
  }
  void release() {
    if (p) { ws_free(p, st); p = nullptr; }
  }
  ~DBuf() { release(); }
  template <typename X>
  X* as() const { return (X*)p; }
};

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Device -> host copies into caller or pageable memory.  Small ones go
// through a pinned staging buffer (asynchronous copy on the stream, stream
// drained, then a host memcpy); larger ones drain the stream and use a
// blocking copy, so the host sees the data when the call returns (an
// asynchronous copy into pageable memory may still be in flight after the
// stream reports idle).
void d2h(hipStream_t st, void* dst, const void* src, size_t bytes) {
  if (bytes <= D2H_STAGE_BYTES) { HIPCK(d2h_small(st, dst, src, bytes)); return; }
  HIPCK(hipStreamSynchronize(st));
  HIPCK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}
void d2h_2d(hipStream_t st, void* dst, size_t dpitch, const void* src, size_t spitch,
            size_t width, size_t height) {
  if (width * height <= D2H_STAGE_BYTES) {
    HIPCK(d2h_small_2d(st, dst, dpitch, src, spitch, width, height));
    return;
  }
  HIPCK(hipStreamSynchronize(st));
  HIPCK(hipMemcpy2D(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToHost));
}
// Diagonal of a device matrix (n entries, leading dimension ld) on the host.
std::vector<double> read_diag(hipStream_t st, const double* M, int64_t n, int64_t ld);

// A matrix argument: either the caller's device buffer or a staged copy.
struct MatIn {
  const void* p = nullptr;
  int64_t ld = 0;
  DBuf own;
  MatIn(const void* src, int64_t rows, int64_t cols, int64_t ld_src, size_t esz, hipStream_t st) {
    if (is_device_ptr(src) || rows * cols == 0) {
      p = src;
      ld = ld_src;
      return;
    }
    own = DBuf((size_t)rows * cols * esz, st);
    HIPCK(hipMemcpy2DAsync(own.p, rows * esz, src, ld_src * esz, rows * esz, cols,
                           hipMemcpyHostToDevice, st));
    p = own.p;
    ld = rows;
  }
  const double* as_d() const { return (const double*)p; }
};

struct MatOut {
  void* user;
  void* p = nullptr;
  int64_t ld = 0, rows, cols, ld_user;
  size_t esz;
  DBuf own;
  hipStream_t st;
  MatOut(void* dst, int64_t r, int64_t c, int64_t ldu, size_t e, hipStream_t s)
      : user(dst), rows(r), cols(c), ld_user(ldu), esz(e), st(s) {
    if (!dst) return;
    if (is_device_ptr(dst)) {
      p = dst;
      ld = ldu;
      return;
    }
    own = DBuf((size_t)std::max<int64_t>(1, r * c) * e, s);
    p = own.p;
    ld = r;
  }
  void commit() {
    if (user && own.p && rows * cols > 0)
      d2h_2d(st, user, ld_user * esz, own.p, rows * esz, rows * esz, cols);
  }
  template <typename X>
  X* as() const { return (X*)p; }
};

// Per-kernel-class timing with events on the context stream.
struct Tm {
  randnla_ctx* c;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  Tm(randnla_ctx* cc, const char* n) : c(cc), name(n) {
    if (!c->timing) return;
    HIPCK(hipEventCreate(&a));
    HIPCK(hipEventCreate(&b));
    HIPCK(hipEventRecord(a, c->stream));
  }
  ~Tm() {
    if (!a) return;
    if (hipEventRecord(b, c->stream) == hipSuccess)
      c->timers[name].pending.push_back({a, b});
  }
};

void resolve_timers(randnla_ctx* c) {
  for (auto& kv : c->timers) {
    for (auto& r : kv.second.pending) {
      (void)hipEventSynchronize(r.b);
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
        kv.second.total_ms += ms;
        kv.second.launches += 1;
      }
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    kv.second.pending.clear();
  }
}

Stream to_stream(const randnla_rng* r, int default_gen, uint64_t default_seed) {
  if (!r) return stream_from_seed(default_gen, default_seed);
  Stream s;
  s.gen = r->gen;
  s.key[0] = r->key[0];
  s.key[1] = r->key[1];
  s.ctr[0] = r->ctr[0];
  s.ctr[1] = r->ctr[1];
  return s;
}

randnla_opts opts_or_default(const randnla_opts* o) {
  randnla_opts d;
  randnla_opts_default(&d);
  return o ? *o : d;
}

int64_t r8(int64_t x) { return (x + 7) / 8 * 8; }

void sync(randnla_ctx* c) { HIPCK(hipStreamSynchronize(c->stream)); }

template <typename F>
int guard(randnla_ctx* c, F&& f) {
  if (!c) return RANDNLA_INVALID_PARAMETERS;
  c->err.clear();
  try {
    HIPCK(hipSetDevice(c->device));
    f();
    return RANDNLA_OK;
  } catch (const Fail& e) {
    (void)hipStreamSynchronize(c->stream);
    return set_err(c, e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return set_err(c, RANDNLA_COMPUTATION_ERROR, "host allocation failed");
  } catch (...) {
    return set_err(c, RANDNLA_COMPUTATION_ERROR, "unexpected exception");
  }
}

std::vector<double> read_diag(hipStream_t st, const double* M, int64_t n, int64_t ld) {
  DBuf dd((size_t)std::max<int64_t>(n, 1) * 8, st);
  HIPCK(hipMemcpy2DAsync(dd.p, 8, M, (size_t)(ld + 1) * 8, 8, n, hipMemcpyDeviceToDevice, st));
  std::vector<double> h(n);
  d2h(st, h.data(), dd.p, (size_t)n * 8);
  return h;
}

// ------------------------------------------------------------ GEMM helpers
int choose_splits(int64_t tiles, int64_t K, int64_t kmin) {
  int64_t s = (512 + tiles - 1) / tiles;
  const int64_t mx = std::max<int64_t>(1, K / kmin);
  if (s > mx) s = mx;
  if (s > 256) s = 256;
  return (int)std::max<int64_t>(1, s);
}

// C (M x N) = alpha * op(A) * B, FP64 MFMA, with automatic split-K for long K.
void gemm64(randnla_ctx* c, int64_t M, int64_t N, int64_t K, const void* A, T ta, int64_t lda,
            bool at, const void* B, T tb, int64_t ldb, void* C, T tc, int64_t ldc,
            double alpha = 1.0, const char* tname = "gemm_f64", int tri = 0) {
  Tm tm(c, tname);
  GemmArgs g;
  g.tri = tri;
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.ta = ta; g.lda = lda; g.a_trans = at;
  g.B = B; g.tb = tb; g.ldb = ldb;
  g.alpha = alpha;
  const int64_t mt = (M + 127) / 128, nt = (N + 63) / 64;
  int64_t tiles = mt * nt;
  if (tri == 1) {   // only tiles that reach the upper triangle compute
    tiles = 0;
    for (int64_t j = 0; j < nt; ++j) tiles += std::min(mt, (64 * j + 64 + 127) / 128);
  }
  const int splits = choose_splits(tiles, K, 128);
  if (splits <= 1) {
    g.C = C; g.tc = tc; g.ldc = ldc;
    HIPCK(gemm_f64(g, c->stream));
    return;
  }
  DBuf part((size_t)splits * M * N * sizeof(double), c->stream);
  g.C = part.p; g.tc = T::F64; g.ldc = M; g.cstride = M * N; g.splits = splits;
  HIPCK(gemm_f64(g, c->stream));
  if (tc == T::F64) {
    HIPCK(reduce_partials_f64(part.as<double>(), M, N, M, M * N, splits, (double*)C, ldc, 1.0,
                              c->stream));
  } else {
    DBuf tmp((size_t)M * N * sizeof(double), c->stream);
    HIPCK(reduce_partials_f64(part.as<double>(), M, N, M, M * N, splits, tmp.as<double>(), M, 1.0,
                              c->stream));
    HIPCK(convert(tmp.p, T::F64, M, N, M, C, tc, ldc, 1.0, c->stream));
  }
}

// ------------------------------------------------------ orthonormalisation
// Gram of Y (m x l) in FP64.
void gram64(randnla_ctx* c, const void* Y, T ty, int64_t m, int64_t l, int64_t ldy, double* G) {
  if (syrk_supported(ty, l, ldy, Y)) {
    Tm tm(c, "gram");
    DBuf part(syrk_workspace(m, l, ty), c->stream);
    HIPCK(syrk_f64(Y, ty, m, l, ldy, G, l, part.as<double>(), c->stream));
    return;
  }
  gemm64(c, l, l, m, Y, ty, ldy, true, Y, ty, ldy, G, T::F64, l, 1.0, "gram", 1);
}

// One Cholesky-QR pass: Q = Y R^{-1} with R = chol(Y^T Y).  Dependent
// columns (pivot <= tau * original diagonal) get a zero row in R and are
// replaced by e_j in Q.  R (l x l, FP64) is returned through Rout if given.
// Everything stays on the stream (no host synchronisation).
void cholqr_pass(randnla_ctx* c, const void* Y, T ty, int64_t m, int64_t l, int64_t ldy, void* Q,
                 T tq, int64_t ldq, double* Rout, double tau) {
  DBuf G((size_t)l * l * sizeof(double), c->stream);
  DBuf X((size_t)l * l * sizeof(double), c->stream);
  DBuf flags((size_t)(l + 1) * sizeof(int), c->stream);
  gram64(c, Y, ty, m, l, ldy, G.as<double>());
  {
    Tm tm(c, "cholesky");
    HIPCK(cholesky_inverse_upper(G.as<double>(), l, l, tau, flags.as<int>(), flags.as<int>() + l,
                                 X.as<double>(), l, c->stream));
  }
  gemm64(c, m, l, l, Y, ty, ldy, false, X.p, T::F64, l, Q, tq, ldq, 1.0, "trmm", 2);
  HIPCK(complete_columns(Q, tq, m, l, ldq, flags.as<int>(), c->stream));
  if (Rout) HIPCK(hipMemcpyAsync(Rout, G.p, (size_t)l * l * 8, hipMemcpyDeviceToDevice, c->stream));
}

// One-sided Jacobi stops when every column pair has |cos| <= sqrt(n) eps
// (the convergence threshold of LAPACK's one-sided Jacobi SVD).
double jacobi_tol(int64_t n) {
  return std::sqrt((double)std::max<int64_t>(n, 1)) * 2.220446049250313e-16;
}

double default_tau(int64_t l) { return 16.0 * (double)std::max<int64_t>(l, 16) * 2.220446049250313e-16; }

// CholeskyQR2 of Y (m x l): Q (tq) and, if Rout, R = R2 R1 (FP64 l x l).
void cholqr2(randnla_ctx* c, const void* Y, T ty, int64_t m, int64_t l, int64_t ldy, void* Q,
             T tq, int64_t ldq, double* Rout) {
  DBuf Q1((size_t)m * l * sizeof(double), c->stream);
  DBuf R1((size_t)l * l * sizeof(double), c->stream);
  DBuf R2((size_t)l * l * sizeof(double), c->stream);
  const double tau = default_tau(l);
  cholqr_pass(c, Y, ty, m, l, ldy, Q1.p, T::F64, m, R1.as<double>(), tau);
  cholqr_pass(c, Q1.p, T::F64, m, l, m, Q, tq, ldq, R2.as<double>(), tau);
  if (Rout)
    gemm64(c, l, l, l, R2.p, T::F64, l, false, R1.p, T::F64, l, Rout, T::F64, l, 1.0, "rprod");
}

// ------------------------------------------------------- rand_svd tail
// Given an orthonormal Q (m x l, type tq) with B^T = A^T Q available as Bt
// (n x l FP64), produce U (m x r), S (r), Vt (r x n).
void svd_tail(randnla_ctx* c, const void* Q, T tq, int64_t ldq, int64_t m, int64_t n, int64_t l,
              const double* Bt, int64_t r, double* U, int64_t ldu, double* S, double* Vt,
              int64_t ldvt) {
  hipStream_t st = c->stream;
  // B^T = Qb Rb  ->  B = Rb^T Qb^T ; SVD(Rb^T) = Ur S Vr^T ; B = Ur S (Qb Vr)^T
  DBuf Qb((size_t)n * l * 8, st), Rb((size_t)l * l * 8, st);
  cholqr2(c, Bt, T::F64, n, l, n, Qb.p, T::F64, n, Rb.as<double>());
  DBuf W((size_t)l * l * 8, st), Vr((size_t)l * l * 8, st), sv((size_t)l * 8, st);
  HIPCK(transpose_f64(Rb.as<double>(), l, l, l, W.as<double>(), l, st));
  HIPCK(set_identity(Vr.as<double>(), l, l, l, st));
  DBuf cnt(sizeof(int), st);
  int sweeps = 0;
  int hist[30] = {0};
  {
    Tm tm(c, "jacobi");
    // An FP32 basis Q (mixed precision) is orthonormal to ~1e-7: rotations
    // below 1e-8 change nothing it can carry (measured on A 65536 x 16384:
    // the same 9 sweeps as at 1e-10, fewer rotations in the last ones, the
    // same error to 8 digits); FP64 bases keep sqrt(l) eps.
    const double tol = tq == T::F32 ? std::max(jacobi_tol(l), 1e-8) : jacobi_tol(l);
    HIPCK(jacobi_svd(W.as<double>(), l, l, Vr.as<double>(), l, sv.as<double>(), cnt.as<int>(), 30,
                     tol, st, &sweeps, hist));
  }
  c->diag[0] = sweeps;
  for (int i = 0; i < 15; ++i) c->diag[1 + i] = hist[i];   // rotations per sweep
  DBuf Ur((size_t)l * l * 8, st), Vs((size_t)l * l * 8, st), ss((size_t)l * 8, st);
  HIPCK(svd_finalize(W.as<double>(), l, l, Vr.as<double>(), l, sv.as<double>(), Ur.as<double>(), l,
                     Vs.as<double>(), l, ss.as<double>(), st));
  // Zero singular values give placeholder columns e_j in Ur: re-orthonormalise.
  std::vector<double> hs(l);
  d2h(st, hs.data(), ss.p, l * 8);
  sync(c);
  if (hs[l - 1] == 0.0) {
    DBuf Ur2((size_t)l * l * 8, st);
    cholqr_pass(c, Ur.p, T::F64, l, l, l, Ur2.p, T::F64, l, nullptr, default_tau(l));
    HIPCK(hipMemcpyAsync(Ur.p, Ur2.p, (size_t)l * l * 8, hipMemcpyDeviceToDevice, st));
  }
  // U = Q Ur[:, :r]
  gemm64(c, m, r, l, Q, tq, ldq, false, Ur.p, T::F64, l, U, T::F64, ldu, 1.0, "svd_u");
  // V = Qb Vs[:, :r], Vt = V^T
  DBuf V((size_t)n * r * 8, st);
  gemm64(c, n, r, l, Qb.p, T::F64, n, false, Vs.p, T::F64, l, V.p, T::F64, n, 1.0, "svd_v");
  HIPCK(transpose_f64(V.as<double>(), n, r, n, Vt, ldvt, st));
  HIPCK(hipMemcpyAsync(S, ss.p, r * 8, hipMemcpyDeviceToDevice, st));
}

// Power-scheme range finder, FP64 products.  Returns Q (m x l, FP64 ld m).
void range_finder_f64(randnla_ctx* c, const void* A, T ta, int64_t m, int64_t n, int64_t lda,
                      int64_t l, int q, const Stream& s, double* Q) {
  hipStream_t st = c->stream;
  DBuf Y((size_t)m * l * 8, st), Z((size_t)n * l * 8, st);
  // Y = A * Omega, Omega generated inside the product when that is cheap,
  // else as one panel (the two give identical values).
  const double regen = (double)((m + 127) / 128) * (double)n * (double)l;
  {
    Tm tm(c, "sketch_gemm");
    if (regen <= 6.4e7) {
      HIPCK(dense_sketch_apply(1, 0, s, A, ta, m, n, lda, l, Y.as<double>(), m, nullptr, 0, st));
    } else {
      DBuf Om((size_t)n * l * 8, st);
      HIPCK(sketch_panel(0, s, n, l, Om.p, T::F64, n, 1.0, st));
      GemmArgs g;
      g.M = m; g.N = l; g.K = n; g.A = A; g.ta = ta; g.lda = lda; g.a_trans = false;
      g.B = Om.p; g.tb = T::F64; g.ldb = n; g.C = Y.p; g.tc = T::F64; g.ldc = m;
      HIPCK(gemm_f64(g, st));
    }
  }
  const double tau = default_tau(l);
  if (q == 0) {
    cholqr2(c, Y.p, T::F64, m, l, m, Q, T::F64, m, nullptr);
    return;
  }
  DBuf Qt((size_t)m * l * 8, st), Qz((size_t)n * l * 8, st);
  cholqr_pass(c, Y.p, T::F64, m, l, m, Qt.p, T::F64, m, nullptr, tau);
  for (int it = 0; it < q; ++it) {
    gemm64(c, n, l, m, A, ta, lda, true, Qt.p, T::F64, m, Z.p, T::F64, n, 1.0, "pass_at");
    cholqr_pass(c, Z.p, T::F64, n, l, n, Qz.p, T::F64, n, nullptr, tau);
    gemm64(c, m, l, n, A, ta, lda, false, Qz.p, T::F64, n, Y.p, T::F64, m, 1.0, "pass_a");
    if (it + 1 < q) cholqr_pass(c, Y.p, T::F64, m, l, m, Qt.p, T::F64, m, nullptr, tau);
  }
  cholqr2(c, Y.p, T::F64, m, l, m, Q, T::F64, m, nullptr);
}

// Power of two s with max|x| * s close to 2^14 (FP16 split scaling).
float pow2_scale(double amax) {
  if (!(amax > 0.0) || !isfinite(amax)) return 1.f;
  const int e = (int)ceil(log2(amax));
  return ldexpf(1.f, 14 - e);
}

// FP8 operands: A (FP32 m x n, column-major) -> row-major (Arm, ld n rounded
// up to 128) and column-major (Acm, ld m rounded up to 128) FP8 copies with
// 128 x 128 block scales SA; optionally a panel X (K x N) -> X8 (ld K rounded
// up to 128; X has leading dimension ldx) with per-column 128-K block scales SX.  Buffers sized by the
// caller; padding is written as zeros.
void fp8_operands(randnla_ctx* c, const float* A, int64_t m, int64_t n, int64_t lda,
                  uint8_t* Arm, uint8_t* Acm, uint8_t* SA, const float* X, int64_t K,
                  int64_t N, int64_t ldx, uint8_t* X8, uint8_t* SX) {
  Tm tm(c, "quantize_fp8");
  hipStream_t st = c->stream;
  const int64_t mp = (m + 127) / 128 * 128, np = (n + 127) / 128 * 128;
  if (A) HIPCK(quantize_fp8_a(A, m, n, lda, Arm, np, Acm, mp, SA, st));
  if (X) HIPCK(quantize_fp8_panel(X, K, N, ldx, X8, (K + 127) / 128 * 128, SX, st));
}

// Rounds of (A^T, A) at the end of the mixed-precision ladder that run on the
// FP32 data.  Measured on A 65536 x 16384 (sigma_i = (1+i)^-2 plus 1e-7
// noise, k = 256, p = 64, q = 3) against the FP64 path on the same bits:
// |e - e_FP64| = 8.2e-7 (BF16 start) and 1.3e-6 (FP8 start) with one clean
// round, 3.6e-7 and 7.7e-7 with two split into hi/lo, 3.5e-7 and 7.7e-7 with
// the first of the two rounding A once to FP16 instead.
constexpr int kCleanRounds = 2;

// Mixed-precision range finder + B^T in split FP16.  A is FP32 (m x n).
// Precision ladder: Y = A Omega (BF16, or FP8-e4m3 with block scales when
// fp8) and the first q - kCleanRounds rounds of (A^T, A) products run on a
// BF16 copy of A; the last kCleanRounds rounds run on the FP32 data, A rounded
// once to FP16 (scaled) in all but the last and split into FP16 hi/lo in the
// last, so that the final basis is computed from A itself (the BF16 rounding
// of A is a perturbation of spectral norm ~1e-5 ||A||, the size of the
// trailing singular values it would otherwise pollute; the FP16 rounding is
// 8x smaller, and the last round contracts what it leaves).  B^T = A^T Q is
// split FP16 with Q split as well.  Returns Q32 (mp x l, FP32) and Bt (n x l
// FP64).
void rand_svd_mixed_core(randnla_ctx* c, const float* A, int64_t m, int64_t n, int64_t lda,
                         int64_t l, int q, const Stream& s, float* Q32, int64_t mp, double* Bt,
                         bool fp8) {
  hipStream_t st = c->stream;
  const int64_t np = r8(n);
  // FP32 operand with 8-aligned shape and leading dimension for the split
  // kernels (zero-padded copy only when A itself does not qualify).
  const float* Aop = A;
  int64_t ldop = lda;
  DBuf Apad;
  if (lda % 8 != 0 || m % 8 != 0 || n % 8 != 0) {
    Apad = DBuf((size_t)mp * np * 4, st);
    HIPCK(hipMemsetAsync(Apad.p, 0, (size_t)mp * np * 4, st));
    HIPCK(convert(A, T::F32, m, n, lda, Apad.p, T::F32, mp, 1.0, st));
    Aop = Apad.as<float>();
    ldop = mp;
  }
  DBuf amax(16, st);
  // max |A| (the split-FP16 scale of A): the ladders get it from the pass
  // that makes the BF16 copy, q = 0 reads A once for it.
  if (q == 0) HIPCK(absmax(Aop, T::F32, mp, np, ldop, amax.as<double>(), st));
  // max |X| of a split-GEMM panel into amax[1]; the kernel derives both
  // power-of-two split scales from amax on the device (no host round trip).
  auto panel_amax = [&](const float* X, int64_t rows, int64_t cols, int64_t ldx) {
    HIPCK(absmax(X, T::F32, rows, cols, ldx, amax.as<double>() + 1, st));
  };

  DBuf Y((size_t)mp * l * 4, st), Z((size_t)np * l * 4, st), Qz32((size_t)np * l * 4, st);
  // Clean products inside the iteration run A split into FP16 hi/lo against
  // the panel rounded to FP16 (mode 2, two products): the panel's rounding
  // perturbs the iterate only.  B^T = A^T Q keeps Q's low part (mode 1).
  // With the row-major copy At (np x mp column-major) the product A X reads
  // K-contiguous rows of A like A^T X does (no transposition on the way into
  // LDS); without it, the M-contiguous A is transposed in the kernel.
  DBuf At;
  // The panel's FP16 copy, or hi and lo planes for B (made inside gemm_lp,
  // loaded straight into LDS).
  DBuf Bh((size_t)2 * l * std::max(mp, np) * 2, st);
  // mode 2 (hi and lo of A), or mode 3 (A rounded once to FP16) for the
  // first of the clean rounds.
  auto pass_a_split = [&](const float* X, int mode) {  // Y = A X, split FP16
    panel_amax(X, np, l, np);
    Tm tm(c, mode == 2 ? "gemm_f16x2" : "gemm_f16");
    LpGemmArgs g;
    g.mode = mode; g.M = mp; g.N = l; g.K = np;
    if (At.p) { g.A = At.p; g.lda = np; g.a_trans = true; }
    else { g.A = Aop; g.lda = ldop; g.a_trans = false; }
    g.B = X; g.ldb = np; g.C = Y.as<float>(); g.ldc = mp; g.cstride = 0;
    g.amax = amax.as<double>();
    g.bh = Bh.as<uint16_t>(); g.ldbh = np;
    HIPCK(gemm_lp(g, st));
  };
  auto pass_at_split = [&](const float* X, int mode) {  // Z = A^T X, split FP16
    panel_amax(X, mp, l, mp);
    const int64_t tiles = (np + 255) / 256;
    const int splits = choose_splits(tiles, mp, 2048);
    DBuf part((size_t)splits * np * l * 4, st);
    {
      Tm tm(c, mode == 2 ? "gemm_f16x2" : "gemm_f16");
      LpGemmArgs g;
      g.mode = mode; g.M = np; g.N = l; g.K = mp; g.A = Aop; g.lda = ldop; g.a_trans = true;
      g.B = X; g.ldb = mp; g.C = part.as<float>(); g.ldc = np; g.cstride = np * l;
      g.splits = splits; g.amax = amax.as<double>();
      g.bh = Bh.as<uint16_t>(); g.ldbh = mp;
      HIPCK(gemm_lp(g, st));
    }
    HIPCK(reduce_partials(part.as<float>(), np, l, np, np * l, splits, Z.p, T::F32, np, 1.0, st));
  };
  const double tau = default_tau(l);
  if (q == 0) {
    HIPCK(hipMemsetAsync(Qz32.p, 0, (size_t)np * l * 4, st));
    HIPCK(sketch_panel(0, s, n, l, Qz32.p, T::F32, np, 1.0, st));
    pass_a_split(Qz32.as<float>(), 2);
  } else {
    // BF16 copy of A (mp x np, zero padded) for the middle rounds and the
    // row-major FP32 copy for the clean rounds' A X, in one pass over A.
    DBuf A16((size_t)mp * np * 2, st);
    At = DBuf((size_t)np * mp * 4, st);
    {
      Tm tm(c, "convert_bf16");
      HIPCK(convert_bf16_transpose_absmax(Aop, mp, np, ldop, A16.as<uint16_t>(), mp,
                                          At.as<float>(), np, amax.as<double>(), st));
    }
    DBuf Q16((size_t)mp * l * 2, st), Qz16((size_t)np * l * 2, st);
    auto pass_a = [&](const void* Bop) {               // Y = A16 Bop
      Tm tm(c, "gemm_bf16_nn");
      LpGemmArgs g;
      g.mode = 0; g.M = mp; g.N = l; g.K = np; g.A = A16.p; g.lda = mp; g.a_trans = false;
      g.B = Bop; g.ldb = np; g.C = Y.as<float>(); g.ldc = mp; g.cstride = 0;
      HIPCK(gemm_lp(g, st));
    };
    auto pass_at = [&](const void* Bop) {              // Z = A16^T Bop
      const int64_t tiles = (np + 255) / 256;
      const int splits = choose_splits(tiles, mp, 4096);
      DBuf part((size_t)splits * np * l * 4, st);
      {
        Tm tm(c, "gemm_bf16_tn");
        LpGemmArgs g;
        g.mode = 0; g.M = np; g.N = l; g.K = mp; g.A = A16.p; g.lda = mp; g.a_trans = true;
        g.B = Bop; g.ldb = mp; g.C = part.as<float>(); g.ldc = np; g.cstride = np * l;
        g.splits = splits;
        HIPCK(gemm_lp(g, st));
      }
      HIPCK(reduce_partials(part.as<float>(), np, l, np, np * l, splits, Z.p, T::F32, np, 1.0, st));
    };
    if (fp8) {
      // FP8-e4m3 ladder: the first product Y = A Omega on an FP8 copy of A
      // (128 x 128 block scales, row-major: K-contiguous for A X) and an FP8
      // Omega (per-column 128-K block scales).  Its perturbation of A
      // (~2^-5 per entry) only changes the starting subspace, which the
      // later rounds contract like any random start; FP8 in a later round
      // would perturb every iterate by ~25 sigma_{k+1} for this spectrum.
      const int64_t m8 = (m + 127) / 128 * 128, n8 = (n + 127) / 128 * 128;
      DBuf Arm((size_t)m8 * n8, st), SA((size_t)(m8 / 128) * (n8 / 128), st),
          X8((size_t)n8 * l, st), SX((size_t)(n8 / 128) * l, st), P32((size_t)np * l * 4, st);
      {
        DBuf Acm((size_t)n8 * m8, st);
        fp8_operands(c, A, m, n, lda, Arm.as<uint8_t>(), Acm.as<uint8_t>(), SA.as<uint8_t>(),
                     nullptr, 0, 0, 0, nullptr, nullptr);
      }
      HIPCK(hipMemsetAsync(P32.p, 0, (size_t)np * l * 4, st));
      HIPCK(sketch_panel(0, s, n, l, P32.p, T::F32, np, 1.0, st));
      fp8_operands(c, nullptr, 0, 0, 0, nullptr, nullptr, nullptr, P32.as<float>(), n, l, np,
                   X8.as<uint8_t>(), SX.as<uint8_t>());
      Tm tm(c, "gemm_fp8_nn");
      Fp8GemmArgs g;
      g.M = mp; g.N = l; g.K = n8; g.A = Arm.as<uint8_t>(); g.lda = n8; g.SA = SA.as<uint8_t>();
      g.sa_r = n8 / 128; g.sa_k = 1; g.B = X8.as<uint8_t>(); g.ldb = n8; g.SB = SX.as<uint8_t>();
      g.sb_j = n8 / 128; g.C = Y.as<float>(); g.ldc = mp;
      HIPCK(gemm_fp8(g, st));
    } else {
      DBuf Om((size_t)np * l * 2, st);
      if (np != n) HIPCK(hipMemsetAsync(Om.p, 0, (size_t)np * l * 2, st));
      HIPCK(sketch_panel(0, s, n, l, Om.p, T::BF16, np, 1.0, st));
      pass_a(Om.p);
    }
    for (int it = 0; it < q; ++it) {
      if (it + kCleanRounds < q) {
        cholqr_pass(c, Y.p, T::F32, mp, l, mp, Q16.p, T::BF16, mp, nullptr, tau);
        pass_at(Q16.p);
        cholqr_pass(c, Z.p, T::F32, np, l, np, Qz16.p, T::BF16, np, nullptr, tau);
        pass_a(Qz16.p);
      } else {
        // The last rounds of the ladder: both products on the FP32 data,
        // rounded once to FP16 in all but the last round, split into hi and
        // lo in the last.
        const int mode = it + 1 < q ? 3 : 2;
        cholqr_pass(c, Y.p, T::F32, mp, l, mp, Q32, T::F32, mp, nullptr, tau);
        pass_at_split(Q32, mode);
        cholqr_pass(c, Z.p, T::F32, np, l, np, Qz32.p, T::F32, np, nullptr, tau);
        pass_a_split(Qz32.as<float>(), mode);
      }
    }
  }
  cholqr2(c, Y.p, T::F32, mp, l, mp, Q32, T::F32, mp, nullptr);
  // B^T = A^T Q: A rounded once to FP16, Q split into FP16 hi/lo (mode 4: two
  // products; Q's low part matters, A's rounding error lies in range(Q) of
  // the residual's complement and moves e by nothing measurable), FP32
  // accumulate, split-K partials summed in FP64.
  panel_amax(Q32, mp, l, mp);
  const int64_t tiles = (np + 255) / 256;
  const int splits = choose_splits(tiles, mp, 2048);
  DBuf part((size_t)splits * np * l * 4, st);
  {
    Tm tm(c, "gemm_bt");
    LpGemmArgs g;
    g.mode = 4; g.M = np; g.N = l; g.K = mp; g.A = Aop; g.lda = ldop; g.a_trans = true;
    g.B = Q32; g.ldb = mp; g.C = part.as<float>(); g.ldc = np; g.cstride = np * l;
    g.splits = splits; g.amax = amax.as<double>();
    g.bh = Bh.as<uint16_t>(); g.ldbh = mp;
    HIPCK(gemm_lp(g, st));
  }
  HIPCK(reduce_partials(part.as<float>(), n, l, np, np * l, splits, Bt, T::F64, n, 1.0, st));
}

void check_svd_params(int64_t m, int64_t n, int64_t k, double eps, int64_t s) {
  if (k == 0) fail(RANDNLA_INVALID_PARAMETERS, "Rank k must be positive, current input is 0");
  if (!(eps > 0.0)) {
    char b[128];
    snprintf(b, sizeof b, "Epsilon must be positive, current input is %g", eps);
    fail(RANDNLA_INVALID_PARAMETERS, b);
  }
  if (s == 0)
    fail(RANDNLA_INVALID_PARAMETERS, "Oversampling parameter s must be positive, current input is 0");
  if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "matrix dimensions must be positive");
  if (k < 0 || s < 0) fail(RANDNLA_NEGATIVE_DIMENSIONS, "k and s must be non-negative");
}

// ------------------------------------------------------------ LS helpers
// Preconditioned operator Ahat = A * P with P upper triangular (n x n) or a
// dense n x k matrix (LSRN / saddle).  Ahat v = A (P v); Ahat^T u = P^T (A^T u).
struct PrecOp {
  randnla_ctx* c;
  const double* A;
  int64_t m, n, lda;
  const double* P;   // n x k
  int64_t k;
  bool tri;          // P upper triangular (k == n)
  double* tn;        // n workspace
  double* work;      // gemv_t workspace (64 * n)
  void apply(const double* v, double* out) {   // out (m) = A P v
    Tm tm(c, "gemv");
    if (P) {
      if (tri) HIPCK(trmv_upper(P, n, n, false, v, tn, c->stream));
      else HIPCK(gemv_n(P, n, k, n, v, tn, 1.0, 0.0, c->stream));
      HIPCK(gemv_n(A, m, n, lda, tn, out, 1.0, 0.0, c->stream));
    } else {
      HIPCK(gemv_n(A, m, n, lda, v, out, 1.0, 0.0, c->stream));
    }
  }
  void apply_t(const double* u, double* out) {   // out (k) = P^T A^T u
    Tm tm(c, "gemv");
    if (P) {
      HIPCK(gemv_t(A, m, n, lda, u, tn, 1.0, 0.0, work, c->stream));
      if (tri) HIPCK(trmv_upper(P, n, n, true, tn, out, c->stream));
      else HIPCK(gemv_t(P, n, k, n, tn, out, 1.0, 0.0, work + 64 * n, c->stream));
    } else {
      HIPCK(gemv_t(A, m, n, lda, u, out, 1.0, 0.0, work, c->stream));
    }
  }
  int64_t cols() const { return P ? k : n; }
  // The fused LSQR pass works on A itself: w = P v before it (n), and
  // out = P^T s after it (k).
  const double* pre(const double* v) {
    if (!P) return v;
    if (tri) HIPCK(trmv_upper(P, n, n, false, v, tn, c->stream));
    else HIPCK(gemv_n(P, n, k, n, v, tn, 1.0, 0.0, c->stream));
    return tn;
  }
  void post_t(const double* s, double* out) {
    if (!P) {
      HIPCK(hipMemcpyAsync(out, s, n * 8, hipMemcpyDeviceToDevice, c->stream));
    } else if (tri) {
      HIPCK(trmv_upper(P, n, n, true, s, out, c->stream));
    } else {
      HIPCK(gemv_t(P, n, k, n, s, out, 1.0, 0.0, work + 64 * n, c->stream));
    }
  }
};

struct SolveOut {
  int istop = 0, itn = 0;
  double r1 = 0, r2 = 0, an = 0, ac = 0, ar = 0, xn = 0;
};

// CGLS on the operator (src/cg.rs:18-61).  x (k) holds x0 on entry.
SolveOut run_cgls(randnla_ctx* c, PrecOp& op, const double* b, double tol, int64_t iters,
                  double* x) {
  hipStream_t st = c->stream;
  const int64_t m = op.m, k = op.cols();
  DBuf r(m * 8, st), ap(m * 8, st), s(k * 8, st), p(k * 8, st), part(2048 * 8, st);
  DBuf Sd(sizeof(CglsState), st);
  // r = b - Ahat x ; s = Ahat^T r ; p = s ; norm_s = s . s
  op.apply(x, ap.as<double>());
  HIPCK(hipMemcpyAsync(r.p, b, m * 8, hipMemcpyDeviceToDevice, st));
  HIPCK(axpby(m, -1.0, ap.as<double>(), 1.0, r.as<double>(), st));
  op.apply_t(r.as<double>(), s.as<double>());
  HIPCK(hipMemcpyAsync(p.p, s.p, k * 8, hipMemcpyDeviceToDevice, st));
  DBuf ns(8, st);
  HIPCK(sqnorm(s.as<double>(), k, ns.as<double>(), part.as<double>(), st));
  CglsState h{};
  double hn = 0;
  d2h(st, &hn, ns.p, 8);
  sync(c);
  h.norm_s = hn;
  h.tol = tol;
  h.max_iters = (int)iters;
  h.done = iters <= 0;
  HIPCK(hipMemcpyAsync(Sd.p, &h, sizeof h, hipMemcpyHostToDevice, st));
  const int poll = 4;
  for (int64_t i = 0; i < iters; ++i) {
    op.apply(p.as<double>(), ap.as<double>());
    HIPCK(cgls_alpha(Sd.as<CglsState>(), ap.as<double>(), m, part.as<double>(), st));
    HIPCK(cgls_xr(Sd.as<CglsState>(), k, m, x, p.as<double>(), r.as<double>(), ap.as<double>(), st));
    op.apply_t(r.as<double>(), s.as<double>());
    HIPCK(cgls_beta_p(Sd.as<CglsState>(), k, p.as<double>(), s.as<double>(), part.as<double>(), st));
    if ((i + 1) % poll == 0 || i + 1 == iters) {
      d2h(st, &h, Sd.p, sizeof h);
      sync(c);
      if (h.done) break;
    }
  }
  d2h(st, &h, Sd.p, sizeof h);
  sync(c);
  SolveOut o;
  o.istop = h.converged ? 1 : 7;
  o.itn = h.iters;
  o.ar = h.last_norm;
  return o;
}

// LSQR on the operator (src/solvers.rs:115-278).  x (k) holds x0 if has_x0.
SolveOut run_lsqr(randnla_ctx* c, PrecOp& op, const double* b, double damp, double atol,
                  double btol, double conlim, int64_t iter_lim, bool calc_var, bool has_x0,
                  double* x, double* var, double* arnorms_host) {
  hipStream_t st = c->stream;
  const int64_t m = op.m, n = op.cols();
  if (iter_lim <= 0) iter_lim = 2 * n;
  DBuf u(m * 8, st), t(m * 8, st), v(n * 8, st), w(n * 8, st), tn(n * 8, st),
      part(4096 * 8, st), arn((iter_lim + 1) * 8, st), sc(16, st);
  if (!has_x0) HIPCK(hipMemsetAsync(x, 0, n * 8, st));
  if (calc_var) HIPCK(hipMemsetAsync(var, 0, n * 8, st));
  HIPCK(hipMemcpyAsync(u.p, b, m * 8, hipMemcpyDeviceToDevice, st));
  HIPCK(sqnorm(u.as<double>(), m, sc.as<double>(), part.as<double>(), st));   // ||b||^2
  if (has_x0) {
    op.apply(x, t.as<double>());
    HIPCK(axpby(m, -1.0, t.as<double>(), 1.0, u.as<double>(), st));
  }
  HIPCK(sqnorm(u.as<double>(), m, sc.as<double>() + 1, part.as<double>(), st));
  double hs[2];
  d2h(st, hs, sc.p, 16);
  sync(c);
  const double bnorm = sqrt(hs[0]);
  const double beta = sqrt(hs[1]);
  HIPCK(axpby(m, 0.0, u.as<double>(), beta > 0 ? 1.0 / beta : 0.0, u.as<double>(), st));
  op.apply_t(u.as<double>(), v.as<double>());
  HIPCK(sqnorm(v.as<double>(), n, sc.as<double>(), part.as<double>(), st));
  double a2 = 0;
  d2h(st, &a2, sc.p, 8);
  sync(c);
  const double alfa = sqrt(a2);
  HIPCK(axpby(n, 0.0, v.as<double>(), alfa > 0 ? 1.0 / alfa : 0.0, v.as<double>(), st));
  HIPCK(hipMemcpyAsync(w.p, v.p, n * 8, hipMemcpyDeviceToDevice, st));
  SolveOut o;
  const double arnorm0 = alfa * beta;
  if (arnorm0 == 0.0) {
    o.istop = 0; o.itn = 0; o.r1 = beta; o.r2 = beta;
    if (arnorms_host) arnorms_host[0] = 0.0;
    return o;
  }
  LsqrState h{};
  h.alfa = alfa; h.beta = beta; h.rhobar = alfa; h.phibar = beta; h.rnorm = beta;
  h.r1norm = beta; h.r2norm = beta; h.damp = damp; h.dampsq = damp * damp;
  h.cs2 = -1.0; h.sn2 = 0.0; h.arnorm = arnorm0; h.bnorm = bnorm;
  h.ctol = conlim > 0 ? 1.0 / conlim : 0.0; h.atol = atol; h.btol = btol;
  h.iter_lim = (int)iter_lim;
  h.umul = 1.0;
  DBuf Sd(sizeof(LsqrState), st);
  HIPCK(hipMemcpyAsync(Sd.p, &h, sizeof h, hipMemcpyHostToDevice, st));
  HIPCK(hipMemcpyAsync(arn.p, &arnorm0, 8, hipMemcpyHostToDevice, st));
  LsqrState* S = Sd.as<LsqrState>();
  const int poll = 8;
  // One pass over A per iteration when A's width allows (n <= 512): the
  // u-update, beta and A^T u from a single read of A.
  const bool fused = op.n <= 512;
  DBuf spart, atu;
  if (fused) {
    spart = DBuf((size_t)lsqr_fused_groups(m) * op.n * 8, st);
    atu = DBuf((size_t)op.n * 8, st);
  }
  for (int64_t it = 0; it < iter_lim; ++it) {
    if (fused) {
      Tm tm(c, "lsqr_pass");
      const double* wv = op.pre(v.as<double>());
      HIPCK(lsqr_fused_update(S, op.A, m, op.n, op.lda, wv, u.as<double>(), part.as<double>(),
                              spart.as<double>(), atu.as<double>(), st));
      op.post_t(atu.as<double>(), tn.as<double>());
      HIPCK(lsqr_vphase(S, n, tn.as<double>(), v.as<double>(), w.as<double>(), x, var, calc_var,
                        arn.as<double>(), (int)iter_lim, st));
    } else {
      op.apply(v.as<double>(), t.as<double>());
      HIPCK(lsqr_u_update(S, m, t.as<double>(), u.as<double>(), part.as<double>(), st));
      op.apply_t(u.as<double>(), tn.as<double>());
      HIPCK(lsqr_v_update(S, n, tn.as<double>(), v.as<double>(), w.as<double>(),
                          part.as<double>(), arn.as<double>(), st));
      HIPCK(lsqr_xw_update(S, n, x, w.as<double>(), v.as<double>(), var, calc_var,
                           arn.as<double>(), (int)iter_lim, st));
    }
    if ((it + 1) % poll == 0 || it + 1 == iter_lim) {
      d2h(st, &h, S, sizeof h);
      sync(c);
      if (h.istop) break;
    }
  }
  d2h(st, &h, S, sizeof h);
  sync(c);
  if (arnorms_host && h.itn > 0)
    HIPCK(hipMemcpy(arnorms_host, arn.p, (size_t)h.itn * 8, hipMemcpyDeviceToHost));
  o.istop = h.istop; o.itn = h.itn; o.r1 = h.r1norm; o.r2 = h.r2norm; o.an = h.anorm;
  o.ac = h.acond; o.ar = h.arnorm; o.xn = h.xnorm;
  return o;
}

void fill_info(randnla_solve_info* info, const SolveOut& o) {
  if (!info) return;
  info->istop = o.istop; info->itn = o.itn; info->r1norm = o.r1; info->r2norm = o.r2;
  info->anorm = o.an; info->acond = o.ac; info->arnorm = o.ar; info->xnorm = o.xn;
}

void check_ls(int64_t m, int64_t n, double eps, int64_t l, double sf) {
  char buf[200];
  if (m < n) {
    snprintf(buf, sizeof buf, "Need more columns than rows, found %lld rows and %lld columns",
             (long long)m, (long long)n);
    fail(RANDNLA_NOT_OVERDETERMINED, buf);
  }
  if (sf < 1.0) {
    snprintf(buf, sizeof buf, "Sampling factor must be greater than 1, current input is %g", sf);
    fail(RANDNLA_INVALID_PARAMETERS, buf);
  }
  if (!(eps > 0.0)) {
    snprintf(buf, sizeof buf, "Epsilon must be positive, current input is %g", eps);
    fail(RANDNLA_INVALID_PARAMETERS, buf);
  }
  if (l <= 0) fail(RANDNLA_INVALID_PARAMETERS, "Number of iterations must be positive, current input is 0");
  if (n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "matrix dimensions must be positive");
}

// out (d x cols) = S * X for the LS sketch selected in opts (X FP64, m rows).
void apply_sketch(randnla_ctx* c, const randnla_opts& o, const Stream& s, int64_t d,
                  const double* X, int64_t m, int64_t cols, int64_t ldx, double* out,
                  int64_t ldo) {
  Tm tm(c, "sketch_apply");
  if (o.sketch == RANDNLA_SPARSE_SIGN) {
    const int64_t need = sparse_sign_work(d, m, cols);
    DBuf w((size_t)need * 8, c->stream);
    HIPCK(sparse_sign_apply(s, d, o.zeta, X, T::F64, m, cols, ldx, out, ldo, w.as<double>(), need,
                            c->stream));
  } else if (cols > 64 && (double)d * (double)m * 8.0 <= 8e9) {
    // Wide operand: the generated-operand GEMM would make S once per 64-column
    // tile; S is made once into memory (the same entries) and multiplied
    DBuf S((size_t)d * m * 8, c->stream);
    HIPCK(sketch_panel(o.sketch, s, d, m, S.p, T::F64, d, 1.0, c->stream));
    gemm64(c, d, cols, m, S.p, T::F64, d, false, X, T::F64, ldx, out, T::F64, ldo, 1.0,
           "sketch_gemm");
  } else {
    const int64_t need = dense_sketch_work(0, m, cols, d);
    DBuf w((size_t)std::max<int64_t>(need, 1) * 8, c->stream);
    HIPCK(dense_sketch_apply(0, o.sketch, s, X, T::F64, m, cols, ldx, d, out, ldo, w.as<double>(),
                             need, c->stream));
  }
}

// [S A | S b] into out (d x (n + 1), leading dimension ldo).  A dense sketch
// is generated inside the GEMM, so A and b are first copied side by side and
// sketched in one product (S is made once, not once more for b); a
// sparse-sign sketch is cheap to regenerate and takes two calls.
void apply_sketch_ab(randnla_ctx* c, const randnla_opts& o, const Stream& s, int64_t d,
                     const double* A, int64_t m, int64_t n, int64_t lda, const double* b,
                     double* out, int64_t ldo) {
  if (o.sketch == RANDNLA_SPARSE_SIGN) {
    apply_sketch(c, o, s, d, A, m, n, lda, out, ldo);
    apply_sketch(c, o, s, d, b, m, 1, m, out + n * ldo, ldo);
    return;
  }
  hipStream_t st = c->stream;
  DBuf Ab((size_t)m * (n + 1) * 8, st);
  HIPCK(hipMemcpy2DAsync(Ab.p, m * 8, A, lda * 8, m * 8, n, hipMemcpyDeviceToDevice, st));
  HIPCK(hipMemcpyAsync(Ab.as<double>() + n * m, b, m * 8, hipMemcpyDeviceToDevice, st));
  apply_sketch(c, o, s, d, Ab.as<double>(), m, n + 1, m, out, ldo);
}

// Householder QR (unpivoted unless pivot) of a device d x n matrix in place;
// returns the reflector panel (tau layout of qrcp_householder).  Unpivoted,
// `extra` columns stored after the n are multiplied by Q^T as well.
DBuf householder(randnla_ctx* c, double* M, int64_t d, int64_t n, bool pivot, int64_t* perm,
                 int64_t extra = 0) {
  Tm tm(c, "householder");
  const int64_t kmax = std::min(d, n);
  DBuf tau((size_t)(n + d * kmax + d) * 8, c->stream);
  DBuf pbuf;
  if (!perm) {
    pbuf = DBuf((size_t)n * 8, c->stream);
    perm = pbuf.as<int64_t>();
  }
  HIPCK(qrcp_householder(M, d, n, d, tau.as<double>(), perm, pivot, c->stream, extra));
  return tau;
}

// Small SVD of a device d x n matrix (n <= d): sigma (n), V (n x n) and
// U (d x n).  CholQR2 gives W = Qw Rw; one-sided Jacobi on Rw^T gives
// Rw^T Vj = Uj S, so W = (Qw Vj) S Uj^T: U = Qw Vj, V = Uj.  Jacobi on Rw^T
// takes about half the sweeps of Jacobi on Rw.  A zero sigma leaves a
// placeholder unit column in V; V is then re-orthonormalised (CholQR with
// unit-vector completion), so V stays orthogonal for every caller.
void small_svd(randnla_ctx* c, const double* Wm, int64_t d, int64_t n, double* U, double* sig,
               double* V) {
  hipStream_t st = c->stream;
  DBuf Qw((size_t)d * n * 8, st), Rw((size_t)n * n * 8, st), Wt((size_t)n * n * 8, st);
  cholqr2(c, Wm, T::F64, d, n, d, Qw.p, T::F64, d, Rw.as<double>());
  HIPCK(transpose_f64(Rw.as<double>(), n, n, n, Wt.as<double>(), n, st));
  DBuf Vr((size_t)n * n * 8, st), sv((size_t)n * 8, st), cnt(8, st);
  HIPCK(set_identity(Vr.as<double>(), n, n, n, st));
  int sweeps = 0;
  {
    Tm tm(c, "jacobi");
    HIPCK(jacobi_svd(Wt.as<double>(), n, n, Vr.as<double>(), n, sv.as<double>(), cnt.as<int>(), 30,
                     jacobi_tol(n), st, &sweeps));
  }
  c->diag[0] = sweeps;
  DBuf Vj((size_t)n * n * 8, st);
  HIPCK(svd_finalize(Wt.as<double>(), n, n, Vr.as<double>(), n, sv.as<double>(), V, n,
                     Vj.as<double>(), n, sig, st));
  double smin = 0.0;
  d2h(st, &smin, sig + (n - 1), 8);
  if (smin == 0.0) {
    DBuf V2((size_t)n * n * 8, st);
    cholqr_pass(c, V, T::F64, n, n, n, V2.p, T::F64, n, nullptr, default_tau(n));
    HIPCK(hipMemcpyAsync(V, V2.p, (size_t)n * n * 8, hipMemcpyDeviceToDevice, st));
  }
  if (U) gemm64(c, d, n, n, Qw.p, T::F64, d, false, Vj.p, T::F64, n, U, T::F64, d, 1.0, "svd_u");
}

// tsog1's products and stabilisations (src/lora_helpers.rs:58-105, intended
// method: the loop starts from the generated test matrix; re-orthonormalisation
// after every passes_per_stab products).  S (n x k, ld n) into Sn.
void tsog1_core(randnla_ctx* c, const void* A, int64_t m, int64_t n, int64_t lda, int64_t k,
                int num_passes, int passes_per_stab, const Stream& s, double* Sn) {
  hipStream_t st = c->stream;
  DBuf Sm((size_t)m * k * 8, st), tmp((size_t)std::max(m, n) * k * 8, st);
  int done = 0;
  const double tau = default_tau(k);
  auto stab = [&](double* X, int64_t rows) {
    if (done % passes_per_stab == 0 && k <= rows) {
      cholqr_pass(c, X, T::F64, rows, k, rows, tmp.p, T::F64, rows, nullptr, tau);
      HIPCK(hipMemcpyAsync(X, tmp.p, (size_t)rows * k * 8, hipMemcpyDeviceToDevice, st));
    }
  };
  if (num_passes % 2 == 0) {
    HIPCK(sketch_panel(0, s, n, k, Sn, T::F64, n, 1.0, st));
  } else {
    HIPCK(sketch_panel(0, s, m, k, Sm.p, T::F64, m, 1.0, st));
    gemm64(c, n, k, m, A, T::F64, lda, true, Sm.p, T::F64, m, Sn, T::F64, n);
    done += 1;
    stab(Sn, n);
  }
  int diff = num_passes - done;
  while (diff >= 2) {
    gemm64(c, m, k, n, A, T::F64, lda, false, Sn, T::F64, n, Sm.p, T::F64, m);
    done += 1;
    stab(Sm.as<double>(), m);
    gemm64(c, n, k, m, A, T::F64, lda, true, Sm.p, T::F64, m, Sn, T::F64, n);
    done += 1;
    stab(Sn, n);
    diff -= 2;
  }
}

// Orth as the reference computes it, nalgebra's thin QR
// (src/lora_helpers.rs:131-133): Householder QR of X (m x n), Q (m x k,
// k = min(m, n)) with column i scaled by sign(R_ii) (+1 for a zero R_ii), so
// that R has a non-negative diagonal; R (k x n, optional) the matching upper
// factor.  A wide X factors its leading m x m block and takes the remaining
// columns as right-hand sides (R = [R_11 | Q^T X_2]).
void orth_householder(randnla_ctx* c, const double* X, int64_t m, int64_t n, int64_t ldx,
                      double* Q, int64_t ldq, double* R, int64_t ldr) {
  hipStream_t st = c->stream;
  const int64_t kq = std::min(m, n);
  DBuf W((size_t)m * n * 8, st);
  HIPCK(hipMemcpy2DAsync(W.p, m * 8, X, ldx * 8, m * 8, n, hipMemcpyDeviceToDevice, st));
  DBuf tau = householder(c, W.as<double>(), m, kq, false, nullptr, n - kq);
  HIPCK(householder_form_q(nullptr, m, kq, m, tau.as<double>(), Q, ldq, kq, st));
  DBuf sg((size_t)kq * 8, st);
  HIPCK(diag_signs(W.as<double>(), kq, m, sg.as<double>(), st));
  HIPCK(scale_columns(Q, m, kq, ldq, sg.as<double>(), 1, st));
  if (R) HIPCK(upper_scaled(W.as<double>(), kq, n, m, sg.as<double>(), R, ldr, st));
}

// Stabilizer as the reference computes it (src/lora_helpers.rs:144-146):
// L (rows x min(rows, cols)) of the full-pivot LU of X (rows x cols, ld ldx)
// into L (ld ldl); returns min(rows, cols).  X is not modified.
int64_t stabilizer_lu(randnla_ctx* c, const double* X, int64_t rows, int64_t cols, int64_t ldx,
                      double* L, int64_t ldl) {
  hipStream_t st = c->stream;
  DBuf W((size_t)rows * cols * 8, st), work((size_t)full_piv_lu_work(rows, cols) * 8, st);
  HIPCK(hipMemcpy2DAsync(W.p, rows * 8, X, ldx * 8, rows * 8, cols, hipMemcpyDeviceToDevice, st));
  HIPCK(full_piv_lu_l(W.as<double>(), rows, cols, rows, L, ldl, work.as<double>(), st));
  return std::min(rows, cols);
}

// Number of columns of the reference's tsog1 result (src/lora_helpers.rs:58-105):
// the loop body is the same in every iteration (S1 does not change), so only
// the last iteration's stabiliser pattern matters.
int64_t tsog1_faithful_cols(int64_t m, int64_t n, int64_t k, int num_passes, int pps) {
  const int done0 = num_passes % 2;
  const int J = (num_passes - done0) / 2;
  if (J <= 0) return k;
  const int64_t k1 = ((done0 + 2 * J - 1) % pps == 0) ? std::min(m, k) : k;
  return ((done0 + 2 * J) % pps == 0) ? std::min(n, k1) : k1;
}

// tsog1 as the reference computes it (src/lora_helpers.rs:58-105).  Even
// num_passes: the Gaussian goes to S and the loop multiplies A by S1, which is
// still zero (:70-72 vs :89).  Odd: S1 = A^T G with G an m x k Gaussian, and
// the stabilised copy of S1 is dropped (:79-82).  Each loop iteration computes
// S = A S1, Stabilizer, S = A^T S, Stabilizer from the same S1, so the result
// is the last iteration's; it is computed once.  Stabilizer is the full-pivot
// LU factor L.  S (n x kc, kc = tsog1_faithful_cols) goes to Sn (ld n, room
// for n x k, columns kc..k-1 zero).  Returns kc.
int64_t tsog1_faithful(randnla_ctx* c, const void* A, T ta, int64_t m, int64_t n, int64_t lda,
                       int64_t k, int num_passes, int pps, const Stream& s, double* Sn) {
  hipStream_t st = c->stream;
  const int done0 = num_passes % 2;
  const int J = (num_passes - done0) / 2;
  HIPCK(fill_f64(Sn, n, k, n, 0.0, st));
  DBuf S1((size_t)n * k * 8, st);
  if (done0 == 0) {
    HIPCK(fill_f64(S1.as<double>(), n, k, n, 0.0, st));
    if (J == 0) HIPCK(sketch_panel(0, s, n, k, Sn, T::F64, n, 1.0, st));   // S is returned as is
  } else {
    DBuf G((size_t)m * k * 8, st);
    HIPCK(sketch_panel(0, s, m, k, G.p, T::F64, m, 1.0, st));
    gemm64(c, n, k, m, A, ta, lda, true, G.p, T::F64, m, S1.p, T::F64, n, 1.0, "pass_at");
  }
  if (J <= 0) return k;
  DBuf Sm((size_t)m * k * 8, st), Lm((size_t)m * k * 8, st);
  gemm64(c, m, k, n, A, ta, lda, false, S1.p, T::F64, n, Sm.p, T::F64, m, 1.0, "pass_a");
  int64_t k1 = k;
  const double* Sa = Sm.as<double>();
  if ((done0 + 2 * J - 1) % pps == 0) {
    k1 = stabilizer_lu(c, Sm.as<double>(), m, k, m, Lm.as<double>(), m);
    Sa = Lm.as<double>();
  }
  gemm64(c, n, k1, m, A, ta, lda, true, Sa, T::F64, m, Sn, T::F64, n, 1.0, "pass_at");
  int64_t kc = k1;
  if ((done0 + 2 * J) % pps == 0) {
    DBuf Ln((size_t)n * k1 * 8, st);
    kc = stabilizer_lu(c, Sn, n, k1, n, Ln.as<double>(), n);
    HIPCK(hipMemcpyAsync(Sn, Ln.p, (size_t)n * kc * 8, hipMemcpyDeviceToDevice, st));
  }
  if (kc < k) HIPCK(fill_f64(Sn + n * kc, n, k - kc, n, 0.0, st));
  return kc;
}

// The range finder behind RF1 / QB1 / rand_svd's FP64 path: Q (m x kc, FP64,
// ld m, room for m x l), kc returned.  Intended method (Gaussian test matrix
// from o.rng, q rounds, CholeskyQR2), kc = min(l, m, n); or, with
// o.faithful, the reference's RF1 (src/lora_helpers.rs:37-44):
// Y = A tsog1(A, l, 2, 1), Q = Orth(Y) by Householder QR, the discarded
// Gaussian drawn from the reference's stream (ThreeFry2x64 seed_from_u64(0)).
int64_t range_finder(randnla_ctx* c, const randnla_opts& o, const void* A, T ta, int64_t m,
                     int64_t n, int64_t lda, int64_t l, double* Q) {
  hipStream_t st = c->stream;
  if (o.faithful) {
    DBuf S((size_t)n * l * 8, st);
    const Stream s = stream_from_seed(GEN_THREEFRY2X64, 0);
    const int64_t kc = tsog1_faithful(c, A, ta, m, n, lda, l, 2, 1, s, S.as<double>());
    DBuf Y((size_t)m * kc * 8, st);
    gemm64(c, m, kc, n, A, ta, lda, false, S.p, T::F64, n, Y.p, T::F64, m, 1.0, "pass_a");
    orth_householder(c, Y.as<double>(), m, kc, m, Q, m, nullptr, 0);
    return std::min(m, kc);
  }
  const int64_t le = std::min(l, std::min(m, n));
  range_finder_f64(c, A, ta, m, n, lda, le, o.power_iters, to_stream(&o.rng, GEN_PHILOX4X32, 42),
                   Q);
  return le;
}

// Indices of lam by |lam| descending, ties kept in index order (the stable
// sort of src/lora_drivers.rs:132-138).
std::vector<int64_t> abs_desc_order(const std::vector<double>& lam) {
  std::vector<int64_t> idx(lam.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int64_t)i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int64_t a, int64_t b) { return std::fabs(lam[a]) > std::fabs(lam[b]); });
  return idx;
}

// ------------------------------------------ interpolative decompositions
// Reference panics of src/id.rs (osid_qrcp :252-254, osid_randomised
// :212-214) and of economic_qrcp (src/pivot_decompositions.rs:198-199).
void check_id_k(int64_t k, int64_t l, int64_t w) {
  if (k <= 0) fail(RANDNLA_PANIC, "k must be positive)");
  if (k > std::min(l, w)) fail(RANDNLA_PANIC, "k must be <= min(l,w)");
}
void check_economic_k(int64_t k, int64_t m, int64_t n) {
  if (k > std::min(m, n)) fail(RANDNLA_PANIC, "k must be <= min(m,n)");
  if (k <= 0) fail(RANDNLA_PANIC, "k must be positive");
}

// The first `steps` steps of Householder QRCP on W (l x w, ld l, in place):
// pivot = first maximum of the recomputed trailing column norms
// (src/pivot_decompositions.rs:194-269).  perm (w, device); the reflector
// flags and panel are in the returned buffer.
DBuf qrcp_steps(randnla_ctx* c, double* W, int64_t l, int64_t w, int64_t steps, int64_t* perm) {
  Tm tm(c, "qrcp");
  const int64_t kmax = std::min(l, w);
  DBuf tau((size_t)(w + l * kmax + l) * 8, c->stream);
  HIPCK(qrcp_householder(W, l, w, l, tau.as<double>(), perm, true, c->stream, 0, steps));
  return tau;
}

// Column ID of Y (l x w, device) by k QRCP steps (src/id.rs:256-296):
// X (k x w, device ld ldx) has columns J[c] = e_c for c < k and
// R1^{-1} R2 after; J (k, device) the pivot columns.
void osid_col(randnla_ctx* c, const double* Y, int64_t l, int64_t w, int64_t ldy, int64_t k,
              double* X, int64_t ldx, int64_t* J) {
  hipStream_t st = c->stream;
  DBuf W((size_t)l * w * 8, st), P((size_t)w * 8, st);
  HIPCK(hipMemcpy2DAsync(W.p, l * 8, Y, ldy * 8, l * 8, w, hipMemcpyDeviceToDevice, st));
  DBuf tau = qrcp_steps(c, W.as<double>(), l, w, k, P.as<int64_t>());
  // solve_upper_triangular(...).unwrap() (src/id.rs:264) fails on a zero R_ii
  DBuf kc(8, st);
  HIPCK(count_diag_above(W.as<double>(), k, l, 0.0, kc.as<int>(), st));
  int nz = 0;
  d2h(st, &nz, kc.p, sizeof(int));
  if (nz < k) fail(RANDNLA_PANIC, "called `Option::unwrap()` on a `None` value");
  DBuf Tm_((size_t)k * std::max<int64_t>(w - k, 1) * 8, st);
  if (w > k) {
    DBuf R1((size_t)k * k * 8, st), Ri((size_t)k * k * 8, st);
    HIPCK(upper_copy(W.as<double>(), k, k, l, R1.as<double>(), k, st));
    HIPCK(trinv_upper(R1.as<double>(), k, k, Ri.as<double>(), k, st));
    gemm64(c, k, w - k, k, Ri.p, T::F64, k, false, W.as<double>() + k * l, T::F64, l, Tm_.p,
           T::F64, k, 1.0, "id_t");
  }
  HIPCK(id_scatter(P.as<int64_t>(), Tm_.as<double>(), k, w, X, ldx, st));
  HIPCK(hipMemcpyAsync(J, P.p, (size_t)k * 8, hipMemcpyDeviceToDevice, st));
}

// Row ID (src/id.rs:298-303): the column ID of Y^T, transposed.  X (l x k).
void osid_row(randnla_ctx* c, const double* Y, int64_t l, int64_t w, int64_t ldy, int64_t k,
              double* X, int64_t ldx, int64_t* I) {
  hipStream_t st = c->stream;
  DBuf Yt((size_t)w * l * 8, st), Xt((size_t)k * l * 8, st);
  HIPCK(transpose_f64(Y, l, w, ldy, Yt.as<double>(), w, st));
  osid_col(c, Yt.as<double>(), w, l, w, k, Xt.as<double>(), k, I);
  HIPCK(transpose_f64(Xt.as<double>(), k, l, k, X, ldx, st));
}

void osid_qrcp_dev(randnla_ctx* c, const double* Y, int64_t l, int64_t w, int64_t ldy, int64_t k,
                   int attr, double* X, int64_t ldx, int64_t* J) {
  check_id_k(k, l, w);
  if (attr == RANDNLA_ROW) osid_row(c, Y, l, w, ldy, k, X, ldx, J);
  else osid_col(c, Y, l, w, ldy, k, X, ldx, J);
}

// src/id.rs:205-234.  Column: Y = S A with S a k x m Gaussian (the dense
// sketch stream), column ID of Y.  Row: S = tsog1(A, k, 2, 1), Y = A S^T
// (defined only for n == k, as in the reference), row ID of Y.
void osid_randomised_dev(randnla_ctx* c, const randnla_opts* opts, const double* A, int64_t m,
                         int64_t n, int64_t lda, int64_t k, int attr, double* X, int64_t ldx,
                         int64_t* J) {
  check_id_k(k, m, n);
  const randnla_opts o = opts_or_default(opts);
  hipStream_t st = c->stream;
  if (attr == RANDNLA_ROW) {
    if (n != k)
      fail(RANDNLA_PANIC, "Matrix multiplication dimensions mismatch: A has n columns, S^T k rows");
    const Stream s = to_stream(&o.rng, GEN_PHILOX4X32, 42);
    DBuf S((size_t)n * k * 8, st), St((size_t)k * n * 8, st), Y((size_t)m * k * 8, st);
    tsog1_core(c, A, m, n, lda, k, 2, 1, s, S.as<double>());
    HIPCK(transpose_f64(S.as<double>(), n, k, n, St.as<double>(), k, st));
    gemm64(c, m, k, n, A, T::F64, lda, false, St.p, T::F64, k, Y.p, T::F64, m, 1.0, "id_y");
    osid_row(c, Y.as<double>(), m, k, m, k, X, ldx, J);
  } else {
    const Stream s = to_stream(opts ? &o.rng : nullptr, GEN_THREEFRY2X64, 0);
    DBuf Y((size_t)k * n * 8, st);
    apply_sketch(c, o, s, k, A, m, n, lda, Y.as<double>(), k);
    osid_col(c, Y.as<double>(), k, n, k, k, X, ldx, J);
  }
}

// Pivot order of the first k QRCP steps of Y (l x w) into P (w, device).
void qrcp_order(randnla_ctx* c, const double* Y, int64_t l, int64_t w, int64_t ldy, int64_t k,
                int64_t* P) {
  check_economic_k(k, l, w);
  DBuf W((size_t)l * w * 8, c->stream);
  HIPCK(hipMemcpy2DAsync(W.p, l * 8, Y, ldy * 8, l * 8, w, hipMemcpyDeviceToDevice, c->stream));
  DBuf tau = qrcp_steps(c, W.as<double>(), l, w, k, P);
}

// pinv(M) B for a tall M (d x k, k <= d, ld d, device) and B (d x k2, ld
// ldb, device) into out (k x k2, ld ldo): Householder QR of [M | B] (Q^T
// applied to B in the same factorisation), one-sided Jacobi SVD of the
// k x k R = Ur S Vr^T, then Vr S^+ Ur^T (Q^T B)[0:k], with S^+ inverting
// every sigma > 0 (nalgebra's pseudo_inverse(0.0)).  The factorisation is
// backward stable, so ill-conditioned M keep their small singular values.
void pinv_apply(randnla_ctx* c, const double* M, int64_t d, int64_t k, const double* B,
                int64_t ldb, int64_t k2, double* out, int64_t ldo) {
  hipStream_t st = c->stream;
  DBuf W((size_t)d * (k + k2) * 8, st);
  HIPCK(hipMemcpyAsync(W.p, M, (size_t)d * k * 8, hipMemcpyDeviceToDevice, st));
  HIPCK(hipMemcpy2DAsync(W.as<double>() + d * k, d * 8, B, ldb * 8, d * 8, k2,
                         hipMemcpyDeviceToDevice, st));
  DBuf tau = householder(c, W.as<double>(), d, k, false, nullptr, k2);
  DBuf R((size_t)k * k * 8, st), Vr((size_t)k * k * 8, st), sv((size_t)k * 8, st), cnt(8, st),
      Ur((size_t)k * k * 8, st), Vs((size_t)k * k * 8, st), ss((size_t)k * 8, st),
      inv((size_t)k * 8, st), T((size_t)k * k2 * 8, st);
  // One-sided Jacobi on R^T (lower triangular): R^T Vj = Uj S, so R = Vj S Uj^T
  // and pinv(M) B = Uj S^+ Vj^T (Q^T B)[0:k].  R's columns follow the pivot
  // order of the calling decomposition, so R is graded and R^T's columns are
  // nearly orthogonal: fewer sweeps than on R (7 against 19 in a scalar model
  // of a kappa = 3e8 case).
  HIPCK(upper_copy(W.as<double>(), k, k, d, R.as<double>(), k, st));
  HIPCK(transpose_f64(R.as<double>(), k, k, k, Ur.as<double>(), k, st));
  HIPCK(set_identity(Vr.as<double>(), k, k, k, st));
  int sweeps = 0;
  {
    Tm tm(c, "jacobi");
    HIPCK(jacobi_svd(Ur.as<double>(), k, k, Vr.as<double>(), k, sv.as<double>(), cnt.as<int>(), 30,
                     jacobi_tol(k), st, &sweeps));
  }
  c->diag[0] = sweeps;
  HIPCK(svd_finalize(Ur.as<double>(), k, k, Vr.as<double>(), k, sv.as<double>(), R.as<double>(), k,
                     Vs.as<double>(), k, ss.as<double>(), st));
  HIPCK(inv_sigma(ss.as<double>(), k, 0.0, 0.0, 2, inv.as<double>(), cnt.as<int>(), st));
  // R now holds Uj (normalised columns of R^T Vj), Vs holds Vj
  gemm64(c, k, k2, k, Vs.p, T::F64, k, true, W.as<double>() + d * k, T::F64, d, T.p, T::F64, k,
         1.0, "pinv");
  HIPCK(scale_rows(T.as<double>(), k, k2, k, inv.as<double>(), st));
  gemm64(c, k, k2, k, R.p, T::F64, k, false, T.p, T::F64, k, out, T::F64, ldo, 1.0, "pinv");
}

// CUR (src/id.rs:33-76 and :154-197): J (k), U (k x k, device ld ldu), I (k),
// all on the device.
void cur_dev(randnla_ctx* c, const randnla_opts* opts, bool randomised, const double* A,
             int64_t m, int64_t n, int64_t lda, int64_t k, int64_t* J, double* U, int64_t ldu,
             int64_t* I) {
  hipStream_t st = c->stream;
  if (m >= n) {
    DBuf X((size_t)k * n * 8, st);
    if (randomised) osid_randomised_dev(c, opts, A, m, n, lda, k, RANDNLA_COLUMN, X.as<double>(), k, J);
    else osid_qrcp_dev(c, A, m, n, lda, k, RANDNLA_COLUMN, X.as<double>(), k, J);
    DBuf AC((size_t)m * k * 8, st), ACt((size_t)k * m * 8, st), P((size_t)m * 8, st);
    HIPCK(gather_cols(A, m, lda, J, k, AC.as<double>(), m, st));
    HIPCK(transpose_f64(AC.as<double>(), m, k, m, ACt.as<double>(), k, st));
    qrcp_order(c, ACt.as<double>(), k, m, k, k, P.as<int64_t>());
    HIPCK(hipMemcpyAsync(I, P.p, (size_t)k * 8, hipMemcpyDeviceToDevice, st));
    // U = X pinv(A[I, :]) = (pinv(A[I, :]^T) X^T)^T
    DBuf AR((size_t)k * n * 8, st), ARt((size_t)n * k * 8, st), Xt((size_t)n * k * 8, st),
        Ut((size_t)k * k * 8, st);
    HIPCK(gather_rows(A, n, lda, I, k, AR.as<double>(), k, st));
    HIPCK(transpose_f64(AR.as<double>(), k, n, k, ARt.as<double>(), n, st));
    HIPCK(transpose_f64(X.as<double>(), k, n, k, Xt.as<double>(), n, st));
    pinv_apply(c, ARt.as<double>(), n, k, Xt.as<double>(), n, k, Ut.as<double>(), k);
    HIPCK(transpose_f64(Ut.as<double>(), k, k, k, U, ldu, st));
  } else {
    DBuf At((size_t)n * m * 8, st), Z((size_t)k * m * 8, st);
    HIPCK(transpose_f64(A, m, n, lda, At.as<double>(), n, st));
    if (randomised)
      osid_randomised_dev(c, opts, At.as<double>(), n, m, n, k, RANDNLA_COLUMN, Z.as<double>(), k, I);
    else osid_qrcp_dev(c, At.as<double>(), n, m, n, k, RANDNLA_COLUMN, Z.as<double>(), k, I);
    DBuf AR((size_t)k * n * 8, st), P((size_t)n * 8, st);
    HIPCK(gather_rows(A, n, lda, I, k, AR.as<double>(), k, st));
    qrcp_order(c, AR.as<double>(), k, n, k, k, P.as<int64_t>());
    HIPCK(hipMemcpyAsync(J, P.p, (size_t)k * 8, hipMemcpyDeviceToDevice, st));
    // U = pinv(A[:, J]) Z^T
    DBuf AC((size_t)m * k * 8, st), Zt((size_t)m * k * 8, st);
    HIPCK(gather_cols(A, m, lda, J, k, AC.as<double>(), m, st));
    HIPCK(transpose_f64(Z.as<double>(), k, m, k, Zt.as<double>(), m, st));
    pinv_apply(c, AC.as<double>(), m, k, Zt.as<double>(), m, k, U, ldu);
  }
}

// k indices from the device to a device or host pointer.
void put_indices(randnla_ctx* c, int64_t* dst, const int64_t* src, int64_t k) {
  if (!dst || k <= 0) return;
  if (is_device_ptr(dst))
    HIPCK(hipMemcpyAsync(dst, src, (size_t)k * 8, hipMemcpyDeviceToDevice, c->stream));
  else
    d2h(c->stream, dst, src, (size_t)k * 8);
}

// r eigenvalues to a device or host pointer.
void put_values(randnla_ctx* c, double* dst, const std::vector<double>& v, int64_t r) {
  if (!dst || r <= 0) return;
  if (is_device_ptr(dst)) {
    HIPCK(hipMemcpyAsync(dst, v.data(), (size_t)r * 8, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
  } else {
    memcpy(dst, v.data(), (size_t)r * 8);
  }
}

}  // namespace

// ===================================================================== ABI
extern "C" {

void randnla_opts_default(randnla_opts* o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->precision = RANDNLA_FP64;
  o->power_iters = 1;             // RF1 = tsog1(A, k, 2, 1): one (A, A^T) round
  o->sketch = RANDNLA_GAUSSIAN;
  o->zeta = 8;
  o->solver = RANDNLA_CGLS;
  o->faithful = 0;
  const Stream s = stream_from_seed(GEN_PHILOX4X32, 42);
  o->rng.gen = s.gen;
  o->rng.key[0] = s.key[0];
  o->rng.key[1] = s.key[1];
  o->rng.ctr[0] = 0;
  o->rng.ctr[1] = 0;
}

int randnla_create(randnla_ctx** out, int device) {
  if (!out) return RANDNLA_INVALID_PARAMETERS;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return RANDNLA_NO_DEVICE;
  }
  if (device < 0 || device >= n) return RANDNLA_NO_DEVICE;
  randnla_ctx* c = new (std::nothrow) randnla_ctx();
  if (!c) return RANDNLA_COMPUTATION_ERROR;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return RANDNLA_HIP_ERROR;
  }
  c->own_stream = true;
  double* scr = nullptr;
  if (scratch_lookup(device, &scr) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return RANDNLA_HIP_ERROR;
  }
  {
    std::lock_guard<std::mutex> g(stage_mu);
    ++live_contexts;
  }
  // Keep freed workspace in the stream-ordered pool: the rand_svd workspaces
  // are GB-sized and would otherwise be unmapped and remapped on every call.
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
    uint64_t thr = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
  }
  *out = c;
  return RANDNLA_OK;
}

// The context's stream is drained before its cached workspace is returned,
// and the pinned staging buffer goes with the last context, so that nothing
// of the library is left for the runtime's own teardown at process exit.
void randnla_destroy(randnla_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  resolve_timers(c);
  ws_flush(c->stream);
  (void)hipStreamSynchronize(c->stream);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
  std::lock_guard<std::mutex> g(stage_mu);
  if (--live_contexts == 0 && stage_buf) {
    (void)hipHostFree(stage_buf);
    stage_buf = nullptr;
  }
}

const char* randnla_last_error(const randnla_ctx* c) { return c ? c->err.c_str() : "no context"; }

// The old stream is drained and its cached workspace returned before it is
// replaced (and destroyed when the context owns it).
int randnla_set_stream(randnla_ctx* c, void* stream) {
  if (!c) return RANDNLA_INVALID_PARAMETERS;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  resolve_timers(c);
  ws_flush(c->stream);
  (void)hipStreamSynchronize(c->stream);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  c->stream = (hipStream_t)stream;
  c->own_stream = false;
  return RANDNLA_OK;
}

int randnla_get_stream(const randnla_ctx* c, void** stream) {
  if (!c || !stream) return RANDNLA_INVALID_PARAMETERS;
  *stream = (void*)c->stream;
  return RANDNLA_OK;
}

int randnla_timing_enable(randnla_ctx* c, int on) {
  if (!c) return RANDNLA_INVALID_PARAMETERS;
  c->timing = on != 0;
  return RANDNLA_OK;
}

int randnla_timing_query(randnla_ctx* c, const char* name, double* total_ms, int64_t* launches) {
  if (!c || !name) return RANDNLA_INVALID_PARAMETERS;
  resolve_timers(c);
  auto it = c->timers.find(name);
  if (total_ms) *total_ms = it == c->timers.end() ? 0.0 : it->second.total_ms;
  if (launches) *launches = it == c->timers.end() ? 0 : it->second.launches;
  return RANDNLA_OK;
}

int randnla_timing_reset(randnla_ctx* c) {
  if (!c) return RANDNLA_INVALID_PARAMETERS;
  resolve_timers(c);
  c->timers.clear();
  return RANDNLA_OK;
}

int randnla_diag(const randnla_ctx* c, int i, double* out) {
  if (!c || i < 0 || i >= 16 || !out) return RANDNLA_INVALID_PARAMETERS;
  *out = c->diag[i];
  return RANDNLA_OK;
}

// ---------------------------------------------------------------- RNG (host)
void randnla_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  u32x4 c{{ctr[0], ctr[1], ctr[2], ctr[3]}};
  const u32x4 o = philox4x32(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = o.v[i];
}
void randnla_philox2x32(const uint32_t ctr[2], uint32_t key, uint32_t out[2]) {
  u32x2 c{{ctr[0], ctr[1]}};
  const u32x2 o = philox2x32(c, key);
  out[0] = o.v[0];
  out[1] = o.v[1];
}
void randnla_philox2x64(const uint64_t ctr[2], uint64_t key, uint64_t out[2]) {
  u64x2 c{{ctr[0], ctr[1]}};
  const u64x2 o = philox2x64(c, key);
  out[0] = o.v[0];
  out[1] = o.v[1];
}
void randnla_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  u64x4 c{{ctr[0], ctr[1], ctr[2], ctr[3]}};
  const u64x4 o = philox4x64(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = o.v[i];
}
void randnla_threefry2x64(const uint64_t ctr[2], const uint64_t key[2], uint64_t out[2]) {
  u64x2 c{{ctr[0], ctr[1]}};
  const u64x2 o = threefry2x64(c, key[0], key[1]);
  out[0] = o.v[0];
  out[1] = o.v[1];
}
void randnla_rng_from_seed(int gen, uint64_t seed, randnla_rng* out) {
  if (!out) return;
  const Stream s = stream_from_seed(gen, seed);
  out->gen = s.gen;
  out->key[0] = s.key[0];
  out->key[1] = s.key[1];
  out->ctr[0] = 0;
  out->ctr[1] = 0;
}

int randnla_stream_words(randnla_ctx* c, const randnla_rng* rng, uint64_t first, int64_t n,
                         uint64_t* out) {
  return guard(c, [&] {
    if (n < 0) fail(RANDNLA_NEGATIVE_DIMENSIONS, "n must be non-negative");
    const Stream s = to_stream(rng, GEN_THREEFRY2X64, 0);
    MatOut o(out, n, 1, n, 8, c->stream);
    HIPCK(stream_words(s, first, n, o.as<uint64_t>(), c->stream));
    o.commit();
    sync(c);
  });
}

// --------------------------------------------------------------- sketches
// src/sketch.rs:102-130
int randnla_sketching_operator(randnla_ctx* c, int dist, int64_t rows, int64_t cols,
                               const randnla_rng* rng, int zeta, double* out, int64_t ld) {
  return guard(c, [&] {
    if (rows <= 0 || cols <= 0)
      fail(RANDNLA_INVALID_DIMENSIONS, "Rows and columns must be greater than 0");
    if (dist < 0 || dist > 3) fail(RANDNLA_INVALID_PARAMETERS, "unknown distribution");
    const Stream s = to_stream(rng, GEN_THREEFRY2X64, 0);
    MatOut o(out, rows, cols, ld, 8, c->stream);
    if (dist == RANDNLA_SPARSE_SIGN) {
      if (zeta < 1 || zeta > SPARSE_MAX_ZETA) fail(RANDNLA_INVALID_PARAMETERS, "zeta out of range");
      if (rows > 16384) fail(RANDNLA_INVALID_DIMENSIONS, "sparse-sign rows must be <= 16384");
      // Densify S = S * I_cols.
      DBuf I((size_t)cols * cols * 8, c->stream);
      HIPCK(set_identity(I.as<double>(), cols, cols, cols, c->stream));
      const int64_t need = sparse_sign_work(rows, cols, cols);
      DBuf w((size_t)need * 8, c->stream);
      HIPCK(sparse_sign_apply(s, rows, zeta, I.p, T::F64, cols, cols, cols, o.as<double>(), o.ld,
                              w.as<double>(), need, c->stream));
    } else {
      HIPCK(sketch_panel(dist, s, rows, cols, o.p, T::F64, o.ld, 1.0, c->stream));
    }
    o.commit();
    sync(c);
  });
}

// src/sketch.rs:45-85
int randnla_haar_sample(randnla_ctx* c, int64_t rows, int64_t cols, int attr,
                        const randnla_rng* rng, double* out, int64_t ld) {
  return guard(c, [&] {
    char buf[160];
    int64_t m, n;
    if (attr == RANDNLA_ROW) {
      if (rows > cols) {
        snprintf(buf, sizeof buf,
                 "Cannot have more rows (%lld) than columns (%lld) for row-orthonormal matrix",
                 (long long)rows, (long long)cols);
        fail(RANDNLA_INVALID_DIMENSIONS, buf);
      }
      m = cols; n = rows;
    } else {
      if (cols > rows) {
        snprintf(buf, sizeof buf,
                 "Cannot have more columns (%lld) than rows (%lld) for column-orthonormal matrix",
                 (long long)cols, (long long)rows);
        fail(RANDNLA_INVALID_DIMENSIONS, buf);
      }
      m = rows; n = cols;
    }
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "Rows and columns must be greater than 0");
    const Stream s = to_stream(rng, GEN_THREEFRY2X64, 0);
    hipStream_t st = c->stream;
    DBuf G((size_t)m * n * 8, st);
    HIPCK(sketch_panel(0, s, m, n, G.p, T::F64, m, 1.0, st));
    // Q of Householder QR with columns scaled by sign(R_ii) (+1 for 0):
    // the QR factor with a non-negative diagonal, unique given G.
    DBuf tau = householder(c, G.as<double>(), m, n, false, nullptr);
    DBuf Q((size_t)m * n * 8, st);
    HIPCK(householder_form_q(nullptr, m, n, m, tau.as<double>(), Q.as<double>(), m, n, st));
    const std::vector<double> diag = read_diag(st, G.as<double>(), n, m);
    std::vector<double> sg(n);
    for (int64_t i = 0; i < n; ++i) sg[i] = (diag[i] < 0.0) ? -1.0 : 1.0;
    DBuf dsg((size_t)n * 8, st);
    HIPCK(hipMemcpyAsync(dsg.p, sg.data(), n * 8, hipMemcpyHostToDevice, st));
    HIPCK(scale_columns(Q.as<double>(), m, n, m, dsg.as<double>(), 1, st));
    MatOut o(out, rows, cols, ld, 8, st);
    if (attr == RANDNLA_ROW) HIPCK(transpose_f64(Q.as<double>(), m, n, m, o.as<double>(), o.ld, st));
    else HIPCK(hipMemcpy2DAsync(o.p, o.ld * 8, Q.p, m * 8, m * 8, n, hipMemcpyDeviceToDevice, st));
    o.commit();
    sync(c);
  });
}

int randnla_sketch_apply(randnla_ctx* c, int side, int dist, int zeta, const randnla_rng* rng,
                         const void* A, int dtype, int64_t m, int64_t n, int64_t lda, int64_t d,
                         double* out, int64_t ldo) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0 || d <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    const T ta = dtype == RANDNLA_F32 ? T::F32 : T::F64;
    const size_t esz = ta == T::F32 ? 4 : 8;
    const Stream s = to_stream(rng, GEN_THREEFRY2X64, 0);
    MatIn a(A, m, n, lda, esz, c->stream);
    if (side == RANDNLA_LEFT) {
      MatOut o(out, d, n, ldo, 8, c->stream);
      Tm tm(c, "sketch_apply");
      if (dist == RANDNLA_SPARSE_SIGN) {
        if (zeta < 1 || zeta > SPARSE_MAX_ZETA) fail(RANDNLA_INVALID_PARAMETERS, "zeta out of range");
        if (d > 16384) fail(RANDNLA_INVALID_DIMENSIONS, "sparse-sign rows must be <= 16384");
        const int64_t need = sparse_sign_work(d, m, n);
        DBuf w((size_t)need * 8, c->stream);
        HIPCK(sparse_sign_apply(s, d, zeta, a.p, ta, m, n, a.ld, o.as<double>(), o.ld,
                                w.as<double>(), need, c->stream));
      } else {
        const int64_t need = dense_sketch_work(0, m, n, d);
        DBuf w((size_t)std::max<int64_t>(need, 1) * 8, c->stream);
        HIPCK(dense_sketch_apply(0, dist, s, a.p, ta, m, n, a.ld, d, o.as<double>(), o.ld,
                                 w.as<double>(), need, c->stream));
      }
      o.commit();
    } else {
      MatOut o(out, m, d, ldo, 8, c->stream);
      Tm tm(c, "sketch_apply");
      if (dist == RANDNLA_SPARSE_SIGN) {
        // A * Omega with Omega (n x d) sparse-sign: densify Omega, then GEMM.
        if (n > 16384) fail(RANDNLA_INVALID_DIMENSIONS, "sparse-sign rows must be <= 16384");
        DBuf I((size_t)d * d * 8, c->stream), Om((size_t)n * d * 8, c->stream);
        HIPCK(set_identity(I.as<double>(), d, d, d, c->stream));
        const int64_t need = sparse_sign_work(n, d, d);
        DBuf w((size_t)need * 8, c->stream);
        HIPCK(sparse_sign_apply(s, n, zeta, I.p, T::F64, d, d, d, Om.as<double>(), n,
                                w.as<double>(), need, c->stream));
        gemm64(c, m, d, n, a.p, ta, a.ld, false, Om.p, T::F64, n, o.p, T::F64, o.ld);
      } else {
        HIPCK(dense_sketch_apply(1, dist, s, a.p, ta, m, n, a.ld, d, o.as<double>(), o.ld, nullptr, 0,
                                 c->stream));
      }
      o.commit();
    }
    sync(c);
  });
}

// No reference counterpart: randnla_sketch_apply for callers that cannot
// allocate.  Same operators and results; the partial sums of the LEFT product
// (sparse-sign row chunks, dense split-K slabs) go to the fixed per-GPU
// scratch array instead of allocated workspace, with as many slabs as fit
// (d * n <= 2^23 doubles, else RANDNLA_INVALID_DIMENSIONS).  A and out must
// be device memory; nothing is staged, timed or synchronised: the work is
// queued on the context's stream.  Every slab is written before it is read,
// so nothing needs clearing between calls; calls that overlap on one GPU
// (other streams or contexts) race on the array.
int randnla_sketch_apply_noalloc(randnla_ctx* c, int side, int dist, int zeta,
                                 const randnla_rng* rng, const void* A, int dtype, int64_t m,
                                 int64_t n, int64_t lda, int64_t d, double* out, int64_t ldo) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0 || d <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    if (!is_device_ptr(A) || !is_device_ptr(out))
      fail(RANDNLA_INVALID_PARAMETERS, "the no-allocation entry point takes device memory only");
    const T ta = dtype == RANDNLA_F32 ? T::F32 : T::F64;
    const int64_t rows = side == RANDNLA_LEFT ? d : m, cols = side == RANDNLA_LEFT ? n : d;
    if (lda < m || ldo < rows) fail(RANDNLA_INVALID_DIMENSIONS, "leading dimension too small");
    (void)cols;
    const Stream s = to_stream(rng, GEN_THREEFRY2X64, 0);
    double* scr = nullptr;
    HIPCK(scratch_lookup(c->device, &scr));
    if (side == RANDNLA_LEFT) {
      if (d * n > NOALLOC_SCRATCH_DOUBLES) {
        char b[160];
        snprintf(b, sizeof b, "d * n = %lld exceeds the no-allocation limit of %lld doubles",
                 (long long)(d * n), (long long)NOALLOC_SCRATCH_DOUBLES);
        fail(RANDNLA_INVALID_DIMENSIONS, b);
      }
      if (dist == RANDNLA_SPARSE_SIGN) {
        if (zeta < 1 || zeta > SPARSE_MAX_ZETA) fail(RANDNLA_INVALID_PARAMETERS, "zeta out of range");
        if (d > 16384) fail(RANDNLA_INVALID_DIMENSIONS, "sparse-sign rows must be <= 16384");
        HIPCK(sparse_sign_apply(s, d, zeta, A, ta, m, n, lda, out, ldo, scr,
                                NOALLOC_SCRATCH_DOUBLES, c->stream));
      } else {
        HIPCK(dense_sketch_apply(0, dist, s, A, ta, m, n, lda, d, out, ldo, scr,
                                 NOALLOC_SCRATCH_DOUBLES, c->stream));
      }
    } else {
      if (dist == RANDNLA_SPARSE_SIGN)
        fail(RANDNLA_INVALID_PARAMETERS,
             "A * Omega with a sparse-sign Omega needs an allocated dense Omega");
      HIPCK(dense_sketch_apply(1, dist, s, A, ta, m, n, lda, d, out, ldo, nullptr, 0, c->stream));
    }
  });
}

// ------------------------------------------------------------- low rank
// src/lora_drivers.rs:30-69
int randnla_rand_svd(randnla_ctx* c, const void* A, int dtype, int64_t m, int64_t n, int64_t lda,
                     int64_t k, double epsilon, int64_t s, const randnla_opts* opts, double* U,
                     int64_t ldu, double* S, double* Vt, int64_t ldvt, int64_t* r_out) {
  return guard(c, [&] {
    check_svd_params(m, n, k, epsilon, s);
    const randnla_opts o = opts_or_default(opts);
    if (o.power_iters < 0) fail(RANDNLA_INVALID_PARAMETERS, "power_iters must be >= 0");
    if (o.faithful && o.precision != RANDNLA_FP64)
      fail(RANDNLA_INVALID_PARAMETERS, "the reference-faithful method runs in FP64 only");
    // The reference's Q has min(k + s, m, n) columns (tsog1's stabilisers and
    // the thin QR shrink it) and r = min(k, ncols(Q)) (src/lora_drivers.rs:51).
    const int64_t l = std::min<int64_t>(k + s, std::min(m, n));
    const int64_t r = std::min(k, l);
    const T ta = dtype == RANDNLA_F32 ? T::F32 : T::F64;
    const size_t esz = ta == T::F32 ? 4 : 8;
    const Stream st_rng = to_stream(&o.rng, GEN_PHILOX4X32, 42);
    hipStream_t st = c->stream;
    MatIn a(A, m, n, lda, esz, st);
    MatOut u(U, m, r, ldu, 8, st), sv(S, r, 1, r, 8, st), vt(Vt, r, n, ldvt, 8, st);
    DBuf Bt((size_t)n * l * 8, st);
    if (o.precision == RANDNLA_FP64) {
      DBuf Q((size_t)m * l * 8, st);
      const int64_t lq = range_finder(c, o, a.p, ta, m, n, a.ld, l, Q.as<double>());
      if (lq != l) fail(RANDNLA_COMPUTATION_ERROR, "range finder rank mismatch");
      gemm64(c, n, l, m, a.p, ta, a.ld, true, Q.p, T::F64, m, Bt.p, T::F64, n, 1.0, "pass_bt");
      svd_tail(c, Q.p, T::F64, m, m, n, l, Bt.as<double>(), r, u.as<double>(), u.ld, sv.as<double>(),
               vt.as<double>(), vt.ld);
    } else if (o.precision == RANDNLA_BF16 || o.precision == RANDNLA_FP8) {
      const float* A32 = (const float*)a.p;
      int64_t ld32 = a.ld;
      DBuf conv;
      if (ta == T::F64) {
        conv = DBuf((size_t)m * n * 4, st);
        HIPCK(convert(a.p, T::F64, m, n, a.ld, conv.p, T::F32, m, 1.0, st));
        A32 = conv.as<float>();
        ld32 = m;
      }
      const int64_t mp = r8(m);
      DBuf Q32((size_t)mp * l * 4, st);
      rand_svd_mixed_core(c, A32, m, n, ld32, l, o.power_iters, st_rng, Q32.as<float>(), mp,
                          Bt.as<double>(), o.precision == RANDNLA_FP8);
      svd_tail(c, Q32.p, T::F32, mp, m, n, l, Bt.as<double>(), r, u.as<double>(), u.ld,
               sv.as<double>(), vt.as<double>(), vt.ld);
    } else {
      fail(RANDNLA_INVALID_PARAMETERS, "unsupported precision");
    }
    u.commit();
    sv.commit();
    vt.commit();
    sync(c);
    if (r_out) *r_out = r;
  });
}

// src/lora_helpers.rs:37-44.  Q has min(k, m, n) columns (the reference's
// thin QR of A tsog1(A, k, 2, 1)); the rest of the m x k output is zero.
// Intended method (Gaussian test matrix, q rounds) or, with opts->faithful,
// the reference's computation.
int randnla_rf1(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda, int64_t k,
                const randnla_opts* opts, double* Q, int64_t ldq) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0 || k <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    const randnla_opts o = opts_or_default(opts);
    if (o.power_iters < 0) fail(RANDNLA_INVALID_PARAMETERS, "power_iters must be >= 0");
    MatIn a(A, m, n, lda, 8, c->stream);
    MatOut q(Q, m, k, ldq, 8, c->stream);
    DBuf Qt((size_t)m * k * 8, c->stream);
    const int64_t kc = range_finder(c, o, a.p, T::F64, m, n, a.ld, k, Qt.as<double>());
    HIPCK(fill_f64(q.as<double>(), m, k, q.ld, 0.0, c->stream));
    HIPCK(hipMemcpy2DAsync(q.p, q.ld * 8, Qt.p, m * 8, m * 8, kc, hipMemcpyDeviceToDevice, c->stream));
    q.commit();
    sync(c);
  });
}

// src/lora_helpers.rs:17-23.  Q (m x kc) and B = Q^T A (kc x n) with
// kc = min(k, m, n); the remaining columns of Q and rows of B are zero.
int randnla_qb1(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda, int64_t k,
                double epsilon, const randnla_opts* opts, double* Q, int64_t ldq, double* B,
                int64_t ldb) {
  (void)epsilon;   // ignored, as in the reference (src/lora_helpers.rs:18-19)
  return guard(c, [&] {
    if (m <= 0 || n <= 0 || k <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    const randnla_opts o = opts_or_default(opts);
    if (o.power_iters < 0) fail(RANDNLA_INVALID_PARAMETERS, "power_iters must be >= 0");
    MatIn a(A, m, n, lda, 8, c->stream);
    MatOut q(Q, m, k, ldq, 8, c->stream), b(B, k, n, ldb, 8, c->stream);
    DBuf Qt((size_t)m * k * 8, c->stream);
    const int64_t kc = range_finder(c, o, a.p, T::F64, m, n, a.ld, k, Qt.as<double>());
    HIPCK(fill_f64(q.as<double>(), m, k, q.ld, 0.0, c->stream));
    HIPCK(fill_f64(b.as<double>(), k, n, b.ld, 0.0, c->stream));
    HIPCK(hipMemcpy2DAsync(q.p, q.ld * 8, Qt.p, m * 8, m * 8, kc, hipMemcpyDeviceToDevice, c->stream));
    // B = Q^T A  (kc x n)
    gemm64(c, kc, n, m, Qt.p, T::F64, m, true, a.p, T::F64, a.ld, b.p, T::F64, b.ld, 1.0, "pass_bt");
    q.commit();
    b.commit();
    sync(c);
  });
}

// src/lora_helpers.rs:58-105.  Intended semantics (opts->faithful == 0): the
// loop starts from the generated test matrix, re-orthonormalisation after
// every passes_per_stab products; S is n x k.  opts->faithful: the
// reference's computation (the even-branch Gaussian is discarded, Stabilizer
// is the full-pivot LU factor L), S has tsog1 columns as the reference's
// (min(n, min(m, k)) with passes_per_stab = 1 and num_passes >= 2), the
// rest of the n x k output is zero; the Gaussian comes from the reference's
// stream ThreeFry2x64 seed_from_u64(0) (opts->rng is not used).
int randnla_tsog1(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda, int64_t k,
                  int num_passes, int passes_per_stab, const randnla_opts* opts, double* S,
                  int64_t lds) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0 || k <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    if (passes_per_stab <= 0) fail(RANDNLA_INVALID_PARAMETERS, "passes_per_stab must be positive");
    if (num_passes < 0) fail(RANDNLA_INVALID_PARAMETERS, "num_passes must be non-negative");
    const randnla_opts o = opts_or_default(opts);
    hipStream_t st = c->stream;
    MatIn a(A, m, n, lda, 8, st);
    MatOut out(S, n, k, lds, 8, st);
    DBuf Sn((size_t)n * k * 8, st);
    if (o.faithful) {
      const Stream s = stream_from_seed(GEN_THREEFRY2X64, 0);
      tsog1_faithful(c, a.p, T::F64, m, n, a.ld, k, num_passes, passes_per_stab, s,
                     Sn.as<double>());
    } else {
      const Stream s = to_stream(&o.rng, GEN_PHILOX4X32, 42);
      tsog1_core(c, a.p, m, n, a.ld, k, num_passes, passes_per_stab, s, Sn.as<double>());
    }
    HIPCK(hipMemcpy2DAsync(out.p, out.ld * 8, Sn.p, n * 8, n * 8, k, hipMemcpyDeviceToDevice, st));
    out.commit();
    sync(c);
  });
}

// src/lora_helpers.rs:131-133: Q (m x min(m, n)) of the Householder thin QR
// with nalgebra's sign convention (R_ii >= 0); R (min(m, n) x n, may be NULL)
// with X = Q R.  Wide X: Q is m x m.
int randnla_orth(randnla_ctx* c, const double* X, int64_t m, int64_t n, int64_t ldx, double* Q,
                 int64_t ldq, double* R, int64_t ldr) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    const int64_t kq = std::min(m, n);
    MatIn x(X, m, n, ldx, 8, c->stream);
    MatOut q(Q, m, kq, ldq, 8, c->stream), r(R, kq, n, ldr, 8, c->stream);
    orth_householder(c, x.as_d(), m, n, x.ld, q.as<double>(), q.ld, R ? r.as<double>() : nullptr,
                     r.ld);
    q.commit();
    r.commit();
    sync(c);
  });
}

// CholeskyQR2 of a tall X (m x n, m >= n): Q (m x n) and R = R2 R1 (n x n,
// may be NULL), FP64 Gram and Cholesky.  A column whose Cholesky pivot falls
// to <= 16 max(n, 16) eps times its diagonal is dependent: it gets a zero row
// in R1 and a unit-vector completion in Q, which the second pass
// re-orthogonalises.  No reference counterpart (the reference's Orth is a
// Householder QR); this is the orthonormalisation of the range finder.
int randnla_cholqr2(randnla_ctx* c, const double* X, int64_t m, int64_t n, int64_t ldx, double* Q,
                    int64_t ldq, double* R, int64_t ldr) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    if (n > m) fail(RANDNLA_INVALID_DIMENSIONS, "CholeskyQR2 needs m >= n");
    MatIn x(X, m, n, ldx, 8, c->stream);
    MatOut q(Q, m, n, ldq, 8, c->stream), r(R, n, n, ldr, 8, c->stream);
    DBuf Rt((size_t)n * n * 8, c->stream);
    cholqr2(c, x.p, T::F64, m, n, x.ld, q.p, T::F64, q.ld, Rt.as<double>());
    if (R) HIPCK(hipMemcpy2DAsync(r.p, r.ld * 8, Rt.p, n * 8, n * 8, n, hipMemcpyDeviceToDevice, c->stream));
    q.commit();
    r.commit();
    sync(c);
  });
}

// src/lora_helpers.rs:144-146: L (m x min(m, n)), the unit lower factor of
// the full-pivot LU of X (nalgebra 0.33 FullPivLU), rows in pivoted order.
int randnla_stabilizer(randnla_ctx* c, const double* X, int64_t m, int64_t n, int64_t ldx,
                       double* Y, int64_t ldy) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    const int64_t kl = std::min(m, n);
    MatIn x(X, m, n, ldx, 8, c->stream);
    MatOut y(Y, m, kl, ldy, 8, c->stream);
    stabilizer_lu(c, x.as_d(), m, n, x.ld, y.as<double>(), y.ld);
    y.commit();
    sync(c);
  });
}

// ------------------------------------------------------ eigendecompositions
// src/lora_drivers.rs:87-151.  Q from the range finder (QB1's RF1), C = Q^T A Q
// from its lower triangle, the eigenpairs of C, the k of largest |lambda|, and
// V = Q U.  C's eigenpairs come from one-sided Jacobi on C + mu I with
// mu = ||C||_F >= rho(C): that matrix is positive semi-definite, so its right
// singular vectors are C's eigenvectors and lambda = sigma - mu (a symmetric
// pair lambda, -lambda stays separated, which it would not be in an SVD of C).
int randnla_rand_evd1(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                      int64_t k, double epsilon, int64_t s, const randnla_opts* opts, double* V,
                      int64_t ldv, double* lambda, int64_t* r_out) {
  return guard(c, [&] {
    if (k == 0) fail(RANDNLA_INVALID_PARAMETERS, "Rank k must be positive, current input is 0");
    if (!(epsilon > 0.0)) {
      char b[128];
      snprintf(b, sizeof b, "Epsilon must be positive, current input is %g", epsilon);
      fail(RANDNLA_INVALID_PARAMETERS, b);
    }
    if (s == 0)
      fail(RANDNLA_INVALID_PARAMETERS, "Oversampling parameter s must be positive, current input is 0");
    if (k < 0 || s < 0) fail(RANDNLA_NEGATIVE_DIMENSIONS, "k and s must be non-negative");
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "matrix dimensions must be positive");
    if (m != n) fail(RANDNLA_NOT_HERMITIAN, "Input matrix is not Hermitian");
    const randnla_opts o = opts_or_default(opts);
    if (o.power_iters < 0) fail(RANDNLA_INVALID_PARAMETERS, "power_iters must be >= 0");
    hipStream_t st = c->stream;
    MatIn a(A, n, n, lda, 8, st);
    {
      DBuf flag(sizeof(int), st);
      HIPCK(sym_mismatch(a.as_d(), n, a.ld, flag.as<int>(), st));
      int bad = 0;
      d2h(st, &bad, flag.p, sizeof(int));
      if (bad) fail(RANDNLA_NOT_HERMITIAN, "Input matrix is not Hermitian");
    }
    const int64_t l = std::min(k + s, n), r = std::min(k, l);
    const Stream st_rng = to_stream(&o.rng, GEN_PHILOX4X32, 42);
    DBuf Q((size_t)n * l * 8, st), AQ((size_t)n * l * 8, st), Cm((size_t)l * l * 8, st);
    range_finder_f64(c, a.p, T::F64, n, n, a.ld, l, o.power_iters, st_rng, Q.as<double>());
    gemm64(c, n, l, n, a.p, T::F64, a.ld, true, Q.p, T::F64, n, AQ.p, T::F64, n, 1.0, "pass_bt");
    gemm64(c, l, l, n, Q.p, T::F64, n, true, AQ.p, T::F64, n, Cm.p, T::F64, l, 1.0, "evd_c");
    double mu = 0.0;
    {
      DBuf part(1024 * 8, st), nrm(8, st);
      HIPCK(sqnorm(Cm.as<double>(), l * l, nrm.as<double>(), part.as<double>(), st));
      d2h(st, &mu, nrm.p, 8);
      mu = std::sqrt(mu);
    }
    if (!std::isfinite(mu)) fail(RANDNLA_COMPUTATION_ERROR, "non-finite entries in Q^T A Q");
    DBuf W((size_t)l * l * 8, st), Vr((size_t)l * l * 8, st), sv((size_t)l * 8, st), cnt(8, st);
    HIPCK(sym_from_lower(Cm.as<double>(), l, l, W.as<double>(), l, mu, st));
    HIPCK(set_identity(Vr.as<double>(), l, l, l, st));
    int sweeps = 0;
    {
      Tm tm(c, "jacobi");
      HIPCK(jacobi_svd(W.as<double>(), l, l, Vr.as<double>(), l, sv.as<double>(), cnt.as<int>(),
                       30, jacobi_tol(l), st, &sweeps));
    }
    c->diag[0] = sweeps;
    DBuf Ur((size_t)l * l * 8, st), Vs((size_t)l * l * 8, st), ss((size_t)l * 8, st);
    HIPCK(svd_finalize(W.as<double>(), l, l, Vr.as<double>(), l, sv.as<double>(), Ur.as<double>(),
                       l, Vs.as<double>(), l, ss.as<double>(), st));
    std::vector<double> lam(l);
    d2h(st, lam.data(), ss.p, (size_t)l * 8);
    for (auto& x : lam) x -= mu;
    const std::vector<int64_t> idx = abs_desc_order(lam);
    std::vector<double> lr(r);
    for (int64_t i = 0; i < r; ++i) lr[i] = lam[idx[i]];
    DBuf ib((size_t)r * 8, st), Us((size_t)l * r * 8, st);
    HIPCK(hipMemcpyAsync(ib.p, idx.data(), (size_t)r * 8, hipMemcpyHostToDevice, st));
    HIPCK(gather_cols(Vs.as<double>(), l, l, ib.as<int64_t>(), r, Us.as<double>(), l, st));
    MatOut v(V, n, r, ldv, 8, st);
    gemm64(c, n, r, l, Q.p, T::F64, n, false, Us.p, T::F64, l, v.p, T::F64, v.ld, 1.0, "evd_v");
    v.commit();
    sync(c);
    put_values(c, lambda, lr, r);
    if (r_out) *r_out = r;
  });
}

// src/lora_drivers.rs:167-224.  S = tsog1(A, k + s, 3, 1); Y = A S shifted by
// nu S with nu = sqrt(n) eps ||Y||_F; R from the Cholesky factorisation of
// S^T Y (lower triangle read); the SVD of B = Y R^{-1}; lambda = sigma^2 - nu
// over the sigma^2 > nu, at most k of them.  The reference's positive
// semi-definiteness test (every eigenvalue of A's lower-triangle symmetric
// matrix >= 0) is decided here by a Cholesky factorisation of that matrix
// shifted by n eps max|a_ii|; a zero diagonal accepts only the zero matrix.
int randnla_rand_evd2(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                      int64_t k, int64_t s, const randnla_opts* opts, double* V, int64_t ldv,
                      double* lambda, int64_t* r_out) {
  return guard(c, [&] {
    if (k == 0) fail(RANDNLA_INVALID_PARAMETERS, "Rank k must be positive, current input is 0");
    if (k < 0 || s < 0) fail(RANDNLA_NEGATIVE_DIMENSIONS, "k and s must be non-negative");
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "matrix dimensions must be positive");
    if (m != n) fail(RANDNLA_INVALID_DIMENSIONS, "rand_evd2 needs a square matrix");
    const int64_t l = k + s;
    if (l > n) fail(RANDNLA_INVALID_PARAMETERS, "k + s must not exceed n");
    const randnla_opts o = opts_or_default(opts);
    hipStream_t st = c->stream;
    const double eps = 2.220446049250313e-16;
    MatIn a(A, n, n, lda, 8, st);
    {
      DBuf G((size_t)n * n * 8, st), dm(8, st), flags((size_t)(n + 1) * 4, st);
      double dmax = 0.0;
      HIPCK(absmax(a.p, T::F64, 1, n, a.ld + 1, dm.as<double>(), st));
      d2h(st, &dmax, dm.p, 8);
      if (dmax == 0.0) {
        double amax = 0.0;
        HIPCK(sym_from_lower(a.as_d(), n, a.ld, G.as<double>(), n, 0.0, st));
        HIPCK(absmax(G.p, T::F64, n, n, n, dm.as<double>(), st));
        d2h(st, &amax, dm.p, 8);
        if (amax != 0.0)
          fail(RANDNLA_NOT_POSITIVE_SEMI_DEFINITE, "Matrix is not positive semi-definite");
      } else if (std::isfinite(dmax)) {
        {
          Tm tm(c, "psd_check");
          HIPCK(sym_from_lower(a.as_d(), n, a.ld, G.as<double>(), n, (double)n * eps * dmax, st));
          HIPCK(cholesky_upper(G.as<double>(), n, n, 0.0, flags.as<int>(), flags.as<int>() + n,
                               st));
        }
        int ndef = 0;
        d2h(st, &ndef, flags.as<int>() + n, sizeof(int));
        if (ndef)
          fail(RANDNLA_NOT_POSITIVE_SEMI_DEFINITE, "Matrix is not positive semi-definite");
      }
    }
    const Stream st_rng = to_stream(&o.rng, GEN_PHILOX4X32, 42);
    DBuf S((size_t)n * l * 8, st), Y((size_t)n * l * 8, st);
    tsog1_core(c, a.p, n, n, a.ld, l, 3, 1, st_rng, S.as<double>());
    gemm64(c, n, l, n, a.p, T::F64, a.ld, false, S.p, T::F64, n, Y.p, T::F64, n, 1.0, "evd_y");
    double nu = 0.0;
    {
      DBuf part(1024 * 8, st), nrm(8, st);
      HIPCK(sqnorm(Y.as<double>(), n * l, nrm.as<double>(), part.as<double>(), st));
      d2h(st, &nu, nrm.p, 8);
      nu = std::sqrt((double)n) * eps * std::sqrt(nu);
    }
    HIPCK(axpby(n * l, nu, S.as<double>(), 1.0, Y.as<double>(), st));
    DBuf SY((size_t)l * l * 8, st), G((size_t)l * l * 8, st), X((size_t)l * l * 8, st),
        flags((size_t)(l + 1) * 4, st);
    gemm64(c, l, l, n, S.p, T::F64, n, true, Y.p, T::F64, n, SY.p, T::F64, l, 1.0, "evd_sy");
    HIPCK(sym_from_lower(SY.as<double>(), l, l, G.as<double>(), l, 0.0, st));
    {
      Tm tm(c, "cholesky");
      HIPCK(cholesky_inverse_upper(G.as<double>(), l, l, 0.0, flags.as<int>(),
                                   flags.as<int>() + l, X.as<double>(), l, st));
    }
    int ndef = 0;
    d2h(st, &ndef, flags.as<int>() + l, sizeof(int));
    if (ndef) fail(RANDNLA_MATRIX_DECOMPOSITION_ERROR, "Cholesky Decomposition Failed");
    DBuf B((size_t)n * l * 8, st), U((size_t)n * l * 8, st), sig((size_t)l * 8, st),
        Vb((size_t)l * l * 8, st);
    gemm64(c, n, l, l, Y.p, T::F64, n, false, X.p, T::F64, l, B.p, T::F64, n, 1.0, "trmm");
    small_svd(c, B.as<double>(), n, l, U.as<double>(), sig.as<double>(), Vb.as<double>());
    std::vector<double> hs(l), lam;
    d2h(st, hs.data(), sig.p, (size_t)l * 8);
    for (double x : hs)
      if (x > 0.0) lam.push_back(x * x);
    int64_t above = 0;
    for (double x : lam) above += x > nu;
    const int64_t r = std::min(k, above);
    std::vector<double> lr(std::max<int64_t>(r, 1));
    for (int64_t i = 0; i < r; ++i) lr[i] = lam[i] - nu;
    MatOut v(V, n, r, ldv, 8, st);
    if (r > 0)
      HIPCK(hipMemcpy2DAsync(v.p, v.ld * 8, U.p, n * 8, n * 8, r, hipMemcpyDeviceToDevice, st));
    v.commit();
    sync(c);
    put_values(c, lambda, lr, r);
    if (r_out) *r_out = r;
  });
}

// ------------------------------------------ interpolative decompositions
// src/pivot_decompositions.rs:194-269: the first k steps of Householder QRCP.
// Q (m x k, may be NULL), R (k x n, upper trapezoidal), perm (n).
int randnla_economic_qrcp(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                          int64_t k, double* Q, int64_t ldq, double* R, int64_t ldr,
                          int64_t* perm) {
  return guard(c, [&] {
    check_economic_k(k, m, n);
    hipStream_t st = c->stream;
    MatIn a(A, m, n, lda, 8, st);
    DBuf W((size_t)m * n * 8, st), P((size_t)n * 8, st);
    HIPCK(hipMemcpy2DAsync(W.p, m * 8, a.p, a.ld * 8, m * 8, n, hipMemcpyDeviceToDevice, st));
    DBuf tau = qrcp_steps(c, W.as<double>(), m, n, k, P.as<int64_t>());
    MatOut ro(R, k, n, ldr, 8, st);
    HIPCK(upper_copy(W.as<double>(), k, n, m, ro.as<double>(), ro.ld, st));
    if (Q) {
      MatOut qo(Q, m, k, ldq, 8, st);
      HIPCK(householder_form_q(nullptr, m, n, m, tau.as<double>(), qo.as<double>(), qo.ld, k, st));
      qo.commit();
    }
    ro.commit();
    put_indices(c, perm, P.as<int64_t>(), n);
    sync(c);
  });
}

// src/id.rs:248-305.  Column: X (k x n), J (k) with Y ~ Y[:, J] X.  Row:
// X (m x k), J (k) row indices with Y ~ X Y[J, :].
int randnla_osid_qrcp(randnla_ctx* c, const double* Y, int64_t m, int64_t n, int64_t ldy,
                      int64_t k, int attr, double* X, int64_t ldx, int64_t* J) {
  return guard(c, [&] {
    check_id_k(k, m, n);
    hipStream_t st = c->stream;
    MatIn y(Y, m, n, ldy, 8, st);
    const bool col = attr != RANDNLA_ROW;
    MatOut xo(X, col ? k : m, col ? n : k, ldx, 8, st);
    DBuf Jd((size_t)k * 8, st);
    osid_qrcp_dev(c, y.as_d(), m, n, y.ld, k, attr, xo.as<double>(), xo.ld, Jd.as<int64_t>());
    xo.commit();
    put_indices(c, J, Jd.as<int64_t>(), k);
    sync(c);
  });
}

// src/id.rs:205-234 (opts: the Column sketch's distribution and stream, the
// Row branch's tsog1 stream).
int randnla_osid_randomised(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                            int64_t k, int attr, const randnla_opts* opts, double* X,
                            int64_t ldx, int64_t* J) {
  return guard(c, [&] {
    check_id_k(k, m, n);
    hipStream_t st = c->stream;
    MatIn a(A, m, n, lda, 8, st);
    const bool col = attr != RANDNLA_ROW;
    MatOut xo(X, col ? k : m, col ? n : k, ldx, 8, st);
    DBuf Jd((size_t)k * 8, st);
    osid_randomised_dev(c, opts, a.as_d(), m, n, a.ld, k, attr, xo.as<double>(), xo.ld,
                        Jd.as<int64_t>());
    xo.commit();
    put_indices(c, J, Jd.as<int64_t>(), k);
    sync(c);
  });
}

namespace {
void two_sided(randnla_ctx* c, const randnla_opts* opts, bool randomised, const double* A,
               int64_t m, int64_t n, int64_t lda, int64_t k, double* Z, int64_t ldz, int64_t* I,
               int64_t* J, double* X, int64_t ldx) {
  check_id_k(k, m, n);
  hipStream_t st = c->stream;
  MatIn a(A, m, n, lda, 8, st);
  MatOut xo(X, k, n, ldx, 8, st), zo(Z, m, k, ldz, 8, st);
  DBuf Jd((size_t)k * 8, st), Id((size_t)k * 8, st), AC((size_t)m * k * 8, st);
  if (randomised)
    osid_randomised_dev(c, opts, a.as_d(), m, n, a.ld, k, RANDNLA_COLUMN, xo.as<double>(), xo.ld,
                        Jd.as<int64_t>());
  else
    osid_qrcp_dev(c, a.as_d(), m, n, a.ld, k, RANDNLA_COLUMN, xo.as<double>(), xo.ld,
                  Jd.as<int64_t>());
  HIPCK(gather_cols(a.as_d(), m, a.ld, Jd.as<int64_t>(), k, AC.as<double>(), m, st));
  if (randomised)
    osid_randomised_dev(c, opts, AC.as<double>(), m, k, m, k, RANDNLA_ROW, zo.as<double>(), zo.ld,
                        Id.as<int64_t>());
  else
    osid_qrcp_dev(c, AC.as<double>(), m, k, m, k, RANDNLA_ROW, zo.as<double>(), zo.ld,
                  Id.as<int64_t>());
  xo.commit();
  zo.commit();
  put_indices(c, I, Id.as<int64_t>(), k);
  put_indices(c, J, Jd.as<int64_t>(), k);
  sync(c);
}

void cur_entry(randnla_ctx* c, const randnla_opts* opts, bool randomised, const double* A,
               int64_t m, int64_t n, int64_t lda, int64_t k, int64_t* J, double* U, int64_t ldu,
               int64_t* I) {
  check_id_k(k, m, n);
  hipStream_t st = c->stream;
  MatIn a(A, m, n, lda, 8, st);
  MatOut uo(U, k, k, ldu, 8, st);
  DBuf Jd((size_t)k * 8, st), Id((size_t)k * 8, st);
  cur_dev(c, opts, randomised, a.as_d(), m, n, a.ld, k, Jd.as<int64_t>(), uo.as<double>(), uo.ld,
          Id.as<int64_t>());
  uo.commit();
  put_indices(c, J, Jd.as<int64_t>(), k);
  put_indices(c, I, Id.as<int64_t>(), k);
  sync(c);
}
}  // namespace

// src/id.rs:125-132: A ~ Z A[I, J] X.
int randnla_two_sided_id(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                         int64_t k, double* Z, int64_t ldz, int64_t* I, int64_t* J, double* X,
                         int64_t ldx) {
  return guard(c, [&] { two_sided(c, nullptr, false, A, m, n, lda, k, Z, ldz, I, J, X, ldx); });
}

// src/id.rs:94-101.
int randnla_two_sided_id_randomised(randnla_ctx* c, const double* A, int64_t m, int64_t n,
                                    int64_t lda, int64_t k, const randnla_opts* opts, double* Z,
                                    int64_t ldz, int64_t* I, int64_t* J, double* X,
                                    int64_t ldx) {
  return guard(c, [&] { two_sided(c, opts, true, A, m, n, lda, k, Z, ldz, I, J, X, ldx); });
}

// src/id.rs:33-76: A ~ A[:, J] U A[I, :].
int randnla_cur(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda, int64_t k,
                int64_t* J, double* U, int64_t ldu, int64_t* I) {
  return guard(c, [&] { cur_entry(c, nullptr, false, A, m, n, lda, k, J, U, ldu, I); });
}

// src/id.rs:154-197.
int randnla_cur_randomised(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                           int64_t k, const randnla_opts* opts, int64_t* J, double* U,
                           int64_t ldu, int64_t* I) {
  return guard(c, [&] { cur_entry(c, opts, true, A, m, n, lda, k, J, U, ldu, I); });
}

// --------------------------------------------------------- least squares
// src/sketch_and_precondition.rs:26-59
int randnla_blendenpik(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                       const double* b, double epsilon, int64_t l, double sf,
                       const randnla_opts* opts, double* x, randnla_solve_info* info) {
  return guard(c, [&] {
    check_ls(m, n, epsilon, l, sf);
    const randnla_opts o = opts_or_default(opts);
    const Stream s = to_stream(opts ? &o.rng : nullptr, GEN_THREEFRY2X64, 0);
    hipStream_t st = c->stream;
    const int64_t d = (sf * (double)n > (double)m) ? m : (int64_t)floor(sf * (double)n);
    MatIn a(A, m, n, lda, 8, st), bb(b, m, 1, m, 8, st);
    MatOut xo(x, n, 1, n, 8, st);
    // [S A | S b]: the QR's reflectors turn the last column into Q^T S b
    DBuf SA((size_t)d * (n + 1) * 8, st);
    apply_sketch_ab(c, o, s, d, a.as_d(), m, n, a.ld, (const double*)bb.p, SA.as<double>(), d);
    DBuf tau = householder(c, SA.as<double>(), d, n, false, nullptr, 1);
    DBuf R((size_t)n * n * 8, st), Rinv((size_t)n * n * 8, st);
    HIPCK(upper_copy(SA.as<double>(), n, n, d, R.as<double>(), n, st));
    // Singular R: the reference's solve_upper_triangular(...).unwrap() panics.
    const std::vector<double> dg = read_diag(st, R.as<double>(), n, n);
    for (int64_t i = 0; i < n; ++i)
      if (dg[i] == 0.0) fail(RANDNLA_SINGULAR_MATRIX, "sketched matrix is rank deficient");
    HIPCK(trinv_upper(R.as<double>(), n, n, Rinv.as<double>(), n, st));
    // z0 = (Q^T S b)[0:n]
    DBuf z((size_t)n * 8, st), tn((size_t)n * 8, st), wk((size_t)130 * n * 8, st);
    HIPCK(hipMemcpyAsync(z.p, SA.as<double>() + n * d, n * 8, hipMemcpyDeviceToDevice, st));
    PrecOp op{c, a.as_d(), m, n, a.ld, Rinv.as<double>(), n, true, tn.as<double>(), wk.as<double>()};
    SolveOut so;
    if (o.solver == RANDNLA_LSQR)
      so = run_lsqr(c, op, (const double*)bb.p, 0.0, epsilon, epsilon, 1e8, l, false, true,
                    z.as<double>(), nullptr, nullptr);
    else
      so = run_cgls(c, op, (const double*)bb.p, epsilon, l, z.as<double>());
    HIPCK(trmv_upper(Rinv.as<double>(), n, n, false, z.as<double>(), xo.as<double>(), st));
    xo.commit();
    sync(c);
    fill_info(info, so);
  });
}

// src/sketch_and_precondition.rs:82-119
int randnla_lsrn(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                 const double* b, double epsilon, int64_t l, double sf, const randnla_opts* opts,
                 double* x, randnla_solve_info* info) {
  return guard(c, [&] {
    check_ls(m, n, epsilon, l, sf);
    const randnla_opts o = opts_or_default(opts);
    const Stream s = to_stream(opts ? &o.rng : nullptr, GEN_THREEFRY2X64, 0);
    hipStream_t st = c->stream;
    const int64_t d = (sf * (double)n > (double)m) ? m : (int64_t)floor(sf * (double)n);
    MatIn a(A, m, n, lda, 8, st), bb(b, m, 1, m, 8, st);
    MatOut xo(x, n, 1, n, 8, st);
    DBuf SA((size_t)d * n * 8, st);
    apply_sketch(c, o, s, d, a.as_d(), m, n, a.ld, SA.as<double>(), d);
    DBuf sig((size_t)n * 8, st), V((size_t)n * n * 8, st), isg((size_t)n * 8, st);
    small_svd(c, SA.as<double>(), d, n, nullptr, sig.as<double>(), V.as<double>());
    // N = V Sigma^+ (zero singular values -> 0), src/sketch_and_precondition.rs:113-114
    HIPCK(inv_sigma(sig.as<double>(), n, 0.0, 0.0, 0, isg.as<double>(), nullptr, st));
    HIPCK(scale_columns(V.as<double>(), n, n, n, isg.as<double>(), 1, st));
    DBuf y((size_t)n * 8, st), tn((size_t)n * 8, st), wk((size_t)130 * n * 8, st);
    HIPCK(hipMemsetAsync(y.p, 0, n * 8, st));
    PrecOp op{c, a.as_d(), m, n, a.ld, V.as<double>(), n, false, tn.as<double>(), wk.as<double>()};
    SolveOut so;
    if (o.solver == RANDNLA_LSQR)
      so = run_lsqr(c, op, bb.as_d(), 0.0, epsilon, epsilon, 1e8, l, false, false, y.as<double>(),
                    nullptr, nullptr);
    else
      so = run_cgls(c, op, bb.as_d(), epsilon, l, y.as<double>());
    HIPCK(gemv_n(V.as<double>(), n, n, n, y.as<double>(), xo.as<double>(), 1.0, 0.0, st));
    xo.commit();
    sync(c);
    fill_info(info, so);
  });
}

// src/sketch_and_solve.rs:24-33.  d = m / 4 sketch rows; [S A | S b] in one
// buffer, so the unpivoted Householder QR of S A turns the last column into
// Q^T S b; x = R^{-1} (Q^T S b) with solve_upper_triangular_system's
// zero-diagonal rule.  For d < n the reference's R is d x n and its x has d
// entries: the leading d x d block is solved and x[d:] is 0.
int randnla_sketched_least_squares_qr(randnla_ctx* c, const double* A, int64_t m, int64_t n,
                                      int64_t lda, const double* b, const randnla_opts* opts,
                                      double* x) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    const int64_t d = m / 4;
    if (d <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "the sketch needs m / 4 >= 1 rows");
    const randnla_opts o = opts_or_default(opts);
    const Stream s = to_stream(opts ? &o.rng : nullptr, GEN_THREEFRY2X64, 0);
    hipStream_t st = c->stream;
    MatIn a(A, m, n, lda, 8, st), bb(b, m, 1, m, 8, st);
    MatOut xo(x, n, 1, n, 8, st);
    DBuf SA((size_t)d * (n + 1) * 8, st);
    apply_sketch_ab(c, o, s, d, a.as_d(), m, n, a.ld, bb.as_d(), SA.as<double>(), d);
    DBuf tau = householder(c, SA.as<double>(), d, n, false, nullptr, 1);
    const int64_t k = std::min(d, n);
    DBuf w((size_t)k * 8, st);
    if (k < n) HIPCK(fill_f64(xo.as<double>(), n, 1, n, 0.0, st));
    HIPCK(backsub(SA.as<double>(), k, d, SA.as<double>() + n * d, xo.as<double>(), w.as<double>(),
                  st));
    xo.commit();
    sync(c);
  });
}

// src/sketch_and_solve.rs:53-65.  Thin SVD of S A (of (S A)^T when d < n,
// so the small SVD always sees a tall matrix): x = V diag(1 / sigma) U^T S b
// with solve_diagonal_system's rule (sigma == 0 -> 0).
int randnla_sketched_least_squares_svd(randnla_ctx* c, const double* A, int64_t m, int64_t n,
                                       int64_t lda, const double* b, const randnla_opts* opts,
                                       double* x) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    const int64_t d = m / 4;
    if (d <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "the sketch needs m / 4 >= 1 rows");
    const randnla_opts o = opts_or_default(opts);
    const Stream s = to_stream(opts ? &o.rng : nullptr, GEN_THREEFRY2X64, 0);
    hipStream_t st = c->stream;
    MatIn a(A, m, n, lda, 8, st), bb(b, m, 1, m, 8, st);
    MatOut xo(x, n, 1, n, 8, st);
    DBuf SAb((size_t)d * (n + 1) * 8, st);
    apply_sketch_ab(c, o, s, d, a.as_d(), m, n, a.ld, bb.as_d(), SAb.as<double>(), d);
    double* SA = SAb.as<double>();
    const double* Sb = SAb.as<double>() + n * d;
    const int64_t k = std::min(d, n);
    // S A = L diag(sig) R^T with L (d x k), R (n x k) orthonormal
    DBuf sig((size_t)k * 8, st), isg((size_t)k * 8, st), Lm((size_t)d * k * 8, st),
        Rm((size_t)n * k * 8, st);
    if (d >= n) {
      small_svd(c, SA, d, n, Lm.as<double>(), sig.as<double>(), Rm.as<double>());
    } else {
      DBuf SAt((size_t)n * d * 8, st);
      HIPCK(transpose_f64(SA, d, n, d, SAt.as<double>(), n, st));
      small_svd(c, SAt.as<double>(), n, d, Rm.as<double>(), sig.as<double>(), Lm.as<double>());
    }
    HIPCK(inv_sigma(sig.as<double>(), k, 0.0, 0.0, 0, isg.as<double>(), nullptr, st));
    HIPCK(scale_columns(Rm.as<double>(), n, k, n, isg.as<double>(), 1, st));   // V Sigma^+
    DBuf t((size_t)k * 8, st), wk((size_t)(64 * k + 64) * 8, st);
    HIPCK(gemv_t(Lm.as<double>(), d, k, d, Sb, t.as<double>(), 1.0, 0.0, wk.as<double>(),
                 st));                                                           // U^T S b
    HIPCK(gemv_n(Rm.as<double>(), n, k, n, t.as<double>(), xo.as<double>(), 1.0, 0.0, st));
    xo.commit();
    sync(c);
  });
}

// src/sketch_and_precondition.rs:150-216
int randnla_saddle(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                   const double* b, const double* cvec, double mu, double epsilon, int64_t l,
                   double sf, const randnla_opts* opts, double* x, double* y,
                   randnla_solve_info* info) {
  return guard(c, [&] {
    check_ls(m, n, epsilon, l, sf);
    const randnla_opts o = opts_or_default(opts);
    const Stream s = to_stream(opts ? &o.rng : nullptr, GEN_THREEFRY2X64, 0);
    hipStream_t st = c->stream;
    int64_t d = (int64_t)floor(sf * (double)n);
    d = std::min<int64_t>(std::max<int64_t>(d, 1), m);
    if (d < n) fail(RANDNLA_INVALID_PARAMETERS, "sketch has fewer rows than columns");
    MatIn a(A, m, n, lda, 8, st), bb(b, m, 1, m, 8, st);
    MatIn cc(cvec, cvec ? n : 0, 1, n, 8, st);
    MatOut xo(x, n, 1, n, 8, st), yo(y, m, 1, m, 8, st);
    DBuf SA((size_t)d * n * 8, st);
    apply_sketch(c, o, s, d, a.as_d(), m, n, a.ld, SA.as<double>(), d);
    DBuf U((size_t)d * n * 8, st), sig((size_t)n * 8, st), V((size_t)n * n * 8, st);
    small_svd(c, SA.as<double>(), d, n, U.as<double>(), sig.as<double>(), V.as<double>());
    // Preconditioner M (n x k): V diag(1/sqrt(s^2+mu)), or V[:, :k] diag(1/s)
    // over the leading s > 1e-10 (src/sketch_and_precondition.rs:185-190).
    DBuf isg((size_t)n * 8, st), kcnt(8, st);
    int64_t k = n;
    if (mu > 0.0) {
      HIPCK(inv_sigma(sig.as<double>(), n, mu, 0.0, 1, isg.as<double>(), nullptr, st));
    } else {
      HIPCK(inv_sigma(sig.as<double>(), n, 0.0, 1e-10, 2, isg.as<double>(), kcnt.as<int>(), st));
      int hk = 0;
      d2h(st, &hk, kcnt.p, sizeof(int));
      sync(c);
      k = hk;
    }
    DBuf M((size_t)n * n * 8, st);
    HIPCK(hipMemcpyAsync(M.p, V.p, (size_t)n * n * 8, hipMemcpyDeviceToDevice, st));
    HIPCK(scale_columns(M.as<double>(), n, n, n, isg.as<double>(), 1, st));
    // b_mod = b - S^T (U (diag * (V^T c)))  (src/sketch_and_precondition.rs:194-205)
    DBuf bmod((size_t)m * 8, st), tn((size_t)n * 8, st), wk((size_t)130 * n * 8 + 256 * 8, st);
    HIPCK(hipMemcpyAsync(bmod.p, bb.p, m * 8, hipMemcpyDeviceToDevice, st));
    if (cvec && n > 0) {
      DBuf bs((size_t)n * 8, st), dg((size_t)n * 8, st), ud((size_t)d * 8, st), stv((size_t)m * 8, st);
      HIPCK(gemv_t(V.as<double>(), n, n, n, cc.as_d(), bs.as<double>(), 1.0, 0.0, wk.as<double>(), st));
      HIPCK(inv_sigma(sig.as<double>(), n, mu, 0.0, mu > 0.0 ? 1 : 0, dg.as<double>(), nullptr, st));
      HIPCK(scale_rows(bs.as<double>(), n, 1, n, dg.as<double>(), st));
      HIPCK(gemv_n(U.as<double>(), d, n, d, bs.as<double>(), ud.as<double>(), 1.0, 0.0, st));
      if (o.sketch == RANDNLA_SPARSE_SIGN)
        HIPCK(sparse_sign_transpose(s, d, o.zeta, m, ud.as<double>(), stv.as<double>(), st));
      else
        HIPCK(dense_sketch_transpose(o.sketch, s, d, m, ud.as<double>(), stv.as<double>(), st));
      HIPCK(axpby(m, -1.0, stv.as<double>(), 1.0, bmod.as<double>(), st));
    }
    // z0 = U^T (S b_mod), truncated to the k columns of M.
    DBuf sb((size_t)d * 8, st), z((size_t)n * 8, st);
    apply_sketch(c, o, s, d, bmod.as<double>(), m, 1, m, sb.as<double>(), d);
    HIPCK(gemv_t(U.as<double>(), d, n, d, sb.as<double>(), z.as<double>(), 1.0, 0.0, wk.as<double>(), st));
    PrecOp op{c, a.as_d(), m, n, a.ld, M.as<double>(), k, false, tn.as<double>(), wk.as<double>()};
    SolveOut so;
    if (k > 0) {
      if (o.solver == RANDNLA_LSQR)
        so = run_lsqr(c, op, bmod.as<double>(), 0.0, epsilon, epsilon, 1e8, l, false, true,
                      z.as<double>(), nullptr, nullptr);
      else
        so = run_cgls(c, op, bmod.as<double>(), epsilon, l, z.as<double>());
    }
    // x = M z ; y = b - A x
    HIPCK(gemv_n(M.as<double>(), n, k, n, z.as<double>(), xo.as<double>(), 1.0, 0.0, st));
    HIPCK(hipMemcpyAsync(yo.p, bb.p, m * 8, hipMemcpyDeviceToDevice, st));
    HIPCK(gemv_n(a.as_d(), m, n, a.ld, xo.as<double>(), yo.as<double>(), -1.0, 1.0, st));
    xo.commit();
    yo.commit();
    sync(c);
    fill_info(info, so);
  });
}

// src/cg.rs:18-61
int randnla_cgls(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                 const double* b, double tolerance, int64_t num_iterations, const double* x0,
                 double* x, randnla_solve_info* info) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    hipStream_t st = c->stream;
    MatIn a(A, m, n, lda, 8, st), bb(b, m, 1, m, 8, st);
    MatOut xo(x, n, 1, n, 8, st);
    if (x0) {
      MatIn xi(x0, n, 1, n, 8, st);
      HIPCK(hipMemcpyAsync(xo.p, xi.p, n * 8, hipMemcpyDeviceToDevice, st));
    } else {
      HIPCK(hipMemsetAsync(xo.p, 0, n * 8, st));
    }
    DBuf tn((size_t)n * 8, st), wk((size_t)130 * n * 8, st);
    PrecOp op{c, a.as_d(), m, n, a.ld, nullptr, n, false, tn.as<double>(), wk.as<double>()};
    const SolveOut so = run_cgls(c, op, bb.as_d(), tolerance, num_iterations, xo.as<double>());
    xo.commit();
    sync(c);
    fill_info(info, so);
  });
}

// src/solvers.rs:115-278
int randnla_lsqr(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda,
                 const double* b, double damp, double atol, double btol, double conlim,
                 int64_t iter_lim, int calc_var, const double* x0, double* x, double* var,
                 double* arnorms, randnla_solve_info* info) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    hipStream_t st = c->stream;
    MatIn a(A, m, n, lda, 8, st), bb(b, m, 1, m, 8, st);
    MatOut xo(x, n, 1, n, 8, st), vo(var, n, 1, n, 8, st);
    if (x0) {
      MatIn xi(x0, n, 1, n, 8, st);
      HIPCK(hipMemcpyAsync(xo.p, xi.p, n * 8, hipMemcpyDeviceToDevice, st));
    }
    DBuf vtmp;
    double* vp = vo.as<double>();
    if (!vp) { vtmp = DBuf((size_t)n * 8, st); vp = vtmp.as<double>(); }
    DBuf tn((size_t)n * 8, st), wk((size_t)130 * n * 8, st);
    PrecOp op{c, a.as_d(), m, n, a.ld, nullptr, n, false, tn.as<double>(), wk.as<double>()};
    const int64_t lim = iter_lim > 0 ? iter_lim : 2 * n;
    std::vector<double> arn((size_t)lim + 1, 0.0);
    const SolveOut so = run_lsqr(c, op, bb.as_d(), damp, atol, btol, conlim, lim, calc_var != 0,
                                 x0 != nullptr, xo.as<double>(), vp, arn.data());
    if (arnorms) {
      const int cnt = so.itn > 0 ? so.itn : 1;
      HIPCK(hipMemcpy(arnorms, arn.data(), (size_t)cnt * 8, hipMemcpyDefault));
    }
    xo.commit();
    if (var) vo.commit();
    sync(c);
    fill_info(info, so);
  });
}

// src/solvers.rs:22-41
int randnla_solve_upper_triangular(randnla_ctx* c, const double* U, int64_t n, int64_t ldu,
                                   const double* y, double* x) {
  return guard(c, [&] {
    if (n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    hipStream_t st = c->stream;
    MatIn u(U, n, n, ldu, 8, st), yy(y, n, 1, n, 8, st);
    MatOut xo(x, n, 1, n, 8, st);
    DBuf w((size_t)n * 8, st);
    HIPCK(backsub(u.as_d(), n, u.ld, yy.as_d(), xo.as<double>(), w.as<double>(), st));
    xo.commit();
    sync(c);
  });
}

int randnla_solve_upper_triangular_f32(randnla_ctx* c, const float* U, int64_t n, int64_t ldu,
                                       const float* y, float* x) {
  return guard(c, [&] {
    if (n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    hipStream_t st = c->stream;
    MatIn u(U, n, n, ldu, 4, st), yy(y, n, 1, n, 4, st);
    MatOut xo(x, n, 1, n, 4, st);
    DBuf w((size_t)n * 4, st);
    HIPCK(backsub_f32((const float*)u.p, n, u.ld, (const float*)yy.p, xo.as<float>(), w.as<float>(),
                      st));
    xo.commit();
    sync(c);
  });
}

// src/pivot_decompositions.rs:106-175
int randnla_qrcp(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda, double* Q,
                 int64_t ldq, double* R, int64_t ldr, int64_t* perm) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    hipStream_t st = c->stream;
    MatIn a(A, m, n, lda, 8, st);
    DBuf W((size_t)m * n * 8, st), P((size_t)n * 8, st);
    HIPCK(hipMemcpy2DAsync(W.p, m * 8, a.p, a.ld * 8, m * 8, n, hipMemcpyDeviceToDevice, st));
    DBuf tau = householder(c, W.as<double>(), m, n, true, P.as<int64_t>());
    MatOut ro(R, m, n, ldr, 8, st);
    HIPCK(upper_copy(W.as<double>(), m, n, m, ro.as<double>(), ro.ld, st));
    if (Q) {
      MatOut qo(Q, m, m, ldq, 8, st);
      HIPCK(householder_form_q(nullptr, m, n, m, tau.as<double>(), qo.as<double>(), qo.ld, m, st));
      qo.commit();
    }
    ro.commit();
    put_indices(c, perm, P.as<int64_t>(), n);
    sync(c);
  });
}

// src/cqrrpt.rs:27-58
int randnla_cqrrpt(randnla_ctx* c, const double* A, int64_t m, int64_t n, int64_t lda, int64_t d,
                   const randnla_opts* opts, double* Q, int64_t ldq, double* R, int64_t ldr,
                   int64_t* J, int64_t* k_out) {
  return guard(c, [&] {
    if (!(n <= d && d <= m)) fail(RANDNLA_PANIC, "d must satisfy n ≤ d ≪ m");
    if (n <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    const randnla_opts o = opts_or_default(opts);
    const Stream s = to_stream(opts ? &o.rng : nullptr, GEN_THREEFRY2X64, 0);
    hipStream_t st = c->stream;
    MatIn a(A, m, n, lda, 8, st);
    DBuf SA((size_t)d * n * 8, st), P((size_t)n * 8, st);
    apply_sketch(c, o, s, d, a.as_d(), m, n, a.ld, SA.as<double>(), d);
    DBuf tau = householder(c, SA.as<double>(), d, n, true, P.as<int64_t>());
    // Numerical rank: |R_ii| > 1e-10 (src/cqrrpt.rs:37-43).
    DBuf kc(8, st);
    HIPCK(count_diag_above(SA.as<double>(), std::min(d, n), d, 1e-10, kc.as<int>(), st));
    int k = 0;
    d2h(st, &k, kc.p, sizeof(int));
    sync(c);
    if (k < n) fail(RANDNLA_PANIC, "Matrix dimension mismatch: sketch rank k < n");
    DBuf Rsk((size_t)n * n * 8, st), Rinv((size_t)n * n * 8, st), AJ((size_t)m * n * 8, st),
        Apre((size_t)m * n * 8, st);
    HIPCK(upper_copy(SA.as<double>(), n, n, d, Rsk.as<double>(), n, st));
    HIPCK(trinv_upper(Rsk.as<double>(), n, n, Rinv.as<double>(), n, st));
    HIPCK(gather_cols(a.as_d(), m, a.ld, P.as<int64_t>(), n, AJ.as<double>(), m, st));
    gemm64(c, m, n, n, AJ.p, T::F64, m, false, Rinv.p, T::F64, n, Apre.p, T::F64, m, 1.0, "trmm");
    // R_pre = chol(A_pre^T A_pre)^T; failure -> the reference's expect() panic.
    DBuf G((size_t)n * n * 8, st), flags((size_t)(n + 1) * 4, st), X((size_t)n * n * 8, st);
    gram64(c, Apre.p, T::F64, m, n, m, G.as<double>());
    HIPCK(cholesky_upper(G.as<double>(), n, n, 0.0, flags.as<int>(), flags.as<int>() + n, st));
    int ndef = 0;
    d2h(st, &ndef, flags.as<int>() + n, sizeof(int));
    sync(c);
    if (ndef) fail(RANDNLA_PANIC, "Cholesky decomposition failed");
    HIPCK(trinv_upper(G.as<double>(), n, n, X.as<double>(), n, st));
    MatOut qo(Q, m, n, ldq, 8, st), ro(R, n, n, ldr, 8, st);
    gemm64(c, m, n, n, Apre.p, T::F64, m, false, X.p, T::F64, n, qo.p, T::F64, qo.ld, 1.0, "trmm");
    gemm64(c, n, n, n, G.p, T::F64, n, false, Rsk.p, T::F64, n, ro.p, T::F64, ro.ld, 1.0, "rprod");
    qo.commit();
    ro.commit();
    put_indices(c, J, P.as<int64_t>(), n);
    sync(c);
    if (k_out) *k_out = k;
  });
}

// src/test_assist.rs (generators for the configurations)
int randnla_test_matrix(randnla_ctx* c, int64_t m, int64_t n, const double* sigma, double noise,
                        const randnla_rng* rng, void* out, int dtype, int64_t ld) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0 || (m & (m - 1)) || (n & (n - 1)) || n > m)
      fail(RANDNLA_INVALID_DIMENSIONS, "m, n must be powers of two with n <= m");
    hipStream_t st = c->stream;
    const Stream s = to_stream(rng, GEN_PHILOX4X32, 7);
    MatIn sg(sigma, n, 1, n, 8, st);
    const T to = dtype == RANDNLA_F32 ? T::F32 : T::F64;
    MatOut o(out, m, n, ld, to == T::F32 ? 4 : 8, st);
    DBuf work((size_t)m * n * 8, st);
    HIPCK(hadamard_test_matrix(m, n, sg.as_d(), noise, s, work.as<double>(), o.p, to, o.ld, st));
    o.commit();
    sync(c);
  });
}

// C (M x N, FP64) = op(A) * B with the library's MFMA kernels.
// precision: RANDNLA_FP64 (FP64 MFMA; A/B FP64 or FP32), RANDNLA_BF16 (A, B
// rounded to BF16, FP32 accumulate), 3 = split FP16 (FP32 operands, three
// FP16 products), 4 = op(A) split FP16, B rounded to FP16 (two products).  The low-precision kernels need K, lda, ldb multiples of 8
// (and M for op(A) = A); operands are staged to meet that.
int randnla_gemm(randnla_ctx* c, int precision, int trans_a, int64_t M, int64_t N, int64_t K,
                 const void* A, int dtype_a, int64_t lda, const void* B, int dtype_b,
                 int64_t ldb, double* Cm, int64_t ldc) {
  return guard(c, [&] {
    if (M <= 0 || N <= 0 || K <= 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    hipStream_t st = c->stream;
    const T ta = dtype_a == RANDNLA_F32 ? T::F32 : T::F64;
    const T tb = dtype_b == RANDNLA_F32 ? T::F32 : T::F64;
    const int64_t ar = trans_a ? K : M, ac = trans_a ? M : K;
    MatIn a(A, ar, ac, lda, ta == T::F32 ? 4 : 8, st), b(B, K, N, ldb, tb == T::F32 ? 4 : 8, st);
    MatOut o(Cm, M, N, ldc, 8, st);
    if (precision == RANDNLA_FP64) {
      gemm64(c, M, N, K, a.p, ta, a.ld, trans_a != 0, b.p, tb, b.ld, o.p, T::F64, o.ld);
    } else if (precision == RANDNLA_BF16 || (precision >= 3 && precision <= 6)) {
      const bool f16x3 = precision != RANDNLA_BF16;
      const int64_t Kp = r8(K), Mp = r8(M);
      const int64_t arp = trans_a ? Kp : Mp, acp = trans_a ? M : Kp;
      // Stage A and B zero-padded: BF16 for mode 0, FP32 for the split mode.
      const T tw = f16x3 ? T::F32 : T::BF16;
      const size_t ew = f16x3 ? 4 : 2;
      DBuf Aw((size_t)arp * acp * ew, st), Bw((size_t)Kp * N * ew, st);
      HIPCK(hipMemsetAsync(Aw.p, 0, (size_t)arp * acp * ew, st));
      HIPCK(hipMemsetAsync(Bw.p, 0, (size_t)Kp * N * ew, st));
      HIPCK(convert(a.p, ta, ar, ac, a.ld, Aw.p, tw, arp, 1.0, st));
      HIPCK(convert(b.p, tb, K, N, b.ld, Bw.p, tw, Kp, 1.0, st));
      float sa = 1.f, sb = 1.f;
      if (f16x3) {
        DBuf am(16, st);
        HIPCK(absmax(Aw.p, T::F32, arp, acp, arp, am.as<double>(), st));
        HIPCK(absmax(Bw.p, T::F32, Kp, N, Kp, am.as<double>() + 1, st));
        double h[2];
        d2h(st, h, am.p, 16);
        sync(c);
        sa = pow2_scale(h[0]);
        sb = pow2_scale(h[1]);
      }
      const int64_t Mo = trans_a ? M : Mp;
      const int64_t tiles = ((Mo + 255) / 256) * ((N + 319) / 320);
      const int splits = choose_splits(tiles, Kp, 2048);
      DBuf part((size_t)splits * Mo * N * 4, st);
      LpGemmArgs g;
      g.mode = precision == 6 ? 4 : precision == 5 ? 3 : precision == 4 ? 2 : (f16x3 ? 1 : 0);
      g.M = Mo; g.N = N; g.K = Kp; g.A = Aw.p; g.lda = arp;
      g.a_trans = trans_a != 0; g.B = Bw.p; g.ldb = Kp; g.C = part.as<float>(); g.ldc = Mo;
      g.cstride = Mo * N; g.splits = splits; g.sa = sa; g.sb = sb; g.alpha = 1.f / (sa * sb);
      DBuf Bh;
      if (f16x3) {              // full-tile products load B's FP16 form straight into LDS
        Bh = DBuf((size_t)2 * N * Kp * 2, st);
        g.bh = Bh.as<uint16_t>(); g.ldbh = Kp;
      }
      {
        Tm tm(c, f16x3 ? "gemm_f16x3" : "gemm_bf16");
        HIPCK(gemm_lp(g, st));
      }
      HIPCK(reduce_partials(part.as<float>(), M, N, Mo, Mo * N, splits, o.p, T::F64, o.ld, 1.0, st));
    } else if (precision == RANDNLA_FP8) {
      // FP32 copies, FP8 quantisation with 128 x 128 block scales for op(A)
      // and per-column 128-K block scales for B, then the scaled-MFMA GEMM.
      DBuf A32((size_t)ar * ac * 4, st), B32((size_t)K * N * 4, st);
      HIPCK(convert(a.p, ta, ar, ac, a.ld, A32.p, T::F32, ar, 1.0, st));
      HIPCK(convert(b.p, tb, K, N, b.ld, B32.p, T::F32, K, 1.0, st));
      const int64_t arp = (ar + 127) / 128 * 128, acp = (ac + 127) / 128 * 128;
      const int64_t Kq = (K + 127) / 128 * 128, kb = Kq / 128;
      DBuf Arm((size_t)arp * acp, st), Acm((size_t)acp * arp, st),
          SA((size_t)(arp / 128) * (acp / 128), st), B8((size_t)Kq * N, st), SB((size_t)N * kb, st);
      fp8_operands(c, A32.as<float>(), ar, ac, ar, Arm.as<uint8_t>(), Acm.as<uint8_t>(),
                   SA.as<uint8_t>(), B32.as<float>(), K, N, K, B8.as<uint8_t>(), SB.as<uint8_t>());
      const int64_t tiles = ((M + 127) / 128) * ((N + 319) / 320);
      const int splits = choose_splits(tiles, Kq, 2048);
      DBuf part((size_t)splits * M * N * 4, st);
      Fp8GemmArgs g;
      g.M = M; g.N = N; g.K = Kq;
      if (!trans_a) {           // rows of A, K along its columns: row-major copy
        g.A = Arm.as<uint8_t>(); g.lda = acp; g.sa_r = acp / 128; g.sa_k = 1;
      } else {                  // columns of A, K along its rows: column-major copy
        g.A = Acm.as<uint8_t>(); g.lda = arp; g.sa_r = 1; g.sa_k = acp / 128;
      }
      g.SA = SA.as<uint8_t>();
      g.B = B8.as<uint8_t>(); g.ldb = Kq; g.SB = SB.as<uint8_t>(); g.sb_j = kb;
      g.C = part.as<float>(); g.ldc = M; g.cstride = M * N; g.splits = splits;
      {
        Tm tm(c, "gemm_fp8");
        HIPCK(gemm_fp8(g, st));
      }
      HIPCK(reduce_partials(part.as<float>(), M, N, M, M * N, splits, o.p, T::F64, o.ld, 1.0, st));
    } else {
      fail(RANDNLA_INVALID_PARAMETERS, "unknown precision");
    }
    o.commit();
    sync(c);
  });
}

// ||A - U diag(S) Vt||_2 and ||A||_2 by Golub-Kahan-Lanczos
// bidiagonalisation of R (R = A - U diag(S) Vt, or A) from a fixed Gaussian
// start, with full reorthogonalisation of both bases (two classical
// Gram-Schmidt passes per vector), `iters` steps (at most min(m, n, 512)).
// The estimate is the largest singular value of the steps x steps upper
// bidiagonal matrix (one-sided Jacobi on the device): a lower bound on
// ||R||_2 that converges from below, much faster than power iteration when
// the top of R's spectrum is flat (a noise floor).
int randnla_spectral_error(randnla_ctx* c, const void* A, int dtype, int64_t m, int64_t n,
                           int64_t lda, const double* U, int64_t ldu, const double* S, int64_t r,
                           const double* Vt, int64_t ldvt, int iters, double* err,
                           double* anorm) {
  return guard(c, [&] {
    if (m <= 0 || n <= 0 || r < 0) fail(RANDNLA_INVALID_DIMENSIONS, "dimensions must be positive");
    if (iters <= 0) fail(RANDNLA_INVALID_PARAMETERS, "iters must be positive");
    hipStream_t st = c->stream;
    const T ta = dtype == RANDNLA_F32 ? T::F32 : T::F64;
    MatIn a(A, m, n, lda, ta == T::F32 ? 4 : 8, st);
    MatIn u(U, m, r, ldu, 8, st), sv(S, r, 1, r, 8, st), vt(Vt, r, n, ldvt, 8, st);
    const int64_t t = std::min<int64_t>(iters, std::min<int64_t>(512, std::min(m, n)));
    DBuf Vb((size_t)n * (t + 1) * 8, st), Ub((size_t)m * t * 8, st), p((size_t)m * 8, st),
        q((size_t)n * 8, st), T1((size_t)std::max<int64_t>(r, 1) * 8, st),
        coef((size_t)(t + 1) * 8, st), work(256 * 8, st), nrm(8, st);
    auto mv = [&](bool resid, const double* x, double* y) {        // y = R x
      gemm64(c, m, 1, n, a.p, ta, a.ld, false, x, T::F64, n, y, T::F64, m);
      if (resid && r > 0) {
        gemm64(c, r, 1, n, vt.p, T::F64, vt.ld, false, x, T::F64, n, T1.p, T::F64, r);
        HIPCK(scale_rows(T1.as<double>(), r, 1, r, sv.as_d(), st));
        GemmArgs g;
        g.M = m; g.N = 1; g.K = r; g.A = u.p; g.ta = T::F64; g.lda = u.ld; g.a_trans = false;
        g.B = T1.p; g.tb = T::F64; g.ldb = r; g.C = y; g.tc = T::F64; g.ldc = m;
        g.alpha = -1.0; g.beta = 1.0;
        HIPCK(gemm_f64(g, st));
      }
    };
    auto mtv = [&](bool resid, const double* y, double* x) {       // x = R^T y
      gemm64(c, n, 1, m, a.p, ta, a.ld, true, y, T::F64, m, x, T::F64, n);
      if (resid && r > 0) {
        gemm64(c, r, 1, m, u.p, T::F64, u.ld, true, y, T::F64, m, T1.p, T::F64, r);
        HIPCK(scale_rows(T1.as<double>(), r, 1, r, sv.as_d(), st));
        GemmArgs g;
        g.M = n; g.N = 1; g.K = r; g.A = vt.p; g.ta = T::F64; g.lda = vt.ld; g.a_trans = true;
        g.B = T1.p; g.tb = T::F64; g.ldb = r; g.C = x; g.tc = T::F64; g.ldc = n;
        g.alpha = -1.0; g.beta = 1.0;
        HIPCK(gemm_f64(g, st));
      }
    };
    auto reorth = [&](double* x, int64_t len, const double* basis, int64_t j) {
      for (int pass = 0; pass < 2 && j > 0; ++pass) {   // x -= B (B^T x)
        gemm64(c, j, 1, len, basis, T::F64, len, true, x, T::F64, len, coef.p, T::F64, j);
        GemmArgs g;
        g.M = len; g.N = 1; g.K = j; g.A = basis; g.ta = T::F64; g.lda = len; g.a_trans = false;
        g.B = coef.p; g.tb = T::F64; g.ldb = j; g.C = x; g.tc = T::F64; g.ldc = len;
        g.alpha = -1.0; g.beta = 1.0;
        HIPCK(gemm_f64(g, st));
      }
    };
    auto norm = [&](const double* x, int64_t len) {
      HIPCK(dot(x, x, len, nrm.as<double>(), work.as<double>(), st));
      double h = 0.0;
      d2h(st, &h, nrm.p, 8);
      return std::sqrt(h);
    };
    auto run = [&](bool resid) -> double {
      double* V = Vb.as<double>();
      double* Uq = Ub.as<double>();
      const Stream s0 = stream_from_seed(GEN_PHILOX4X32, 12345);
      HIPCK(sketch_panel(0, s0, n, 1, V, T::F64, n, 1.0, st));
      const double v0 = norm(V, n);
      HIPCK(axpby(n, 0.0, V, 1.0 / v0, V, st));
      std::vector<double> al, be;
      for (int64_t j = 0; j < t; ++j) {
        double* vj = V + j * n;
        double* uj = Uq + j * m;
        mv(resid, vj, p.as<double>());
        if (j > 0) HIPCK(axpby(m, -be[j - 1], Uq + (j - 1) * m, 1.0, p.as<double>(), st));
        reorth(p.as<double>(), m, Uq, j);
        const double aj = norm(p.as<double>(), m);
        if (!(aj > 0.0)) break;
        al.push_back(aj);
        HIPCK(hipMemcpyAsync(uj, p.p, (size_t)m * 8, hipMemcpyDeviceToDevice, st));
        HIPCK(axpby(m, 0.0, p.as<double>(), 1.0 / aj, uj, st));
        mtv(resid, uj, q.as<double>());
        HIPCK(axpby(n, -aj, vj, 1.0, q.as<double>(), st));
        reorth(q.as<double>(), n, V, j + 1);
        const double bj = j + 1 < t ? norm(q.as<double>(), n) : 0.0;
        if (!(bj > 0.0)) break;
        be.push_back(bj);
        HIPCK(hipMemcpyAsync(V + (j + 1) * n, q.p, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
        HIPCK(axpby(n, 0.0, q.as<double>(), 1.0 / bj, V + (j + 1) * n, st));
      }
      const int64_t te = (int64_t)al.size();
      if (te == 0) return 0.0;
      std::vector<double> hb((size_t)te * te, 0.0);
      for (int64_t i = 0; i < te; ++i) {
        hb[i + i * te] = al[i];
        if (i + 1 < te) hb[i + (i + 1) * te] = be[i];
      }
      DBuf W((size_t)te * te * 8, st), Vj((size_t)te * te * 8, st), cn((size_t)te * 8, st),
          cnt(8, st);
      HIPCK(hipMemcpyAsync(W.p, hb.data(), (size_t)te * te * 8, hipMemcpyHostToDevice, st));
      HIPCK(set_identity(Vj.as<double>(), te, te, te, st));
      int sweeps = 0;
      HIPCK(jacobi_svd(W.as<double>(), te, te, Vj.as<double>(), te, cn.as<double>(),
                       cnt.as<int>(), 60, jacobi_tol(te), st, &sweeps));
      HIPCK(column_norms(W.as<double>(), te, te, te, cn.as<double>(), st));
      std::vector<double> h(te);
      d2h(st, h.data(), cn.p, te * 8);
      sync(c);
      return *std::max_element(h.begin(), h.end());
    };
    if (err) *err = run(true);
    if (anorm) *anorm = run(false);
  });
}

// ------------------------------------------------ batched least squares
// No reference counterpart (the reference solves one problem per call,
// src/sketch_and_precondition.rs:26-59); each problem follows that function
// with the LSQR solver.
static_assert(sizeof(randnla_ls_info) == 56, "randnla_ls_info layout");
static_assert(sizeof(randnla_ls_problem) == 56, "randnla_ls_problem layout");

}  // extern "C"

struct randnla_batch_ls {
  randnla_ctx* ctx = nullptr;
  rna::BatchPlan* plan = nullptr;
  bool autotune_pending = false;
};

extern "C" {

int randnla_batch_ls_create(randnla_ctx* c, const randnla_ls_problem* problems, int64_t count,
                            double sf, double tol, int64_t iter_lim, const randnla_opts* opts,
                            int autotune, randnla_batch_ls** out) {
  return guard(c, [&] {
    if (!out) fail(RANDNLA_INVALID_PARAMETERS, "plan output pointer is null");
    *out = nullptr;
    if (count <= 0 || !problems)
      fail(RANDNLA_INVALID_DIMENSIONS, "a batch needs at least one problem");
    if (count > (1 << 24)) fail(RANDNLA_INVALID_DIMENSIONS, "at most 2^24 problems per batch");
    std::vector<BatchProbDesc> desc(count);
    char buf[200];
    for (int64_t i = 0; i < count; ++i) {
      const randnla_ls_problem& q = problems[i];
      if (q.m < 0 || q.n < 0) fail(RANDNLA_NEGATIVE_DIMENSIONS, "negative dimension");
      check_ls(q.m, q.n, tol, iter_lim, sf);
      if (q.n > 256) {
        snprintf(buf, sizeof buf, "batched problems need n <= 256, problem %lld has n = %lld",
                 (long long)i, (long long)q.n);
        fail(RANDNLA_INVALID_DIMENSIONS, buf);
      }
      if (q.m > (1ll << 30)) fail(RANDNLA_INVALID_DIMENSIONS, "batched problems need m < 2^30");
      if (q.lda < q.m) {
        snprintf(buf, sizeof buf, "problem %lld: lda %lld < m %lld", (long long)i,
                 (long long)q.lda, (long long)q.m);
        fail(RANDNLA_INVALID_DIMENSIONS, buf);
      }
      if (q.a_offset < 0 || q.b_offset < 0 || q.x_offset < 0)
        fail(RANDNLA_INVALID_PARAMETERS, "negative offset");
      if ((q.n + 256) * q.lda >= (1ll << 29)) {
        snprintf(buf, sizeof buf, "problem %lld: (n + 256) * lda must stay below 2^29",
                 (long long)i);
        fail(RANDNLA_INVALID_DIMENSIONS, buf);
      }
      desc[i] = BatchProbDesc{q.m, q.n, q.lda, q.a_offset, q.b_offset, q.x_offset, q.stream_id};
    }
    const Stream s = to_stream(opts ? &opts->rng : nullptr, GEN_PHILOX4X32, 42);
    if (s.gen != GEN_PHILOX4X32)
      fail(RANDNLA_INVALID_PARAMETERS, "batched sketches use the Philox4x32 generator");
    BatchPlan* pl = nullptr;
    HIPCK(batch_plan_create(desc.data(), count, sf, tol, iter_lim, s, c->stream, &pl));
    randnla_batch_ls* b = new randnla_batch_ls();
    b->ctx = c; b->plan = pl; b->autotune_pending = autotune != 0;
    *out = b;
  });
}

int randnla_batch_ls_execute(randnla_batch_ls* b, const float* A, const double* bb, double* x,
                             randnla_ls_info* info) {
  if (!b) return RANDNLA_INVALID_PARAMETERS;
  randnla_ctx* c = b->ctx;
  return guard(c, [&] {
    if (!is_device_ptr(A) || !is_device_ptr(bb) || !is_device_ptr(x))
      fail(RANDNLA_INVALID_PARAMETERS, "batched least squares takes device pointers for A, b, x");
    hipStream_t st = c->stream;
    {
      Tm tm(c, "batch_ls");
      if (b->autotune_pending) {
        HIPCK(batch_plan_autotune(b->plan, A, bb, x, st));
        b->autotune_pending = false;
      } else {
        HIPCK(batch_plan_execute(b->plan, A, bb, x, st));
      }
    }
    // Host info arrays are filled from a device copy through a blocking copy.
    const bool info_dev = info && is_device_ptr(info);
    const size_t info_bytes = sizeof(randnla_ls_info) * (size_t)batch_plan_count(b->plan);
    DBuf dinfo;
    if (info && !info_dev) dinfo = DBuf(info_bytes, st);
    int* herr = nullptr;
    HIPCK(hipHostMalloc((void**)&herr, sizeof(int), hipHostMallocDefault));
    *herr = 0;
    hipError_t e = batch_plan_info(b->plan, info_dev ? info : (randnla_ls_info*)dinfo.p,
                                   info != nullptr, herr, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    const int werr = *herr;
    (void)hipHostFree(herr);
    HIPCK(e);
    if (info && !info_dev) d2h(st, info, dinfo.p, info_bytes);
    if (werr) fail(RANDNLA_COMPUTATION_ERROR, "a batched solve hit its internal wait limit");
  });
}

int randnla_batch_ls_sketch(randnla_batch_ls* b, float* out, int64_t out_elems) {
  if (!b) return RANDNLA_INVALID_PARAMETERS;
  return guard(b->ctx, [&] {
    if (!out || out_elems < batch_plan_sketch_size(b->plan))
      fail(RANDNLA_INVALID_DIMENSIONS, "sketch output buffer too small");
    hipStream_t st = b->ctx->stream;
    if (is_device_ptr(out)) {
      HIPCK(batch_plan_sketch(b->plan, out, st));
      sync(b->ctx);
    } else {
      const size_t bytes = (size_t)batch_plan_sketch_size(b->plan) * 4;
      DBuf tmp(bytes, st);
      HIPCK(batch_plan_sketch(b->plan, tmp.as<float>(), st));
      d2h(st, out, tmp.p, bytes);
    }
  });
}

int randnla_batch_ls_config(const randnla_batch_ls* b, int* row_step_shift, int* wgs_per_cu,
                            double* best_ms) {
  if (!b) return RANDNLA_INVALID_PARAMETERS;
  batch_plan_config(b->plan, row_step_shift, wgs_per_cu, best_ms);
  return RANDNLA_OK;
}

void randnla_batch_ls_destroy(randnla_batch_ls* b) {
  if (!b) return;
  if (b->ctx) {
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
  }
  batch_plan_destroy(b->plan);
  delete b;
}

int randnla_batch_ls_solve(randnla_ctx* c, const randnla_ls_problem* problems, int64_t count,
                     const float* A, const double* b, double* x, double sf, double tol,
                     int64_t iter_lim, const randnla_opts* opts, randnla_ls_info* info) {
  randnla_batch_ls* pl = nullptr;
  int st = randnla_batch_ls_create(c, problems, count, sf, tol, iter_lim, opts, 0, &pl);
  if (st != RANDNLA_OK) return st;
  st = randnla_batch_ls_execute(pl, A, b, x, info);
  randnla_batch_ls_destroy(pl);
  return st;
}

// ------------------------------------------- learned (k, p, q) proposer
// No reference counterpart: the reference leaves k, p and q to the caller
// (src/lora_drivers.rs:30-69 takes k and s; the power-iteration count is
// fixed by tsog1's num_passes).

int randnla_mlp_param_count(void) { return mlp_param_count(); }

int randnla_mlp_forward(randnla_ctx* c, const float* params, const float* X, int64_t B,
                        float* Y) {
  return guard(c, [&] {
    if (B <= 0 || B > (1 << 24)) fail(RANDNLA_INVALID_DIMENSIONS, "batch size out of range");
    hipStream_t st = c->stream;
    const int np = mlp_param_count();
    MatIn pp(params, np, 1, np, 4, st), x(X, 34, B, 34, 4, st);
    MatOut y(Y, 3, B, 3, 4, st);
    {
      Tm tm(c, "mlp_forward");
      HIPCK(mlp_forward((const float*)pp.p, (const float*)x.p, (int)B, (float*)y.p, st));
    }
    y.commit();
    sync(c);
  });
}

int randnla_mlp_grad(randnla_ctx* c, const float* params, const float* X, const float* Tg,
                     const float* wo, int64_t B, float* grad, float* loss) {
  return guard(c, [&] {
    if (B <= 0 || B > (1 << 24)) fail(RANDNLA_INVALID_DIMENSIONS, "batch size out of range");
    hipStream_t st = c->stream;
    const int np = mlp_param_count();
    MatIn pp(params, np, 1, np, 4, st), x(X, 34, B, 34, 4, st), t(Tg, 3, B, 3, 4, st),
        w(wo, 3, 1, 3, 4, st);
    MatOut g(grad, np, 1, np, 4, st), l(loss, 1, 1, 1, 4, st);
    const int64_t nwg = (B + mlp_group_samples() - 1) / mlp_group_samples();
    DBuf work((size_t)nwg * (np + 1) * 4, st), lbuf(4, st);
    {
      Tm tm(c, "mlp_grad");
      HIPCK(mlp_grad((const float*)pp.p, (const float*)x.p, (const float*)t.p,
                     (const float*)w.p, (int)B, (float*)g.p,
                     loss ? (float*)l.p : lbuf.as<float>(), work.as<float>(), st));
    }
    g.commit();
    l.commit();
    sync(c);
  });
}

int randnla_mlp_train_step(randnla_ctx* c, float* params, float* adam_m, float* adam_v,
                           const float* X, const float* Tg, const float* wo, int64_t B, int step,
                           float lr, float* loss) {
  return guard(c, [&] {
    if (B <= 0 || B > (1 << 24)) fail(RANDNLA_INVALID_DIMENSIONS, "batch size out of range");
    if (step < 1) fail(RANDNLA_INVALID_PARAMETERS, "Adam step counts from 1");
    if (!is_device_ptr(params) || !is_device_ptr(adam_m) || !is_device_ptr(adam_v) ||
        !is_device_ptr(X) || !is_device_ptr(Tg) || !is_device_ptr(wo))
      fail(RANDNLA_INVALID_PARAMETERS, "training takes device pointers");
    hipStream_t st = c->stream;
    const int np = mlp_param_count();
    const int64_t nwg = (B + mlp_group_samples() - 1) / mlp_group_samples();
    DBuf work((size_t)nwg * (np + 1) * 4, st), g((size_t)np * 4, st), lb(4, st);
    MatOut l(loss, 1, 1, 1, 4, st);
    {
      Tm tm(c, "mlp_train");
      HIPCK(mlp_grad(params, X, Tg, wo, (int)B, g.as<float>(), loss ? (float*)l.p : lb.as<float>(),
                     work.as<float>(), st));
      HIPCK(mlp_adam(params, adam_m, adam_v, g.as<float>(), step, lr, 0.9f, 0.999f, 1e-8f, st));
    }
    l.commit();
    if (loss) sync(c);
  });
}

// Probe features of A for the proposer: Y0 = A Omega0 (Omega0 n x 48
// Gaussian from rng, Philox4x32 seed 42 by default); s = singular values of
// Y0 (CholQR2 + Jacobi, descending); out[i] = log10(max(s_i / s_0, 1e-12)) /
// 12 for i < 32 (all -1 when s_0 = 0), out[32] = ln(m) / 10, out[33] =
// ln(n) / 10.
int randnla_probe_features(randnla_ctx* c, const void* A, int dtype, int64_t m, int64_t n,
                           int64_t lda, const randnla_rng* rng, float* out) {
  return guard(c, [&] {
    if (m < 48 || n < 48) fail(RANDNLA_INVALID_DIMENSIONS, "the probe needs m, n >= 48");
    if (dtype != RANDNLA_F64 && dtype != RANDNLA_F32)
      fail(RANDNLA_INVALID_PARAMETERS, "dtype must be F64 or F32");
    hipStream_t st = c->stream;
    const Stream s = to_stream(rng, GEN_PHILOX4X32, 42);
    const T ta = dtype == RANDNLA_F64 ? T::F64 : T::F32;
    MatIn a(A, m, n, lda, dtype == RANDNLA_F64 ? 8 : 4, st);
    const int64_t l = 48;
    DBuf Y((size_t)m * l * 8, st), sig((size_t)l * 8, st);
    {
      Tm tm(c, "probe");
      DBuf Om((size_t)n * l * 8, st);
      HIPCK(sketch_panel(0, s, n, l, Om.p, T::F64, n, 1.0, st));
      GemmArgs g;
      g.M = m; g.N = l; g.K = n; g.A = a.p; g.ta = ta; g.lda = a.ld; g.a_trans = false;
      g.B = Om.p; g.tb = T::F64; g.ldb = n; g.C = Y.p; g.tc = T::F64; g.ldc = m;
      HIPCK(gemm_f64(g, st));
      DBuf V((size_t)l * l * 8, st);
      small_svd(c, Y.as<double>(), m, l, nullptr, sig.as<double>(), V.as<double>());
    }
    std::vector<double> hs(l);
    d2h(st, hs.data(), sig.p, l * 8);
    sync(c);
    float f[34];
    for (int i = 0; i < 32; ++i) {
      const double r = hs[0] > 0.0 ? hs[i] / hs[0] : 0.0;
      f[i] = (float)(std::log10(std::max(r, 1e-12)) / 12.0);
    }
    f[32] = (float)(std::log((double)m) / 10.0);
    f[33] = (float)(std::log((double)n) / 10.0);
    if (is_device_ptr(out)) {
      HIPCK(hipMemcpyAsync(out, f, sizeof f, hipMemcpyHostToDevice, st));
      sync(c);
    } else {
      memcpy(out, f, sizeof f);
    }
  });
}

}  // extern "C"
