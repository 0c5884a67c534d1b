"""randnla_amd -- MI355X (gfx950) sketch-and-factor library, Python binding.

A thin ctypes layer over ``librandnla_amd.so`` (C ABI in include/randnla.h)
that keeps the names, argument meaning and error behaviour of the reference
crate ``randblas`` (divijkhaitan/randnla):

    sketching_operator, haar_sample          src/sketch.rs:45-130
    rand_svd                                 src/lora_drivers.rs:30-69
    QB1, RF1, tsog1, Orth, Stabilizer        src/lora_helpers.rs:17-146
    blendenpik_overdetermined, lsrn_overdetermined,
    sketch_saddle_point_precondition         src/sketch_and_precondition.rs:26-216
    cgls                                     src/cg.rs:18-61
    lsqr, solve_upper_triangular_system      src/solvers.rs:22-278
    qrcp                                     src/pivot_decompositions.rs:106-175
    sap_chol_qrcp                            src/cqrrpt.rs:27-58

Errors: a non-zero status raises ``RandNLAError`` whose ``kind`` is the
reference's enum variant name (src/errors.rs:3-14); conditions the reference
reports by panicking raise ``RandNLAPanic`` carrying the same message.

Arrays: numpy arrays (host) or torch tensors on the GPU.  Matrices are
column-major like nalgebra's DMatrix; results come back as numpy arrays for
host inputs and as column-major torch tensors for device inputs.  There is no
CPU fallback: without the compiled library or a GPU every call raises.
"""
import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librandnla_amd.so")

GAUSSIAN, UNIFORM, RADEMACHER, SPARSE_SIGN = 0, 1, 2, 3
ROW, COLUMN = 0, 1
THREEFRY2X64, PHILOX4X32 = 0, 1
FP64, BF16, FP8 = 0, 1, 2
F64, F32 = 0, 1
CGLS, LSQR = 0, 1
LEFT, RIGHT = 0, 1

_KINDS = {1: "InvalidParameters", 2: "InvalidDimensions", 3: "NegativeDimensions",
          4: "NotOverdetermined", 5: "NotSquare", 6: "SingularMatrix",
          7: "MatrixDecompositionError", 8: "NotHermitian", 9: "NotPositiveSemiDefinite",
          10: "ComputationError", 100: "HipError", 101: "NoDevice"}
PANIC = 11


class RandNLAError(Exception):
    """Status from the library; ``kind`` names the RandNLAError variant."""

    def __init__(self, code, msg):
        self.code = code
        self.kind = _KINDS.get(code, f"Status{code}")
        super().__init__(f"{self.kind}: {msg}")


class RandNLAPanic(RuntimeError):
    """A precondition the reference enforces with a panic."""


class Rng(C.Structure):
    _fields_ = [("gen", C.c_int32), ("key", C.c_uint64 * 2), ("ctr", C.c_uint64 * 2)]


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("power_iters", C.c_int32), ("sketch", C.c_int32),
                ("zeta", C.c_int32), ("solver", C.c_int32), ("faithful", C.c_int32),
                ("rng", Rng)]


class SolveInfo(C.Structure):
    _fields_ = [("istop", C.c_int32), ("itn", C.c_int32), ("r1norm", C.c_double),
                ("r2norm", C.c_double), ("anorm", C.c_double), ("acond", C.c_double),
                ("arnorm", C.c_double), ("xnorm", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LsProblem(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("lda", C.c_int64),
                ("a_offset", C.c_int64), ("b_offset", C.c_int64), ("x_offset", C.c_int64),
                ("stream_id", C.c_int64)]


class LsInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("istop", C.c_int32), ("itn", C.c_int32),
                ("d", C.c_int32), ("r2norm", C.c_double), ("arnorm", C.c_double),
                ("anorm", C.c_double), ("acond", C.c_double), ("xnorm", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None
P = C.c_void_p
I64 = C.c_int64
D = C.c_double

_SIGS = {
    "randnla_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "randnla_destroy": (None, [P]),
    "randnla_last_error": (C.c_char_p, [P]),
    "randnla_set_stream": (C.c_int, [P, P]),
    "randnla_get_stream": (C.c_int, [P, C.POINTER(C.c_void_p)]),
    "randnla_opts_default": (None, [C.POINTER(Opts)]),
    "randnla_timing_enable": (C.c_int, [P, C.c_int]),
    "randnla_timing_query": (C.c_int, [P, C.c_char_p, C.POINTER(D), C.POINTER(I64)]),
    "randnla_timing_reset": (C.c_int, [P]),
    "randnla_diag": (C.c_int, [P, C.c_int, C.POINTER(D)]),
    "randnla_philox4x32": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "randnla_philox2x32": (None, [C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_uint32)]),
    "randnla_philox2x64": (None, [C.POINTER(C.c_uint64), C.c_uint64, C.POINTER(C.c_uint64)]),
    "randnla_philox4x64": (None, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "randnla_threefry2x64": (None, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "randnla_rng_from_seed": (None, [C.c_int, C.c_uint64, C.POINTER(Rng)]),
    "randnla_stream_words": (C.c_int, [P, C.POINTER(Rng), C.c_uint64, I64, P]),
    "randnla_sketching_operator": (C.c_int, [P, C.c_int, I64, I64, C.POINTER(Rng), C.c_int, P, I64]),
    "randnla_haar_sample": (C.c_int, [P, I64, I64, C.c_int, C.POINTER(Rng), P, I64]),
    "randnla_sketch_apply": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(Rng), P, C.c_int,
                                       I64, I64, I64, I64, P, I64]),
    "randnla_sketch_apply_noalloc": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(Rng), P,
                                               C.c_int, I64, I64, I64, I64, P, I64]),
    "randnla_rand_svd": (C.c_int, [P, P, C.c_int, I64, I64, I64, I64, D, I64, C.POINTER(Opts),
                                   P, I64, P, P, I64, C.POINTER(I64)]),
    "randnla_qb1": (C.c_int, [P, P, I64, I64, I64, I64, D, C.POINTER(Opts), P, I64, P, I64]),
    "randnla_rf1": (C.c_int, [P, P, I64, I64, I64, I64, C.POINTER(Opts), P, I64]),
    "randnla_tsog1": (C.c_int, [P, P, I64, I64, I64, I64, C.c_int, C.c_int, C.POINTER(Opts), P, I64]),
    "randnla_economic_qrcp": (C.c_int, [P, P, I64, I64, I64, I64, P, I64, P, I64, P]),
    "randnla_osid_qrcp": (C.c_int, [P, P, I64, I64, I64, I64, C.c_int, P, I64, P]),
    "randnla_osid_randomised": (C.c_int, [P, P, I64, I64, I64, I64, C.c_int, C.POINTER(Opts), P,
                                          I64, P]),
    "randnla_two_sided_id": (C.c_int, [P, P, I64, I64, I64, I64, P, I64, P, P, P, I64]),
    "randnla_two_sided_id_randomised": (C.c_int, [P, P, I64, I64, I64, I64, C.POINTER(Opts), P,
                                                  I64, P, P, P, I64]),
    "randnla_cur": (C.c_int, [P, P, I64, I64, I64, I64, P, P, I64, P]),
    "randnla_cur_randomised": (C.c_int, [P, P, I64, I64, I64, I64, C.POINTER(Opts), P, P, I64, P]),
    "randnla_rand_evd1": (C.c_int, [P, P, I64, I64, I64, I64, D, I64, C.POINTER(Opts), P, I64, P,
                                    C.POINTER(I64)]),
    "randnla_rand_evd2": (C.c_int, [P, P, I64, I64, I64, I64, I64, C.POINTER(Opts), P, I64, P,
                                    C.POINTER(I64)]),
    "randnla_orth": (C.c_int, [P, P, I64, I64, I64, P, I64, P, I64]),
    "randnla_stabilizer": (C.c_int, [P, P, I64, I64, I64, P, I64]),
    "randnla_cholqr2": (C.c_int, [P, P, I64, I64, I64, P, I64, P, I64]),
    "randnla_blendenpik": (C.c_int, [P, P, I64, I64, I64, P, D, I64, D, C.POINTER(Opts), P,
                                     C.POINTER(SolveInfo)]),
    "randnla_lsrn": (C.c_int, [P, P, I64, I64, I64, P, D, I64, D, C.POINTER(Opts), P,
                               C.POINTER(SolveInfo)]),
    "randnla_sketched_least_squares_qr": (C.c_int, [P, P, I64, I64, I64, P, C.POINTER(Opts), P]),
    "randnla_sketched_least_squares_svd": (C.c_int, [P, P, I64, I64, I64, P, C.POINTER(Opts), P]),
    "randnla_saddle": (C.c_int, [P, P, I64, I64, I64, P, P, D, D, I64, D, C.POINTER(Opts), P, P,
                                 C.POINTER(SolveInfo)]),
    "randnla_cgls": (C.c_int, [P, P, I64, I64, I64, P, D, I64, P, P, C.POINTER(SolveInfo)]),
    "randnla_lsqr": (C.c_int, [P, P, I64, I64, I64, P, D, D, D, D, I64, C.c_int, P, P, P, P,
                               C.POINTER(SolveInfo)]),
    "randnla_solve_upper_triangular": (C.c_int, [P, P, I64, I64, P, P]),
    "randnla_solve_upper_triangular_f32": (C.c_int, [P, P, I64, I64, P, P]),
    "randnla_qrcp": (C.c_int, [P, P, I64, I64, I64, P, I64, P, I64, P]),
    "randnla_cqrrpt": (C.c_int, [P, P, I64, I64, I64, I64, C.POINTER(Opts), P, I64, P, I64, P,
                                 C.POINTER(I64)]),
    "randnla_test_matrix": (C.c_int, [P, I64, I64, P, D, C.POINTER(Rng), P, C.c_int, I64]),
    "randnla_gemm": (C.c_int, [P, C.c_int, C.c_int, I64, I64, I64, P, C.c_int, I64, P, C.c_int,
                               I64, P, I64]),
    "randnla_spectral_error": (C.c_int, [P, P, C.c_int, I64, I64, I64, P, I64, P, I64, P, I64,
                                         C.c_int, C.POINTER(D), C.POINTER(D)]),
    "randnla_batch_ls_create": (C.c_int, [P, C.POINTER(LsProblem), I64, D, D, I64,
                                          C.POINTER(Opts), C.c_int, C.POINTER(C.c_void_p)]),
    "randnla_batch_ls_execute": (C.c_int, [P, P, P, P, C.POINTER(LsInfo)]),
    "randnla_batch_ls_sketch": (C.c_int, [P, P, I64]),
    "randnla_batch_ls_config": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(D)]),
    "randnla_batch_ls_destroy": (None, [P]),
    "randnla_batch_ls_solve": (C.c_int, [P, C.POINTER(LsProblem), I64, P, P, P, D, D, I64,
                                         C.POINTER(Opts), C.POINTER(LsInfo)]),
    "randnla_mlp_param_count": (C.c_int, []),
    "randnla_mlp_forward": (C.c_int, [P, P, P, I64, P]),
    "randnla_mlp_grad": (C.c_int, [P, P, P, P, P, I64, P, P]),
    "randnla_mlp_train_step": (C.c_int, [P, P, P, P, P, P, P, I64, C.c_int, C.c_float, P]),
    "randnla_probe_features": (C.c_int, [P, P, C.c_int, I64, I64, I64, C.POINTER(Rng), P]),
}


def lib():
    """Load the compiled library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m randnla_amd.build`")
        # The library and PyTorch must share one HIP runtime (same soname):
        # load PyTorch's first so that device tensors and the library agree.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
            if name in _CTX_CALLS:
                setattr(L, name, _stream_ordered(f, _ctx_of_handle))
            elif name in _PLAN_CALLS:
                setattr(L, name, _stream_ordered(f, _ctx_of_plan))
        _lib = L
    return _lib


# ------------------------------------------- ordering against torch streams
# The library runs on its context's own stream.  Device tensors handed to it
# are produced (and its outputs consumed) on torch's current stream, and
# torch's caching allocator may hand out blocks whose previous use is still
# queued there.  Every call that takes a context (or a plan) therefore makes
# the context's stream wait for torch's current stream before the call and
# torch's current stream wait for the context's stream after it.
_NO_ORDER = {"randnla_create", "randnla_destroy", "randnla_last_error", "randnla_set_stream",
             "randnla_get_stream", "randnla_opts_default", "randnla_timing_enable",
             "randnla_timing_query", "randnla_timing_reset", "randnla_diag",
             "randnla_philox4x32", "randnla_philox2x32", "randnla_philox2x64",
             "randnla_philox4x64", "randnla_threefry2x64", "randnla_rng_from_seed",
             "randnla_mlp_param_count", "randnla_batch_ls_config", "randnla_batch_ls_destroy"}
_PLAN_CALLS = {"randnla_batch_ls_execute", "randnla_batch_ls_sketch"}
_CTX_CALLS = {n for n in _SIGS if n not in _NO_ORDER and n not in _PLAN_CALLS}
_CONTEXTS = weakref.WeakValueDictionary()   # handle value -> Context
_PLANS = weakref.WeakValueDictionary()      # plan handle value -> BatchLS


def _hval(h):
    return h.value if isinstance(h, C.c_void_p) else h


def _ctx_of_handle(h):
    return _CONTEXTS.get(_hval(h))


def _ctx_of_plan(h):
    plan = _PLANS.get(_hval(h))
    return plan.ctx if plan is not None else None


def _stream_ordered(f, find_ctx):
    def call(*args):
        torch = sys.modules.get("torch")
        pair = None
        if torch is not None and args and torch.cuda.is_initialized():
            ctx = find_ctx(args[0])
            if ctx is not None:
                pair = ctx._torch_streams(torch)
        if pair is None:
            return f(*args)
        cur, ext = pair
        ext.wait_stream(cur)
        try:
            return f(*args)
        finally:
            cur.wait_stream(ext)
    call.__name__ = getattr(f, "__name__", "call")
    return call


EXPORTED = tuple(_SIGS)


# ------------------------------------------------------------------ context
class Context:
    """One library context (HIP stream + diagnostics) on a device."""

    def __init__(self, device=0):
        h = C.c_void_p()
        st = lib().randnla_create(C.byref(h), device)
        if st != 0:
            raise RandNLAError(st, "randnla_create failed (no GPU or HIP runtime error)")
        self.h = h
        self.device = device
        self._ext = None
        _CONTEXTS[h.value] = self

    def close(self):
        if self.h:
            _CONTEXTS.pop(self.h.value, None)
            lib().randnla_destroy(self.h)
            self.h = None
            self._ext = None

    def stream(self):
        """The HIP stream (handle value) this context's work runs on."""
        p = C.c_void_p()
        self.check(lib().randnla_get_stream(self.h, C.byref(p)))
        return p.value or 0

    def _torch_streams(self, torch):
        """(torch's current stream, the context's stream as a torch stream),
        or None when they are the same stream."""
        cur = torch.cuda.current_stream(self.device)
        if self._ext is None:
            self._ext = torch.cuda.ExternalStream(self.stream(), device=self.device)
        if self._ext.cuda_stream == cur.cuda_stream:
            return None
        return cur, self._ext

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st):
        if st == 0:
            return
        msg = lib().randnla_last_error(self.h).decode("utf-8", "replace")
        if st == PANIC:
            raise RandNLAPanic(msg)
        raise RandNLAError(st, msg)

    def set_stream(self, stream_handle):
        self.check(lib().randnla_set_stream(self.h, C.c_void_p(stream_handle)))
        self._ext = None

    def timing(self, on=True):
        self.check(lib().randnla_timing_enable(self.h, 1 if on else 0))

    def timing_reset(self):
        self.check(lib().randnla_timing_reset(self.h))

    def timer(self, name):
        ms, n = D(), I64()
        self.check(lib().randnla_timing_query(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def diag(self, i):
        v = D()
        self.check(lib().randnla_diag(self.h, i, C.byref(v)))
        return v.value


_ctx = None


def context():
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


@atexit.register
def _close_all():
    """Release plans and contexts while the HIP runtime is still up (atexit
    handlers run before the runtime's own teardown; __del__ during
    interpreter shutdown may not)."""
    global _ctx
    for plan in list(_PLANS.values()):
        plan.close()
    for ctx in list(_CONTEXTS.values()):
        ctx.close()
    _ctx = None


# ------------------------------------------------------------ array helpers
def _is_torch(x):
    return type(x).__module__.startswith("torch")


class _Arr:
    """Column-major view of an input/output array for the C ABI."""

    def __init__(self, x, dtype=np.float64):
        if x is None:
            self.obj, self.ptr, self.ld = None, None, 0
            return
        if _is_torch(x):
            import torch
            tdt = torch.float64 if dtype == np.float64 else (torch.float32 if dtype == np.float32 else dtype)
            t = x.to(tdt)
            if t.dim() == 1:
                t = t.contiguous()
                self.ld = t.shape[0]
            else:
                if not (t.stride(0) == 1 and t.stride(1) >= max(1, t.shape[0])):
                    t = t.t().contiguous().t()
                self.ld = max(1, t.stride(1))
            self.obj, self.ptr = t, t.data_ptr()
        else:
            a = np.asarray(x, dtype=dtype)
            if a.ndim == 2:
                a = np.asfortranarray(a)
                self.ld = max(1, a.shape[0])
            else:
                a = np.ascontiguousarray(a)
                self.ld = max(1, a.shape[0] if a.ndim else 1)
            self.obj, self.ptr = a, a.ctypes.data


def _like(ref, shape, dtype=np.float64):
    """Column-major output on the same side (host/device) as ref."""
    if _is_torch(ref):
        import torch
        tdt = {np.float64: torch.float64, np.float32: torch.float32, np.int64: torch.int64,
               np.uint64: torch.int64}[dtype]
        if len(shape) == 1:
            return torch.empty(shape, dtype=tdt, device=ref.device)
        return torch.empty((shape[1], shape[0]), dtype=tdt, device=ref.device).t()
    return np.zeros(shape, dtype=dtype, order="F")


def _ptr(x):
    if x is None:
        return None
    return x.data_ptr() if _is_torch(x) else x.ctypes.data


def _ld(x):
    if _is_torch(x):
        return max(1, x.stride(1)) if x.dim() == 2 else x.shape[0]
    return max(1, x.shape[0])


def _dtype_of(x):
    if _is_torch(x):
        import torch
        return F32 if x.dtype == torch.float32 else F64
    return F32 if np.asarray(x).dtype == np.float32 else F64


def _opts(precision=FP64, power_iters=None, sketch=None, zeta=8, solver=CGLS, rng=None,
          faithful=False):
    o = Opts()
    lib().randnla_opts_default(C.byref(o))
    o.precision = precision
    o.faithful = 1 if faithful else 0
    if power_iters is not None:
        o.power_iters = power_iters
    if sketch is not None:
        o.sketch = sketch
    o.zeta = zeta
    o.solver = solver
    if rng is not None:
        o.rng = rng
    return o


def rng_from_seed(gen, seed):
    """``Rng::seed_from_u64(seed)`` (rand_core 0.6.4) for the given generator."""
    r = Rng()
    lib().randnla_rng_from_seed(gen, seed, C.byref(r))
    return r


def rng(gen, key, ctr=(0, 0)):
    r = Rng()
    r.gen = gen
    r.key[0], r.key[1] = int(key[0]), int(key[1])
    r.ctr[0], r.ctr[1] = int(ctr[0]), int(ctr[1])
    return r


# --------------------------------------------------------------- RNG (host)
def philox4x32(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().randnla_philox4x32(c, k, o)
    return list(o)


def philox2x32(ctr, key):
    c = (C.c_uint32 * 2)(*ctr)
    o = (C.c_uint32 * 2)()
    lib().randnla_philox2x32(c, key, o)
    return list(o)


def philox2x64(ctr, key):
    c = (C.c_uint64 * 2)(*ctr)
    o = (C.c_uint64 * 2)()
    lib().randnla_philox2x64(c, key, o)
    return list(o)


def philox4x64(ctr, key):
    c = (C.c_uint64 * 4)(*ctr)
    k = (C.c_uint64 * 2)(*key)
    o = (C.c_uint64 * 4)()
    lib().randnla_philox4x64(c, k, o)
    return list(o)


def threefry2x64(ctr, key):
    c = (C.c_uint64 * 2)(*ctr)
    k = (C.c_uint64 * 2)(*key)
    o = (C.c_uint64 * 2)()
    lib().randnla_threefry2x64(c, k, o)
    return list(o)


def stream_words(r, first, n):
    """u64 words [first, first + n) of stream r, generated on the GPU."""
    ctx = context()
    out = np.zeros(n, dtype=np.uint64)
    ctx.check(lib().randnla_stream_words(ctx.h, C.byref(r), first, n, out.ctypes.data))
    return out


# --------------------------------------------------------------- sketches
def sketching_operator(dist_type, rows, cols, rng=None, zeta=8, device=False):
    """src/sketch.rs:102-130.  ``rng=None`` uses the reference's stream
    ThreeFry2x64Rng::seed_from_u64(0)."""
    ctx = context()
    if device:
        import torch
        out = torch.empty((max(cols, 1), max(rows, 1)), dtype=torch.float64, device="cuda").t()
    else:
        out = np.zeros((max(rows, 1), max(cols, 1)), order="F")
    ctx.check(lib().randnla_sketching_operator(ctx.h, dist_type, rows, cols,
                                               C.byref(rng) if rng is not None else None,
                                               zeta, _ptr(out), _ld(out)))
    return out


def haar_sample(rows, columns, attr, rng=None):
    """src/sketch.rs:45-85."""
    ctx = context()
    out = np.zeros((max(rows, 1), max(columns, 1)), order="F")
    ctx.check(lib().randnla_haar_sample(ctx.h, rows, columns, attr,
                                        C.byref(rng) if rng is not None else None,
                                        out.ctypes.data, out.shape[0]))
    return out


def sketch_apply(side, dist, A, d, rng=None, zeta=8):
    """Fused S*A (LEFT, d x n) or A*Omega (RIGHT, m x d)."""
    ctx = context()
    dt = _dtype_of(A)
    a = _Arr(A, np.float32 if dt == F32 else np.float64)
    m, n = A.shape
    out = _like(A, (d, n) if side == LEFT else (m, d))
    ctx.check(lib().randnla_sketch_apply(ctx.h, side, dist, zeta,
                                         C.byref(rng) if rng is not None else None,
                                         a.ptr, dt, m, n, a.ld, d, _ptr(out), _ld(out)))
    return out


def sketch_apply_noalloc(side, dist, A, d, rng=None, zeta=8, out=None, ctx=None):
    """randnla_sketch_apply_noalloc: S*A (LEFT) or A*Omega (RIGHT) without any
    device allocation by the library; A (and out, if given) are CUDA tensors.
    The output is allocated here (by torch) when not given."""
    import torch
    ctx = ctx or context()
    dt = _dtype_of(A)
    m, n = A.shape
    if not (A.stride(0) == 1 and A.stride(1) >= max(1, m)):
        raise ValueError("A must be a column-major device tensor")
    if out is None:
        shape = (d, n) if side == LEFT else (m, d)
        out = torch.empty((shape[1], shape[0]), dtype=torch.float64, device=A.device).t()
    ctx.check(lib().randnla_sketch_apply_noalloc(ctx.h, side, dist, zeta,
                                                 C.byref(rng) if rng is not None else None,
                                                 A.data_ptr(), dt, m, n, max(1, A.stride(1)), d,
                                                 out.data_ptr(), max(1, out.stride(1))))
    return out


# -------------------------------------------------------------- low rank
def rand_svd(A, k, epsilon, s, precision=FP64, power_iters=None, rng=None, ctx=None,
             faithful=False):
    """src/lora_drivers.rs:30-69.  Returns (U m x r, S r x r diagonal, Vt r x n)
    like the reference (S as a diagonal matrix), r = min(k, m, n).
    ``faithful`` runs the reference's own computation (tsog1 with its
    even-branch discard, full-pivot LU stabiliser, Householder Orth)."""
    U, sv, Vt = rand_svd_factors(A, k, epsilon, s, precision, power_iters, rng, ctx, faithful)
    if _is_torch(sv):
        import torch
        return U, torch.diag(sv), Vt
    return U, np.diag(sv), Vt


def rand_svd_factors(A, k, epsilon, s, precision=FP64, power_iters=None, rng=None, ctx=None,
                     faithful=False):
    """rand_svd returning the singular values as a vector."""
    ctx = ctx or context()
    dt = _dtype_of(A)
    a = _Arr(A, np.float32 if dt == F32 else np.float64)
    m, n = A.shape
    r = max(0, min(k, k + s, m, n))
    U = _like(A, (m, max(r, 1)))
    sv = _like(A, (max(r, 1),))
    Vt = _like(A, (max(r, 1), n))
    o = _opts(precision=precision, power_iters=power_iters, rng=rng, faithful=faithful)
    rout = I64()
    ctx.check(lib().randnla_rand_svd(ctx.h, a.ptr, dt, m, n, a.ld, k, epsilon, s, C.byref(o),
                                     _ptr(U), _ld(U), _ptr(sv), _ptr(Vt), _ld(Vt), C.byref(rout)))
    r = rout.value
    return U[:, :r], sv[:r], Vt[:r, :]


def QB1(A, k, epsilon, power_iters=None, rng=None, faithful=False):
    """src/lora_helpers.rs:17-23 (epsilon ignored, as in the reference):
    Q (m x kc), B (kc x n), kc = min(k, m, n)."""
    ctx = context()
    a = _Arr(A)
    m, n = A.shape
    Q = _like(A, (m, k))
    B = _like(A, (k, n))
    o = _opts(power_iters=power_iters, rng=rng, faithful=faithful)
    ctx.check(lib().randnla_qb1(ctx.h, a.ptr, m, n, a.ld, k, epsilon, C.byref(o), _ptr(Q),
                                _ld(Q), _ptr(B), _ld(B)))
    kc = min(k, m, n)
    return Q[:, :kc], B[:kc, :]


def RF1(A, k, power_iters=None, rng=None, faithful=False):
    """src/lora_helpers.rs:37-44: Q (m x min(k, m, n))."""
    ctx = context()
    a = _Arr(A)
    m, n = A.shape
    Q = _like(A, (m, k))
    o = _opts(power_iters=power_iters, rng=rng, faithful=faithful)
    ctx.check(lib().randnla_rf1(ctx.h, a.ptr, m, n, a.ld, k, C.byref(o), _ptr(Q), _ld(Q)))
    return Q[:, :min(k, m, n)]


def tsog1_cols(m, n, k, num_passes, passes_per_stab):
    """Columns of the reference's tsog1 result (src/lora_helpers.rs:58-105):
    k when the loop does not run, else what the last iteration's stabilisers
    (full-pivot LU factors, min(rows, cols) columns) leave."""
    done0 = num_passes % 2
    j = (num_passes - done0) // 2
    if j <= 0:
        return k
    k1 = min(m, k) if (done0 + 2 * j - 1) % passes_per_stab == 0 else k
    return min(n, k1) if (done0 + 2 * j) % passes_per_stab == 0 else k1


def tsog1(A, k, num_passes, passes_per_stab, rng=None, faithful=False):
    """src/lora_helpers.rs:58-105.  Intended method (the loop starts from the
    generated test matrix; n x k), or with ``faithful`` the reference's own
    result (n x tsog1_cols(...)), drawn from the reference's stream
    ThreeFry2x64 seed_from_u64(0) (``rng`` is then not used)."""
    ctx = context()
    a = _Arr(A)
    m, n = A.shape
    S = _like(A, (n, k))
    o = _opts(rng=rng, faithful=faithful)
    ctx.check(lib().randnla_tsog1(ctx.h, a.ptr, m, n, a.ld, k, num_passes, passes_per_stab,
                                  C.byref(o), _ptr(S), _ld(S)))
    if faithful:
        return S[:, :tsog1_cols(m, n, k, num_passes, passes_per_stab)]
    return S


def _evd(fn, A, extra, k, s, power_iters, rng, ctx):
    ctx = ctx or context()
    a = _Arr(A)
    m, n = A.shape
    r = max(1, min(k, n)) if k > 0 else 1
    V = _like(A, (n, r))
    lam = _like(A, (r,))
    o = _opts(power_iters=power_iters, rng=rng)
    rout = I64()
    ctx.check(fn(ctx.h, a.ptr, m, n, a.ld, k, *extra, C.byref(o), _ptr(V), _ld(V), _ptr(lam),
                 C.byref(rout)))
    r = rout.value
    return V[:, :r], lam[:r]


def rand_evd1(A, k, epsilon, s, power_iters=None, rng=None, ctx=None):
    """src/lora_drivers.rs:87-151.  Returns (V n x r, lambda r) with lambda
    ordered by |lambda| descending, r = min(k, k + s, n)."""
    return _evd(lib().randnla_rand_evd1, A, (epsilon, s), k, s, power_iters, rng, ctx)


def rand_evd2(A, k, s, rng=None, ctx=None):
    """src/lora_drivers.rs:167-224.  Returns (V n x r, lambda r), lambda
    descending, r = min(k, number of sigma^2 above the nu shift)."""
    return _evd(lib().randnla_rand_evd2, A, (s,), k, s, None, rng, ctx)


def _idx(ref, k):
    if _is_torch(ref):
        import torch
        return torch.empty(max(k, 1), dtype=torch.int64, device=ref.device)
    return np.zeros(max(k, 1), dtype=np.int64)


def _opt_ptr(rng, sketch=None, zeta=8):
    if rng is None and sketch is None:
        return None
    return C.byref(_opts(sketch=sketch, zeta=zeta, rng=rng))


def economic_qrcp(a, k, ctx=None):
    """src/pivot_decompositions.rs:194-269.  Returns (Q m x k, R k x n, p)."""
    ctx = ctx or context()
    aa = _Arr(a)
    m, n = a.shape
    kk = max(k, 1)
    Q, R, p = _like(a, (m, kk)), _like(a, (kk, n)), _idx(a, n)
    ctx.check(lib().randnla_economic_qrcp(ctx.h, aa.ptr, m, n, aa.ld, k, _ptr(Q), _ld(Q), _ptr(R),
                                          _ld(R), _ptr(p)))
    return Q, R, p


def osid_qrcp(y, k, attr, ctx=None):
    """src/id.rs:248-305.  Returns (X, J): Column X k x n, Row X m x k."""
    ctx = ctx or context()
    a = _Arr(y)
    m, n = y.shape
    kk = max(k, 1)
    X = _like(y, (kk, n) if attr == COLUMN else (m, kk))
    J = _idx(y, kk)
    ctx.check(lib().randnla_osid_qrcp(ctx.h, a.ptr, m, n, a.ld, k, attr, _ptr(X), _ld(X), _ptr(J)))
    return X, J[:k]


def osid_randomised(a, k, attr, rng=None, sketch=None, ctx=None):
    """src/id.rs:205-234."""
    ctx = ctx or context()
    aa = _Arr(a)
    m, n = a.shape
    kk = max(k, 1)
    X = _like(a, (kk, n) if attr == COLUMN else (m, kk))
    J = _idx(a, kk)
    ctx.check(lib().randnla_osid_randomised(ctx.h, aa.ptr, m, n, aa.ld, k, attr,
                                            _opt_ptr(rng, sketch), _ptr(X), _ld(X), _ptr(J)))
    return X, J[:k]


def _two_sided(randomised, a, k, rng, ctx=None):
    ctx = ctx or context()
    aa = _Arr(a)
    m, n = a.shape
    kk = max(k, 1)
    Z, X = _like(a, (m, kk)), _like(a, (kk, n))
    I, J = _idx(a, kk), _idx(a, kk)
    if randomised:
        ctx.check(lib().randnla_two_sided_id_randomised(ctx.h, aa.ptr, m, n, aa.ld, k,
                                                        _opt_ptr(rng), _ptr(Z), _ld(Z), _ptr(I),
                                                        _ptr(J), _ptr(X), _ld(X)))
    else:
        ctx.check(lib().randnla_two_sided_id(ctx.h, aa.ptr, m, n, aa.ld, k, _ptr(Z), _ld(Z),
                                             _ptr(I), _ptr(J), _ptr(X), _ld(X)))
    return Z, I[:k], J[:k], X


def two_sided_id(a, k, ctx=None):
    """src/id.rs:125-132.  Returns (Z, I, J, X) with A ~ Z A[I, J] X."""
    return _two_sided(False, a, k, None, ctx)


def two_sided_id_randomised(a, k, rng=None, ctx=None):
    """src/id.rs:94-101."""
    return _two_sided(True, a, k, rng, ctx)


def _cur(randomised, a, k, rng, ctx=None):
    ctx = ctx or context()
    aa = _Arr(a)
    m, n = a.shape
    kk = max(k, 1)
    U = _like(a, (kk, kk))
    I, J = _idx(a, kk), _idx(a, kk)
    if randomised:
        ctx.check(lib().randnla_cur_randomised(ctx.h, aa.ptr, m, n, aa.ld, k, _opt_ptr(rng),
                                               _ptr(J), _ptr(U), _ld(U), _ptr(I)))
    else:
        ctx.check(lib().randnla_cur(ctx.h, aa.ptr, m, n, aa.ld, k, _ptr(J), _ptr(U), _ld(U),
                                    _ptr(I)))
    return J[:k], U, I[:k]


def cur(a, k, ctx=None):
    """src/id.rs:33-76.  Returns (J, U, I) with A ~ A[:, J] U A[I, :]."""
    return _cur(False, a, k, None, ctx)


def cur_randomised(a, k, rng=None, ctx=None):
    """src/id.rs:154-197."""
    return _cur(True, a, k, rng, ctx)


def Orth(X, return_r=False):
    """src/lora_helpers.rs:131-133: Q (m x min(m, n)) of the Householder thin
    QR with nalgebra's sign convention (R_ii >= 0); with ``return_r`` also
    R (min(m, n) x n)."""
    ctx = context()
    x = _Arr(X)
    m, n = X.shape
    kq = min(m, n)
    Q = _like(X, (m, kq))
    R = _like(X, (kq, n)) if return_r else None
    ctx.check(lib().randnla_orth(ctx.h, x.ptr, m, n, x.ld, _ptr(Q), _ld(Q),
                                 _ptr(R) if R is not None else None, _ld(R) if R is not None else 0))
    return (Q, R) if return_r else Q


def cholqr2(X):
    """CholeskyQR2 of a tall X (no reference counterpart; the range finder's
    orthonormalisation): (Q m x n, R n x n) with X = Q R."""
    ctx = context()
    x = _Arr(X)
    m, n = X.shape
    Q = _like(X, (m, n))
    R = _like(X, (n, n))
    ctx.check(lib().randnla_cholqr2(ctx.h, x.ptr, m, n, x.ld, _ptr(Q), _ld(Q), _ptr(R), _ld(R)))
    return Q, R


def Stabilizer(X):
    """src/lora_helpers.rs:144-146: L (m x min(m, n)), the unit lower factor
    of X's full-pivot LU (nalgebra FullPivLU), rows in pivoted order."""
    ctx = context()
    x = _Arr(X)
    m, n = X.shape
    Y = _like(X, (m, min(m, n)))
    ctx.check(lib().randnla_stabilizer(ctx.h, x.ptr, m, n, x.ld, _ptr(Y), _ld(Y)))
    return Y


# --------------------------------------------------------- least squares
def _ls_common(fn, A, b, epsilon, l, sampling_factor, sketch, zeta, solver, rng, ctx=None):
    ctx = ctx or context()
    a = _Arr(A)
    m, n = A.shape
    bb = _Arr(np.asarray(b).reshape(-1) if not _is_torch(b) else b.reshape(-1))
    x = _like(A, (n,))
    info = SolveInfo()
    o = _opts(sketch=sketch, zeta=zeta, solver=solver, rng=rng)
    ctx.check(fn(ctx.h, a.ptr, m, n, a.ld, bb.ptr, epsilon, l, sampling_factor,
                 C.byref(o) if (rng is not None or sketch not in (None, GAUSSIAN) or solver != CGLS) else None,
                 _ptr(x), C.byref(info)))
    return x, info.as_dict()


def blendenpik_overdetermined(a, b, epsilon, l, sampling_factor, sketch=None, zeta=8,
                              solver=CGLS, rng=None, return_info=False, ctx=None):
    """src/sketch_and_precondition.rs:26-59."""
    x, info = _ls_common(lib().randnla_blendenpik, a, b, epsilon, l, sampling_factor, sketch,
                         zeta, solver, rng, ctx)
    return (x, info) if return_info else x


def lsrn_overdetermined(a, b, epsilon, l, sampling_factor, sketch=None, zeta=8, solver=CGLS,
                        rng=None, return_info=False, ctx=None):
    """src/sketch_and_precondition.rs:82-119."""
    x, info = _ls_common(lib().randnla_lsrn, a, b, epsilon, l, sampling_factor, sketch, zeta,
                         solver, rng, ctx)
    return (x, info) if return_info else x


def _sketch_solve(fn, A, b, sketch, zeta, rng, ctx):
    ctx = ctx or context()
    a = _Arr(A)
    m, n = A.shape
    bb = _Arr(np.asarray(b).reshape(-1) if not _is_torch(b) else b.reshape(-1))
    x = _like(A, (n,))
    o = _opts(sketch=sketch, zeta=zeta, rng=rng)
    ctx.check(fn(ctx.h, a.ptr, m, n, a.ld, bb.ptr,
                 C.byref(o) if (rng is not None or sketch not in (None, GAUSSIAN)) else None,
                 _ptr(x)))
    return x


def sketched_least_squares_qr(a, b, sketch=None, zeta=8, rng=None, ctx=None):
    """src/sketch_and_solve.rs:24-33 (m / 4 sketch rows, QR of the sketch)."""
    return _sketch_solve(lib().randnla_sketched_least_squares_qr, a, b, sketch, zeta, rng, ctx)


def sketched_least_squares_svd(a, b, sketch=None, zeta=8, rng=None, ctx=None):
    """src/sketch_and_solve.rs:53-65 (m / 4 sketch rows, SVD of the sketch)."""
    return _sketch_solve(lib().randnla_sketched_least_squares_svd, a, b, sketch, zeta, rng, ctx)


def sketch_saddle_point_precondition(a, b, c, mu, epsilon, l, sampling_factor, sketch=None,
                                     zeta=8, solver=CGLS, rng=None):
    """src/sketch_and_precondition.rs:150-216.  Returns (x, y = b - A x)."""
    ctx = context()
    A = _Arr(a)
    m, n = a.shape
    bb = _Arr(np.asarray(b).reshape(-1))
    cc = None
    if c is not None and np.asarray(c).size:
        cc = _Arr(np.asarray(c).reshape(-1))
    x = _like(a, (n,))
    y = _like(a, (m,))
    info = SolveInfo()
    o = _opts(sketch=sketch, zeta=zeta, solver=solver, rng=rng)
    use = rng is not None or sketch not in (None, GAUSSIAN) or solver != CGLS
    ctx.check(lib().randnla_saddle(ctx.h, A.ptr, m, n, A.ld, bb.ptr, cc.ptr if cc else None, mu,
                                   epsilon, l, sampling_factor, C.byref(o) if use else None,
                                   _ptr(x), _ptr(y), C.byref(info)))
    return x, y


def cgls(a, b, tolerance, num_iterations, x=None, return_info=False):
    """src/cg.rs:18-61."""
    ctx = context()
    A = _Arr(a)
    m, n = a.shape
    bb = _Arr(np.asarray(b).reshape(-1))
    x0 = _Arr(np.asarray(x).reshape(-1)) if x is not None else None
    out = _like(a, (n,))
    info = SolveInfo()
    ctx.check(lib().randnla_cgls(ctx.h, A.ptr, m, n, A.ld, bb.ptr, tolerance, num_iterations,
                                 x0.ptr if x0 else None, _ptr(out), C.byref(info)))
    return (out, info.as_dict()) if return_info else out


def lsqr(a, b, damp=0.0, atol=1e-8, btol=1e-8, conlim=1e8, iter_lim=None, calc_var=False,
         x0=None):
    """src/solvers.rs:115-278.  Returns the reference's 10-tuple
    (x, istop, itn, r1norm, r2norm, anorm, acond, arnorms, xnorm, var)."""
    ctx = context()
    A = _Arr(a)
    m, n = a.shape
    bb = _Arr(np.asarray(b).reshape(-1))
    xi = _Arr(np.asarray(x0).reshape(-1)) if x0 is not None else None
    x = np.zeros(n)
    var = np.zeros(n)
    lim = iter_lim if iter_lim is not None else 2 * n
    arn = np.zeros(max(lim, 1) + 1)
    info = SolveInfo()
    ctx.check(lib().randnla_lsqr(ctx.h, A.ptr, m, n, A.ld, bb.ptr, damp, atol, btol, conlim,
                                 lim, 1 if calc_var else 0, xi.ptr if xi else None,
                                 x.ctypes.data, var.ctypes.data, arn.ctypes.data, C.byref(info)))
    itn = info.itn
    arnorms = list(arn[:max(itn, 1)]) if itn > 0 else [0.0]
    return (x, info.istop, itn, info.r1norm, info.r2norm, info.anorm, info.acond, arnorms,
            info.xnorm, var)


def solve_upper_triangular_system(u, y):
    """src/solvers.rs:22-41 (zero diagonal -> x_i = 0); f32 inputs stay f32."""
    ctx = context()
    f32 = np.asarray(u).dtype == np.float32
    dt = np.float32 if f32 else np.float64
    U = _Arr(u, dt)
    n = np.asarray(u).shape[0]
    yy = _Arr(np.asarray(y, dtype=dt).reshape(-1), dt)
    x = np.zeros(n, dtype=dt)
    fn = lib().randnla_solve_upper_triangular_f32 if f32 else lib().randnla_solve_upper_triangular
    ctx.check(fn(ctx.h, U.ptr, n, U.ld, yy.ptr, x.ctypes.data))
    return x.reshape(-1, 1)


def qrcp(a):
    """src/pivot_decompositions.rs:106-175: (Q m x m, R m x n, perm)."""
    ctx = context()
    A = _Arr(a)
    m, n = a.shape
    Q = np.zeros((m, m), order="F")
    R = np.zeros((m, n), order="F")
    p = np.zeros(n, dtype=np.int64)
    ctx.check(lib().randnla_qrcp(ctx.h, A.ptr, m, n, A.ld, Q.ctypes.data, m, R.ctypes.data, m,
                                 p.ctypes.data))
    return Q, R, p


def sap_chol_qrcp(a, d, sketch=None, rng=None, ctx=None):
    """src/cqrrpt.rs:27-58: (Q m x k, R k x n, J).  Raises RandNLAPanic
    "d must satisfy n ≤ d ≪ m" for a bad d, as the reference panics."""
    ctx = ctx or context()
    A = _Arr(a)
    m, n = a.shape
    Q = _like(a, (m, max(n, 1)))
    R = _like(a, (max(n, 1), max(n, 1)))
    J = _idx(a, n)
    kk = I64()
    use = rng is not None or sketch not in (None, GAUSSIAN)
    o = _opts(sketch=sketch, rng=rng)
    ctx.check(lib().randnla_cqrrpt(ctx.h, A.ptr, m, n, A.ld, d, C.byref(o) if use else None,
                                   _ptr(Q), _ld(Q), _ptr(R), _ld(R), _ptr(J), C.byref(kk)))
    k = kk.value
    return Q[:, :k], R[:k, :n], J[:n]


# ------------------------------------------------------------ test matrices
def test_matrix(m, n, sigma, noise=0.0, rng=None, dtype=F64, device=True):
    """D_r H_m[:, J] diag(sigma) (H_n D_c)^T / sqrt(mn) + noise * G (m, n powers of 2)."""
    ctx = context()
    sg = np.ascontiguousarray(np.asarray(sigma, dtype=np.float64))
    if device:
        import torch
        tdt = torch.float32 if dtype == F32 else torch.float64
        out = torch.empty((n, m), dtype=tdt, device="cuda").t()
    else:
        out = np.zeros((m, n), dtype=np.float32 if dtype == F32 else np.float64, order="F")
    ctx.check(lib().randnla_test_matrix(ctx.h, m, n, sg.ctypes.data, noise,
                                        C.byref(rng) if rng is not None else None, _ptr(out), dtype,
                                        _ld(out)))
    return out


F16X3 = 3
F16X2 = 4
F16 = 5      # op(A) and B each rounded once to FP16 (power-of-two scales), one product
F16_BSPLIT = 6   # op(A) rounded once to FP16, B split into FP16 hi/lo, two products


def gemm(A, B, trans_a=False, precision=FP64):
    """C = op(A) @ B (FP64 result) on the library's MFMA kernels."""
    ctx = context()
    da, db = _dtype_of(A), _dtype_of(B)
    a = _Arr(A, np.float32 if da == F32 else np.float64)
    b = _Arr(B, np.float32 if db == F32 else np.float64)
    K, N = B.shape
    M = A.shape[1] if trans_a else A.shape[0]
    out = _like(A, (M, N))
    ctx.check(lib().randnla_gemm(ctx.h, precision, 1 if trans_a else 0, M, N, K, a.ptr, da, a.ld,
                                 b.ptr, db, b.ld, _ptr(out), _ld(out)))
    return out


def spectral_error(A, U, S, Vt, iters=64, ctx=None):
    """(||A - U diag(S) Vt||_2, ||A||_2) by Golub-Kahan-Lanczos with full
    reorthogonalisation (``iters`` steps); lower bounds converging from
    below."""
    ctx = ctx or context()
    dt = _dtype_of(A)
    a = _Arr(A, np.float32 if dt == F32 else np.float64)
    m, n = A.shape
    u, s, vt = _Arr(U), _Arr(S), _Arr(Vt)
    r = S.shape[0]
    err, an = D(), D()
    ctx.check(lib().randnla_spectral_error(ctx.h, a.ptr, dt, m, n, a.ld, u.ptr, u.ld, s.ptr, r,
                                           vt.ptr, vt.ld, iters, C.byref(err), C.byref(an)))
    return err.value, an.value


# ------------------------------------------------------ batched least squares
class BatchLS:
    """Plan for a batch of independent overdetermined least-squares problems
    solved in one persistent-kernel launch (replayed hipGraph).

    ``shapes`` is a list of (m, n).  Problems are packed column-major, one
    after the other (A_i with lda = m_i in one FP32 buffer, b_i and x_i in
    FP64 buffers); :meth:`pack` builds those buffers on the GPU.  Each problem
    follows blendenpik_overdetermined (src/sketch_and_precondition.rs:26-59)
    with the LSQR solver; ``sampling_factor`` and ``tol`` keep that
    function's meaning (d = min(m, floor(sf n)), atol = btol = tol)."""

    def __init__(self, shapes, sampling_factor=2.0, tol=1e-10, iter_lim=None, rng=None,
                 autotune=False, stream_ids=None, ctx=None):
        self.ctx = ctx or context()
        self.shapes = [(int(m), int(n)) for m, n in shapes]
        if stream_ids is None:
            stream_ids = range(len(self.shapes))
        self.stream_ids = [int(i) for i in stream_ids]
        probs = (LsProblem * len(self.shapes))()
        ao = bo = xo = 0
        self.a_offsets, self.b_offsets, self.x_offsets = [], [], []
        for i, (m, n) in enumerate(self.shapes):
            probs[i] = LsProblem(m, n, m, ao, bo, xo, self.stream_ids[i])
            self.a_offsets.append(ao); self.b_offsets.append(bo); self.x_offsets.append(xo)
            ao += m * n; bo += m; xo += n
        self.a_elems, self.b_elems, self.x_elems = ao, bo, xo
        self._ds = [m if sampling_factor * n > m else int(np.floor(sampling_factor * n))
                    for m, n in self.shapes]
        if iter_lim is None:
            iter_lim = 2 * max([n for _, n in self.shapes] + [1])
        o = _opts(rng=rng if rng is not None else rng_from_seed(PHILOX4X32, 42))
        h = C.c_void_p()
        self.ctx.check(lib().randnla_batch_ls_create(self.ctx.h, probs, len(self.shapes),
                                                     float(sampling_factor), float(tol),
                                                     int(iter_lim), C.byref(o), int(autotune),
                                                     C.byref(h)))
        self.h = h
        _PLANS[h.value] = self

    def pack(self, mats, rhs):
        """Device buffers (A FP32, b FP64) holding the given matrices/vectors."""
        import torch
        A = torch.empty(self.a_elems, dtype=torch.float32, device="cuda")
        b = torch.empty(self.b_elems, dtype=torch.float64, device="cuda")
        for i, (m, n) in enumerate(self.shapes):
            a_i = torch.as_tensor(np.asarray(mats[i], dtype=np.float32).reshape(m, n).T.copy())
            A[self.a_offsets[i]:self.a_offsets[i] + m * n] = a_i.reshape(-1).to("cuda")
            b[self.b_offsets[i]:self.b_offsets[i] + m] = torch.as_tensor(
                np.asarray(rhs[i], dtype=np.float64).reshape(m)).to("cuda")
        return A, b

    def solve(self, A, b, x=None):
        """Solve every problem; returns (x buffer, list of info dicts)."""
        import torch
        if x is None:
            x = torch.empty(self.x_elems, dtype=torch.float64, device="cuda")
        info = (LsInfo * len(self.shapes))()
        self.ctx.check(lib().randnla_batch_ls_execute(self.h, _ptr(A), _ptr(b), _ptr(x), info))
        return x, [info[i].as_dict() for i in range(len(self.shapes))]

    def unpack(self, x):
        xh = x.cpu().numpy() if hasattr(x, "cpu") else np.asarray(x)
        return [xh[o:o + n].copy() for o, (_, n) in zip(self.x_offsets, self.shapes)]

    def sketch(self):
        """[S_i A_i | S_i b_i] of the last solve, one (d_i, n_i + 1) array each."""
        out = np.empty(self._sketch_elems(), dtype=np.float32)
        self.ctx.check(lib().randnla_batch_ls_sketch(self.h, out.ctypes.data, out.size))
        res, o = [], 0
        for (m, n), d in zip(self.shapes, self._ds):
            res.append(out[o:o + d * (n + 1)].reshape(n + 1, d).T.copy())
            o += d * (n + 1)
        return res

    def _sketch_elems(self):
        return sum(d * (n + 1) for (_, n), d in zip(self.shapes, self._ds))

    def config(self):
        cap, wpc, best = C.c_int(), C.c_int(), C.c_double()
        self.ctx.check(lib().randnla_batch_ls_config(self.h, C.byref(cap), C.byref(wpc),
                                                     C.byref(best)))
        return {"row_step_shift": cap.value, "wgs_per_cu": wpc.value, "best_ms": best.value}

    def close(self):
        if getattr(self, "h", None):
            _PLANS.pop(self.h.value, None)
            lib().randnla_batch_ls_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_ls(mats, rhs, sampling_factor=2.0, tol=1e-10, iter_lim=None, rng=None):
    """Solve min ||A_i x_i - b_i|| for every i in one batched launch; returns
    (list of x_i, list of info dicts)."""
    shapes = [np.shape(a) for a in mats]
    plan = BatchLS(shapes, sampling_factor, tol, iter_lim, rng)
    try:
        A, b = plan.pack(mats, rhs)
        x, info = plan.solve(A, b)
        return plan.unpack(x), info
    finally:
        plan.close()


# ------------------------------------------------ learned (k, p, q) proposer
def mlp_param_count():
    return lib().randnla_mlp_param_count()


def _rowmajor_f32(x):
    """Sample-major (C-order) FP32 buffer on the same side as x."""
    if _is_torch(x):
        import torch
        return x.to(torch.float32).contiguous()
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


def mlp_forward(params, X, ctx=None):
    """Y (B x 3) of the proposer network (params FP32, X B x 34, sample-major)."""
    ctx = ctx or context()
    p, x = _rowmajor_f32(params), _rowmajor_f32(X)
    B = x.shape[0]
    if _is_torch(x):
        import torch
        Y = torch.empty((B, 3), dtype=torch.float32, device=x.device)
    else:
        Y = np.zeros((B, 3), dtype=np.float32)
    ctx.check(lib().randnla_mlp_forward(ctx.h, _ptr(p), _ptr(x), B, _ptr(Y)))
    return Y


def mlp_grad(params, X, T, wo, ctx=None):
    """(loss, gradient) of sum_o wo[o] (Y - T)^2 / B."""
    ctx = ctx or context()
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
    x = np.ascontiguousarray(np.asarray(X, dtype=np.float32))
    t = np.ascontiguousarray(np.asarray(T, dtype=np.float32))
    w = np.ascontiguousarray(np.asarray(wo, dtype=np.float32))
    g = np.zeros(mlp_param_count(), dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    ctx.check(lib().randnla_mlp_grad(ctx.h, p.ctypes.data, x.ctypes.data, t.ctypes.data,
                                     w.ctypes.data, x.shape[0], g.ctypes.data, loss.ctypes.data))
    return float(loss[0]), g


def mlp_train_step(params, m, v, X, T, wo, step, lr, ctx=None, want_loss=True):
    """One Adam step on device tensors; returns the minibatch loss."""
    ctx = ctx or context()
    loss = np.zeros(1, dtype=np.float32)
    ctx.check(lib().randnla_mlp_train_step(ctx.h, _ptr(params), _ptr(m), _ptr(v), _ptr(X),
                                           _ptr(T), _ptr(wo), X.shape[0], int(step), float(lr),
                                           loss.ctypes.data if want_loss else None))
    return float(loss[0])


def probe_features(A, rng=None, ctx=None):
    """The proposer's 34 inputs for A (see randnla_probe_features)."""
    ctx = ctx or context()
    dt = _dtype_of(A)
    a = _Arr(A, np.float32 if dt == F32 else np.float64)
    m, n = A.shape
    out = np.zeros(34, dtype=np.float32)
    ctx.check(lib().randnla_probe_features(ctx.h, a.ptr, dt, m, n, a.ld,
                                           C.byref(rng) if rng is not None else None,
                                           out.ctypes.data))
    return out
