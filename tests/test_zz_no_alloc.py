"""GPU: randnla_sketch_apply_noalloc (no counterpart in the reference; for
callers that cannot allocate) against randnla_sketch_apply, through the C
ABI.  Each case makes two calls of the no-allocation entry point in a row,
the second on a smaller input, and compares both with the allocating entry
point on the same input to the FP64 sketch-product tolerance (the sparse-sign
sums are LDS atomics, so their order, and the last bits, vary from call to
call); plus the entry point's error paths."""
import ctypes as C

import numpy as np
import pytest

import randnla_amd as L

pytestmark = pytest.mark.gpu

# (side, dist, zeta, m, n, d); pairs run big then small
CASES = [
    (L.LEFT, L.SPARSE_SIGN, 8, 131072, 512, 2048),   # the LS configuration's sketch
    (L.LEFT, L.SPARSE_SIGN, 5, 20000, 100, 256),
    (L.LEFT, L.GAUSSIAN, 8, 16384, 256, 1024),
    (L.LEFT, L.UNIFORM, 8, 3000, 40, 96),
    (L.RIGHT, L.GAUSSIAN, 8, 4000, 300, 64),
    (L.RIGHT, L.RADEMACHER, 8, 1000, 50, 16),
]


def _dev(m, n, seed, f32=False):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    t = torch.randn((n, m), dtype=torch.float64, device="cuda", generator=g).t()
    return t.float() if f32 else t


@pytest.mark.parametrize("pair", [(0, 1), (2, 3), (4, 5)])
@pytest.mark.parametrize("f32", [False, True])
def test_noalloc_matches_allocating_entry(pair, f32):
    import torch
    rng = L.rng(L.PHILOX4X32, (3, 4))
    for idx in pair:
        side, dist, zeta, m, n, d = CASES[idx]
        A = _dev(m, n, 10 * idx + int(f32), f32)
        if f32:
            A = A.t().contiguous().t()
        want = L.sketch_apply(side, dist, A, d, rng=rng, zeta=zeta)
        got = L.sketch_apply_noalloc(side, dist, A, d, rng=rng, zeta=zeta)
        torch.cuda.synchronize()
        assert got.shape == want.shape
        scale = float(want.abs().max())
        err = float((got - want).abs().max())
        assert err <= 1e-12 * scale, (idx, f32, err, scale)


def test_noalloc_error_paths():
    import torch
    lib, ctx = L.lib(), L.context()
    rng = L.rng(L.PHILOX4X32, (3, 4))
    # host memory is refused (nothing is staged)
    a = np.asfortranarray(np.ones((64, 8)))
    o = np.zeros((16, 8), order="F")
    st = lib.randnla_sketch_apply_noalloc(ctx.h, L.LEFT, L.GAUSSIAN, 8, C.byref(rng),
                                          a.ctypes.data, L.F64, 64, 8, 64, 16, o.ctypes.data, 16)
    assert L._KINDS[st] == "InvalidParameters"
    # d * n beyond the fixed array
    A = _dev(16, 1000, 1)
    with pytest.raises(L.RandNLAError) as e:
        L.sketch_apply_noalloc(L.LEFT, L.GAUSSIAN, A, 9000, rng=rng)
    assert e.value.kind == "InvalidDimensions"
    # A * Omega with a sparse-sign Omega would need an allocated dense Omega
    with pytest.raises(L.RandNLAError) as e:
        L.sketch_apply_noalloc(L.RIGHT, L.SPARSE_SIGN, _dev(100, 20, 2), 8, rng=rng)
    assert e.value.kind == "InvalidParameters"
    torch.cuda.synchronize()
