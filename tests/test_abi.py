"""CPU: the C-ABI library loads and exports every entry point that
include/randnla.h declares; no compute calls (no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "randnla.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(randnla_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("randnla_rand_svd", "randnla_sketching_operator", "randnla_blendenpik",
                 "randnla_lsqr", "randnla_cqrrpt", "randnla_qrcp", "randnla_haar_sample"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import randnla_amd as L
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import randnla_amd as L
    assert set(_declared()) <= set(L.EXPORTED)


def test_no_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import randnla_amd as L
    with pytest.raises(L.RandNLAError) as e:
        L.Context(0)
    assert e.value.kind == "NoDevice"


def test_opts_default_host():
    import randnla_amd as L
    o = L._opts()
    assert o.precision == L.FP64 and o.power_iters == 1 and o.zeta == 8
    r = L.rng_from_seed(L.PHILOX4X32, 42)
    assert (o.rng.key[0], o.rng.key[1]) == (r.key[0], r.key[1])
