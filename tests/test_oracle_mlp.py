"""The proposer network's CPU oracle, pinned by finite differences (the FP64
function) and by the algebra of its rounding emulation."""
import numpy as np

from oracle import mlp


def test_parameter_count_matches_architecture():
    # 34 -> 128 -> 128 -> 3 with biases
    assert mlp.NPARAM == 34 * 128 + 128 + 128 * 128 + 128 + 3 * 128 + 3 == 21379


def test_gelu_derivative_by_finite_differences():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mlp.gelu_d(z), fd, rtol=0, atol=1e-8)
    # closed forms: GELU(0) = 0, GELU(z) -> z for large z, -> 0 for very negative z
    assert mlp.gelu(np.array(0.0)) == 0.0
    np.testing.assert_allclose(mlp.gelu(np.array([10.0, -10.0])), [10.0, 0.0], atol=1e-12)


def test_fp64_gradient_by_finite_differences():
    r = np.random.default_rng(0)
    p = mlp.init_params(1) + 0.05 * r.standard_normal(mlp.NPARAM)
    X = r.standard_normal((7, 34))
    T = r.standard_normal((7, 3))
    wo = np.array([1.0, 0.5, 2.0])
    _, g = mlp.loss_grad(p, X, T, wo, emulate=False)
    idx = np.r_[r.choice(mlp.NPARAM, 60, replace=False), [0, 4351, 4352, 21378, 21000]]
    h = 1e-6
    for i in idx:
        e = np.zeros(mlp.NPARAM)
        e[i] = h
        fp = mlp.loss_grad(p + e, X, T, wo, emulate=False)[0]
        fm = mlp.loss_grad(p - e, X, T, wo, emulate=False)[0]
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even (1)
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even (1 + 2^-6)
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5])
    np.testing.assert_array_equal(mlp.bf16(x), [1.0, 1.0, 1.0 + 2.0 ** -6, -2.5])
    v = np.random.default_rng(1).standard_normal(1000)
    b = mlp.bf16(v)
    assert np.all(np.abs(b - v) <= 2.0 ** -8 * np.abs(v))     # half an ulp of an 8-bit significand
    np.testing.assert_array_equal(mlp.bf16(b), b)            # idempotent


def test_emulated_gradient_close_to_exact():
    r = np.random.default_rng(2)
    p = mlp.init_params(3)
    X = r.standard_normal((64, 34))
    T = r.standard_normal((64, 3))
    wo = np.ones(3)
    l0, g0 = mlp.loss_grad(p, X, T, wo, emulate=False)
    l1, g1 = mlp.loss_grad(p, X, T, wo, emulate=True)
    assert abs(l1 - l0) <= 2e-2 * l0
    assert np.linalg.norm(g1 - g0) <= 5e-2 * np.linalg.norm(g0)
