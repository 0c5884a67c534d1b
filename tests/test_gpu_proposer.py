"""GPU parity for the learned (k, p, q) proposer: the network's forward and
backward kernels against the CPU oracle (oracle/mlp.py, itself pinned by
finite differences), training on the library's Adam kernel, the probe
features against numpy, and the A/B test against plain rand_svd."""
import os

import numpy as np
import pytest

import randnla_amd as L
from oracle import mlp, rng as R

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _params(seed, scale=0.05):
    r = np.random.default_rng(seed)
    return (mlp.init_params(seed) + scale * r.standard_normal(mlp.NPARAM)).astype(np.float32)


def test_mlp_forward_matches_oracle():
    p = _params(1)
    X = np.random.default_rng(2).standard_normal((70, 34)).astype(np.float32)   # 3 groups, ragged
    y = np.asarray(L.mlp_forward(p, X))
    yo, _ = mlp.forward(p.astype(np.float64), X.astype(np.float64), emulate=True)
    assert y.shape == (70, 3)
    # same BF16 rounding points; only FP32 vs FP64 summation order differs
    assert np.linalg.norm(y - yo) <= 1e-3 * np.linalg.norm(yo)
    # and the BF16 network tracks the exact FP64 one
    ye, _ = mlp.forward(p.astype(np.float64), X.astype(np.float64), emulate=False)
    assert np.linalg.norm(y - ye) <= 3e-2 * np.linalg.norm(ye)


def test_mlp_gradient_matches_oracle():
    p = _params(3)
    r = np.random.default_rng(4)
    X = r.standard_normal((45, 34)).astype(np.float32)
    T = r.standard_normal((45, 3)).astype(np.float32)
    wo = np.array([1.0, 0.25, 1.0], dtype=np.float32)
    loss, g = L.mlp_grad(p, X, T, wo)
    lo, go = mlp.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64),
                           wo.astype(np.float64), emulate=True)
    assert abs(loss - lo) <= 1e-4 * lo
    assert np.linalg.norm(g - go) <= 5e-3 * np.linalg.norm(go)
    # each parameter block separately (a transposed or misplaced block fails here)
    o = 0
    for name, shape in mlp.SIZES:
        k = int(np.prod(shape))
        gb, ob = g[o:o + k], go[o:o + k]
        assert np.linalg.norm(gb - ob) <= 1e-2 * np.linalg.norm(ob) + 1e-9, name
        o += k


def test_mlp_training_fits_a_teacher():
    teacher = _params(5, scale=0.2)
    r = np.random.default_rng(6)
    F = r.standard_normal((512, 34)).astype(np.float32)
    T = np.asarray(L.mlp_forward(teacher, F))
    from randnla_amd import proposer
    prop = proposer.Proposer(seed=7)
    hist = prop.fit(F, T, epochs=60, batch=64, lr=3e-3)
    assert hist[-1] < 0.1 * hist[0], (hist[0], hist[-1])
    # the on-device weights moved and the forward pass agrees with the oracle
    p = prop.params()
    assert np.isfinite(p).all()
    y = np.asarray(L.mlp_forward(p, F[:64]))
    yo, _ = mlp.forward(p.astype(np.float64), F[:64].astype(np.float64), emulate=True)
    assert np.linalg.norm(y - yo) <= 1e-3 * np.linalg.norm(yo)


def test_probe_features_match_numpy():
    m, n = 1024, 256
    sig = 0.8 ** np.arange(n)
    A = L.test_matrix(m, n, sig, rng=L.rng(L.PHILOX4X32, (3, 4)), dtype=L.F64, device=False)
    f = L.probe_features(A)
    om = R.gaussian(R.Stream.from_seed(R.PHILOX4X32, 42), n * 48).reshape(48, n).T
    s = np.linalg.svd(A @ om, compute_uv=False)
    want = np.log10(np.maximum(s[:32] / s[0], 1e-12)) / 12.0
    np.testing.assert_allclose(f[:32], want, atol=1e-6)
    np.testing.assert_allclose(f[32:], [np.log(m) / 10, np.log(n) / 10], rtol=1e-6)
    # scale invariance and the rank-1 closed form
    np.testing.assert_allclose(L.probe_features(7.5 * A), f, atol=1e-6)
    u = np.random.default_rng(1).standard_normal(m)
    v = np.random.default_rng(2).standard_normal(n)
    f1 = L.probe_features(np.outer(u, v))
    assert f1[0] == 0.0 and np.all(f1[1:32] <= -0.9)


@pytest.mark.skipif(not os.path.exists(os.path.join(HERE, "golden", "proposer_dataset.npz")),
                    reason="training set not generated")
def test_proposer_beats_plain_rand_svd():
    from randnla_amd import proposer
    d = np.load(os.path.join(HERE, "golden", "proposer_dataset.npz"))
    ok = d["meta"][:, 3].astype(bool)
    F, T = d["features"][ok], d["targets"][ok]
    prop = proposer.Proposer(seed=0)
    hist = prop.fit(F, T, epochs=300, batch=64, lr=2e-3)
    assert hist[-1] < 0.25 * hist[0]
    # held-out matrices: fresh seeds, same families and sizes
    rng = np.random.default_rng(2024)
    res = []
    for i in range(16):
        m, n = proposer.SIZES[int(rng.integers(len(proposer.SIZES)))]
        sig, noise = proposer.spectrum(proposer.FAMILIES[i % 4], n, rng)
        if np.sum(sig > proposer.TOL) + 32 > min(m, n) // 2:
            continue                                  # infeasible for both methods
        A = proposer.make_matrix(m, n, sig, 900000 + i, noise)
        a = prop.solve(A)
        b = proposer.plain(A)
        res.append((a, b))
        # both meet the target under the same a-posteriori estimator
        assert a["err"] <= proposer.TOL and b["err"] <= proposer.TOL, (i, a, b)
    assert len(res) >= 10
    gain = sum(b["cost"] for _, b in res) / sum(a["cost"] for a, _ in res)
    print(f"proposer gain (plain flops / proposer flops) over {len(res)} matrices: {gain:.2f}")
    assert gain > 1.2
