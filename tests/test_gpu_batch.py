"""GPU parity: batched uneven least squares (one persistent launch) against
the per-problem oracle (blendenpik_overdetermined with LSQR,
src/sketch_and_precondition.rs:26-59 + src/solvers.rs:115-278), the unique
least-squares solution, and the batched path's own invariants."""
import numpy as np
import pytest

import randnla_amd as L
from oracle import la, rng as R

pytestmark = pytest.mark.gpu

KEY = R.Stream.from_seed(R.PHILOX4X32, 42)


def problem(rng, m, n, kappa=1e4):
    """Gaussian A with column scaling of condition number kappa, rounded to
    FP32 (the stored precision); b standard normal."""
    a = rng.standard_normal((m, n)) * 10.0 ** (np.log10(kappa) * rng.random(n))
    return a.astype(np.float32).astype(np.float64), rng.standard_normal(m)


SHAPES = [(64, 16), (300, 1), (1000, 40), (257, 64), (2048, 256), (20, 16), (999, 7),
          (4096, 130)]


def d_of(m, n, sf=2.0):
    return m if sf * n > m else int(np.floor(sf * n))


def test_batch_matches_oracle_and_lstsq():
    rng = np.random.default_rng(7)
    probs = [problem(rng, m, n) for m, n in SHAPES]
    xs, info = L.batch_ls([a for a, _ in probs], [b for _, b in probs], 2.0, 1e-12, 400)
    for i, ((a, b), x, inf) in enumerate(zip(probs, xs, info)):
        m, n = a.shape
        assert inf["status"] == 0 and inf["d"] == d_of(m, n)
        x_ls = np.linalg.lstsq(a, b, rcond=None)[0]
        # kappa(A) = 1e4 and LSQR to 1e-12 on the preconditioned operator:
        # the forward error stays below 1e-7 relative.
        assert np.linalg.norm(x - x_ls) <= 1e-7 * np.linalg.norm(x_ls), i
        S = R.batch_gauss_sketch(KEY, i, d_of(m, n), m)
        xo, istop, itn = la.blendenpik_lsqr(a, b, 1e-12, 400, 2.0, S)
        assert inf["istop"] == istop, (i, inf, istop)
        # the FP32 sketch perturbs R at the 1e-6 level: iteration counts agree
        # to within two
        assert abs(inf["itn"] - itn) <= 2, (i, inf["itn"], itn)
        assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)


def test_batch_sketch_matches_oracle():
    rng = np.random.default_rng(8)
    shapes = [(130, 9), (64, 32), (515, 100), (1031, 256)]
    probs = [problem(rng, m, n, kappa=1.0) for m, n in shapes]
    plan = L.BatchLS(shapes, 2.0, 1e-10, 50)
    try:
        A, b = plan.pack([a for a, _ in probs], [b for _, b in probs])
        plan.solve(A, b)
        sk = plan.sketch()
    finally:
        plan.close()
    for i, ((a, b), s) in enumerate(zip(probs, sk)):
        m, n = a.shape
        S = R.batch_gauss_sketch(KEY, i, d_of(m, n), m)
        want = S @ np.c_[a, b]
        assert s.shape == want.shape
        # FP32 products over m terms and FP32 transcendental evaluation
        err = np.abs(s - want).max(axis=0) / np.linalg.norm(np.c_[a, b], axis=0)
        assert err.max() <= 2e-5, (i, err.max())


def test_batch_equals_single_problem_bitwise():
    rng = np.random.default_rng(9)
    shapes = [(800, 50), (300, 20), (5000, 200)]
    probs = [problem(rng, m, n) for m, n in shapes]
    xs, info = L.batch_ls([a for a, _ in probs], [b for _, b in probs], 2.0, 1e-10, 200)
    for i, (a, b) in enumerate(probs):
        plan = L.BatchLS([a.shape], 2.0, 1e-10, 200, stream_ids=[i])
        try:
            A, bb = plan.pack([a], [b])
            x1, inf1 = plan.solve(A, bb)
            x1 = plan.unpack(x1)[0]
        finally:
            plan.close()
        assert np.array_equal(x1, xs[i]), i
        assert inf1[0]["itn"] == info[i]["itn"]


def test_batch_graph_replay_and_autotune():
    import torch
    rng = np.random.default_rng(10)
    shapes = [(int(256 * 8 ** rng.random()), int(rng.integers(16, 64))) for _ in range(40)]
    shapes = [(max(m, 4 * n), n) for m, n in shapes]
    probs = [problem(rng, m, n) for m, n in shapes]
    plan = L.BatchLS(shapes, 2.0, 1e-10, 200, autotune=True)
    try:
        A, b = plan.pack([a for a, _ in probs], [b for _, b in probs])
        x0, info0 = plan.solve(A, b)           # tunes, then solves with the winner
        cfg = plan.config()
        assert cfg["best_ms"] > 0 and cfg["row_step_shift"] in (0, 1) and cfg["wgs_per_cu"] >= 1
        ref = x0.clone()
        for _ in range(3):
            torch.randn(1000, 1000, device="cuda").sum()   # eager work between replays
            torch.cuda.synchronize()
            x1, info1 = plan.solve(A, b)
            assert torch.equal(x1, ref)
            assert [i["itn"] for i in info1] == [i["itn"] for i in info0]
        xs = plan.unpack(ref)
    finally:
        plan.close()
    for (a, bvec), x in zip(probs, xs):
        x_ls = np.linalg.lstsq(a, bvec, rcond=None)[0]
        assert np.linalg.norm(x - x_ls) <= 1e-6 * np.linalg.norm(x_ls)


def test_batch_rank_deficient_problem_flagged():
    rng = np.random.default_rng(11)
    a0, b0 = problem(rng, 200, 10)
    a1, b1 = problem(rng, 300, 12)
    a1[:, 5] = 0.0
    xs, info = L.batch_ls([a0, a1], [b0, b1], 2.0, 1e-10, 100)
    assert info[0]["status"] == 0 and info[1]["status"] == 1
    assert not np.any(xs[1])
    np.testing.assert_allclose(xs[0], np.linalg.lstsq(a0, b0, rcond=None)[0], rtol=1e-7)


def test_batch_errors():
    with pytest.raises(L.RandNLAError) as e:
        L.BatchLS([(1000, 300)])
    assert e.value.kind == "InvalidDimensions"
    with pytest.raises(L.RandNLAError) as e:
        L.BatchLS([(10, 20)])
    assert e.value.kind == "NotOverdetermined"
    with pytest.raises(L.RandNLAError) as e:
        L.BatchLS([(100, 10)], sampling_factor=0.5)
    assert e.value.kind == "InvalidParameters"
    plan = L.BatchLS([(100, 10)])
    try:
        with pytest.raises(L.RandNLAError) as e:
            plan.solve(np.zeros(1000, np.float32), np.zeros(100), np.zeros(10))
        assert e.value.kind == "InvalidParameters"
    finally:
        plan.close()
