"""Benchmark: mixed-precision rand_svd vs the FP64 path on one MI355X.

Workload (default): A is 65536 x 16384 stored FP32, sigma_i = (1+i)^-2 plus
1e-7 Gaussian noise, k = 256, p = 64, q = 3.  One step = one complete
rand_svd call (BF16 range finder, split-FP16 B = Q^T A, FP64 CholeskyQR2 and
small SVD, U and V^T formed), inputs resident in HBM.  The FP64 path runs on
the same FP32 bits for the accuracy and speed comparison.

Prints one JSON line (see the repository's measurement notes).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


# HBM bytes per launch of the BF16 NN GEMM (full-tile variant) at the default
# shape (FETCH_SIZE x 2.0 + WRITE_SIZE, scaled against a 1 GiB copy), from the
# two --pmc runs summarised in profiles/rand_svd_mixed_65536x16384_traffic.md.
TRAFFIC_BF16_NN = (2231.59 + 84.28) * 1e6
# HBM bytes per launch of the clean-round split-FP16 GEMM (mode 2, A^T X form)
# averaged over its launches at the default shape, from the FETCH_SIZE and
# WRITE_SIZE runs summarised in profiles/rand_svd_mixed_65536x16384_traffic.md
# (same calibration).
TRAFFIC_F16X2 = (4559.83 + 213.35) * 1e6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--m", type=int, default=65536)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--k", type=int, default=256)
    ap.add_argument("--p", type=int, default=64)
    ap.add_argument("--q", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--fp8-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--err-iters", type=int, default=64,
                    help="Lanczos steps of the spectral-error estimate")
    ap.add_argument("--only", default="all",
                    help="comma list of workloads to run (profiling): svd (the headline), c1, "
                         "ls, cqrrpt, sls, evd, id, batch, mlp; default all")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device-to-device 3 times (HBM counter calibration)")
    return ap.parse_args()


def copy_probe():
    """A plain 1 GiB device copy: its FETCH_SIZE / WRITE_SIZE per dispatch
    against the 1 GiB it moves calibrate the HBM traffic counters."""
    import torch
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": src.numel() * 4}))


def bench_ls(L, torch, ctx, cpu=True):
    """Overdetermined LS, 131072 x 512 FP64, sigma log-uniform over kappa 1e8,
    b = A x* + 1e-6 N(0,1); sparse-sign sketch d = 4n, zeta = 8; LSQR 1e-12."""
    m, n = 131072, 512
    sig = 10.0 ** (-8.0 * np.arange(n) / (n - 1))
    A = L.test_matrix(m, n, sig, rng=L.rng(L.PHILOX4X32, (77, 1)), dtype=L.F64)
    g = torch.Generator(device="cuda").manual_seed(5)
    xs = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    b = A @ xs + 1e-6 * torch.randn(m, dtype=torch.float64, device="cuda", generator=g)
    torch.cuda.synchronize()

    def solve():
        return L.blendenpik_overdetermined(A, b, 1e-12, 400, 4.0, sketch=L.SPARSE_SIGN, zeta=8,
                                           solver=L.LSQR, return_info=True, ctx=ctx)
    solve()
    torch.cuda.synchronize()
    ctx.timing(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    x, info = solve()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    gemv_ms, gemv_n = ctx.timer("gemv")
    pass_ms, _ = ctx.timer("lsqr_pass")
    fused = n <= 512
    sk_ms, _ = ctx.timer("sketch_apply")
    hh_ms, _ = ctx.timer("householder")
    ctx.timing(False)
    r = A @ x - b
    it = info["itn"]
    # LSQR reads A twice per iteration (src/solvers.rs:194, :201); the fused
    # pass (n <= 512) reads it once, plus the initial A^T u
    lsqr_ms = gemv_ms + pass_ms
    lsqr_bytes = (it + 1) * m * n * 8.0 if fused else (it + 1) * 2.0 * m * n * 8
    # the least-squares solution of the same system (FP64 QR on the GPU)
    xls = torch.linalg.lstsq(A, b.reshape(-1, 1)).solution.reshape(-1)
    d = 4 * n
    sk_bytes = m * n * 8.0 + m * 8.0 + d * (n + 1) * 8.0   # A and b read once, [SA | Sb] written
    out = {"config": f"blendenpik + LSQR, A {m}x{n} FP64, kappa 1e8, sparse-sign d={d} zeta=8, tol 1e-12",
           "ms": dt * 1e3, "itn": it, "istop": info["istop"],
           "rel_resid_normal_eq": float((A.t() @ r).norm() / (A.norm() * r.norm())),
           "rel_err_vs_lstsq": float((x - xls).norm() / xls.norm()),
           "resid_over_lstsq_resid": float(r.norm() / (A @ xls - b).norm()),
           "rel_err_vs_xstar": float((x - xs).norm() / xs.norm()),
           "breakdown_ms": {"sketch_apply": sk_ms, "householder": hh_ms, "lsqr": lsqr_ms},
           "sketch_roofline": {"bound": "hbm", "achieved": sk_bytes / (sk_ms * 1e-3) / 1e9,
                               "peak": 8000.0, "unit": "GB/s",
                               "frac": sk_bytes / (sk_ms * 1e-3) / 8e12,
                               "bytes": sk_bytes} if sk_ms > 0 else None,
           "lsqr_passes_over_A_per_iteration": 1 if fused else 2,
           "lsqr_hbm_GBps": lsqr_bytes / (lsqr_ms * 1e-3) / 1e9 if lsqr_ms > 0 else None,
           "lsqr_hbm_frac": lsqr_bytes / (lsqr_ms * 1e-3) / 8e12 if lsqr_ms > 0 else None}
    if cpu:
        # The reference's path (src/sketch_and_precondition.rs:26-59: dense
        # Gaussian sketch, QR, CGLS on A R^-1) in the oracle, and the intended
        # one (sparse-sign sketch, LSQR), on the first m_s rows of the same
        # construction; S is given (its generation is not timed).
        from oracle import la
        ms_ = 16384
        rs = np.random.default_rng(7)
        u, _ = np.linalg.qr(rs.standard_normal((ms_, n)))
        v, _ = np.linalg.qr(rs.standard_normal((n, n)))
        a = (u * sig) @ v.T
        bb = a @ rs.standard_normal(n) + 1e-6 * rs.standard_normal(ms_)
        s_dense = rs.standard_normal((d, ms_))
        from oracle import rng as R
        s_sparse = R.sketching_operator(3, d, ms_, R.Stream.from_seed(R.PHILOX4X32, 3), zeta=8)
        allc = _cpu_threads()
        t_ref = _timed(lambda: la.blendenpik_overdetermined(a, bb, 1e-12, 400, 4.0, s_op=s_dense))
        t_ref1 = _timed(lambda: la.blendenpik_overdetermined(a, bb, 1e-12, 400, 4.0,
                                                             s_op=s_dense), 1)
        t_int = _timed(lambda: la.blendenpik_lsqr(a, bb, 1e-12, 400, 4.0, s_sparse))
        scale = m / ms_
        out["cpu_baseline"] = {
            "value": t_ref * scale * 1e3, "unit": "ms (extrapolated)", "cores": allc,
            "kind": "port", "cpu": _cpu_model(),
            "sample": f"oracle blendenpik_overdetermined (dense Gaussian S given, QR, CGLS; numpy "
                      f"FP64) on A {ms_}x{n} with the same spectrum, time x {scale:g} (rows)",
            "sample_ms_all_cores": t_ref * 1e3, "sample_ms_single_thread": t_ref1 * 1e3,
            "sample_ms_intended_sparse_lsqr_all_cores": t_int * 1e3}
        out["gpu_over_cpu"] = t_ref * scale / dt
    return out


def _cpu_threads():
    try:
        from threadpoolctl import threadpool_info
        return int(max((d.get("num_threads", 1) for d in threadpool_info()), default=1))
    except Exception:
        return os.cpu_count() or 1


def _timed(fn, threads=None):
    """Wall time of fn() with the BLAS pools limited to `threads` (None: as
    configured)."""
    from threadpoolctl import threadpool_limits
    if threads is None:
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_c1(L, torch, ctx, cpu=True):
    """The plain rand_svd configuration in FP64: A 2000 x 1000 = U diag(sigma)
    V^T (U, V from QR of seeded Gaussians), sigma_i = 10^(-i/10) for i < 50
    then 1e-8, k = 50, p = 10, q = 2, Gaussian Omega from Philox4x32
    seed_from_u64(42).  GPU: the intended method and the reference's own
    computation (faithful mode); CPU: the oracle's restatement of both, on one
    core (the reference's matrixmultiply is single-threaded) and on all."""
    from oracle import la, rng as R
    m, n, k, p, q = 2000, 1000, 50, 10, 2
    sig = np.r_[10.0 ** (-np.arange(50) / 10.0), np.full(n - 50, 1e-8)]
    a = la.lowrank_matrix(m, n, sig, R.Stream.from_seed(R.PHILOX4X32, 1),
                          R.Stream.from_seed(R.PHILOX4X32, 2))
    A = torch.as_tensor(np.asfortranarray(a).T.copy(), device="cuda").t()
    torch.cuda.synchronize()
    out = {"config": f"rand_svd FP64, A {m}x{n}, sigma_i=10^(-i/10) (i<50) then 1e-8, "
                     f"k={k}, p={p}, q={q}"}
    flops = (2 * q + 2) * 2.0 * m * n * (k + p)
    for name, kw in (("gpu_fp64", dict(power_iters=q)), ("gpu_fp64_faithful", dict(faithful=True))):
        fn = lambda: L.rand_svd_factors(A, k, 1e-6, p, precision=L.FP64, ctx=ctx, **kw)
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            U, sv, Vt = fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        err = float(torch.linalg.matrix_norm(A - (U * sv) @ Vt, ord=2))
        out[name] = {"ms": dt * 1e3, "rel_spectral_err": err / sig[0],
                     "GFLOP/s_intended_method": flops / dt / 1e9}
    if cpu:
        om = R.gaussian(R.Stream.from_seed(R.PHILOX4X32, 42), n * (k + p)).reshape(k + p, n).T
        allc = _cpu_threads()
        rec = {"kind": "port", "cpu": _cpu_model(), "threads_all": allc}
        for tag, th in (("single_thread", 1), ("all_cores", allc)):
            t_int = min(_timed(lambda: la.rand_svd_power(a, k, p, q, om), th) for _ in range(3))
            t_fai = min(_timed(lambda: la.rand_svd(a, k, 1e-6, p), th) for _ in range(3))
            rec[tag] = {"intended_ms": t_int * 1e3, "faithful_ms": t_fai * 1e3}
        out["cpu_baseline"] = rec
        out["gpu_over_cpu_single_thread_faithful"] = (rec["single_thread"]["faithful_ms"]
                                                      / out["gpu_fp64_faithful"]["ms"])
        out["gpu_over_cpu_all_cores_intended"] = (rec["all_cores"]["intended_ms"]
                                                  / out["gpu_fp64"]["ms"])
    del A
    return out


def bench_cqrrpt(L, torch, ctx, cpu=True):
    """CQRRPT (src/cqrrpt.rs:27-58) at the least-squares shape: A 131072 x 512
    FP64 (the same kappa-1e8 construction), d = 2048: the dense Gaussian
    sketch S A (the reference's operator), QRCP of the 2048 x 512 sketch,
    A_pre = A[:, J] R_sk^-1, Cholesky of A_pre^T A_pre, Q = A_pre R_pre^-1.
    Also the sparse-sign sketch variant.  The reference's own harness times
    CQRRPT against QRCP (src/main.rs:22-79)."""
    m, n, d = 131072, 512, 2048
    sig = 10.0 ** (-8.0 * np.arange(n) / (n - 1))
    A = L.test_matrix(m, n, sig, rng=L.rng(L.PHILOX4X32, (77, 1)), dtype=L.F64)
    torch.cuda.synchronize()
    out = {"config": f"sap_chol_qrcp, A {m}x{n} FP64, kappa 1e8, d={d}"}
    for name, sk in (("gaussian", None), ("sparse_sign", L.SPARSE_SIGN)):
        fn = lambda: L.sap_chol_qrcp(A, d, sketch=sk, ctx=ctx)
        fn()
        torch.cuda.synchronize()
        ctx.timing(True)
        ctx.timing_reset()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            Q, R, J = fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        rec = {"ms": dt * 1e3, "rows_per_s": m / dt}
        for t in ("sketch_apply", "sketch_gemm", "householder", "trmm", "gram", "rprod"):
            ms_, cnt = ctx.timer(t)
            if cnt:
                rec[t + "_ms"] = ms_ / reps
        ctx.timing(False)
        # invariants of src/cqrrpt.rs:95-109: orthonormal Q, A[:, J] = Q R
        eye = torch.eye(n, dtype=Q.dtype, device=Q.device)
        rec["orth_err"] = float((Q.t() @ Q - eye).abs().max())
        rec["recon_rel_err"] = float((A[:, J] - Q @ R).norm() / A.norm())
        sk_ms = rec.get("sketch_apply_ms", 0.0) + rec.get("sketch_gemm_ms", 0.0)
        if sk_ms > 0 and sk is None:
            fl = 2.0 * d * m * n
            rec["sketch_roofline"] = {"bound": "mfma", "achieved": fl / (sk_ms * 1e-3) / 1e12,
                                      "peak": 78.6, "unit": "TFLOP/s",
                                      "frac": fl / (sk_ms * 1e-3) / 78.6e12}
        elif sk_ms > 0:
            by = m * n * 8.0 + d * n * 8.0
            rec["sketch_roofline"] = {"bound": "hbm", "achieved": by / (sk_ms * 1e-3) / 1e9,
                                      "peak": 8000.0, "unit": "GB/s",
                                      "frac": by / (sk_ms * 1e-3) / 8e12}
        out[name] = rec
        del Q, R, J
    del A
    torch.cuda.empty_cache()
    if cpu:
        from oracle import la
        ms_ = 16384
        rs = np.random.default_rng(8)
        u, _ = np.linalg.qr(rs.standard_normal((ms_, n)))
        v, _ = np.linalg.qr(rs.standard_normal((n, n)))
        a = (u * sig) @ v.T
        s_dense = rs.standard_normal((d, ms_))
        t = _timed(lambda: la.sap_chol_qrcp(a, d, s_op=s_dense))
        out["cpu_baseline"] = {"value": ms_ / t, "unit": "rows/s", "cores": _cpu_threads(),
                               "kind": "port", "cpu": _cpu_model(),
                               "sample": f"oracle sap_chol_qrcp (numpy FP64, Gaussian S given, "
                                         f"QRCP with reflector-accumulated Q) A {ms_}x{n}, d={d}: "
                                         f"{t * 1e3:.0f} ms"}
        out["gpu_over_cpu_rows_per_s"] = out["gaussian"]["rows_per_s"] / (ms_ / t)
    return out


def bench_sketch_solve(L, torch, ctx, cpu=True):
    """Sketch-and-solve (src/sketch_and_solve.rs:24-65), A 16384 x 256 FP64,
    kappa 1e4, the reference's Gaussian sketch with m / 4 = 4096 rows (fused
    generation in the FP64 GEMM), then QR (single launch, Sb appended) or the
    small SVD.  Quality: ||A x - b|| over the exact least-squares residual."""
    m, n = 16384, 256
    d = m // 4
    sig = 10.0 ** (-4.0 * np.arange(n) / (n - 1))
    A = L.test_matrix(m, n, sig, rng=L.rng(L.PHILOX4X32, (78, 1)), dtype=L.F64)
    g = torch.Generator(device="cuda").manual_seed(6)
    xs = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    b = A @ xs + 1e-3 * torch.randn(m, dtype=torch.float64, device="cuda", generator=g)
    xls = torch.linalg.lstsq(A, b.reshape(-1, 1)).solution.reshape(-1)
    r_ls = float((A @ xls - b).norm())
    out = {"config": f"sketch-and-solve, A {m}x{n} FP64, kappa 1e4, Gaussian sketch d={d}"}
    for variant in ("qr", "svd"):
        fn = getattr(L, "sketched_least_squares_" + variant)
        fn(A, b, ctx=ctx)
        torch.cuda.synchronize()
        ctx.timing(True)
        ctx.timing_reset()
        t0 = time.perf_counter()
        x = fn(A, b, ctx=ctx)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        sk_ms, _ = ctx.timer("sketch_apply")
        hh_ms, _ = ctx.timer("householder")
        ctx.timing(False)
        out[variant] = {"ms": dt * 1e3, "sketch_ms": sk_ms,
                        "sketch_TFLOPs": 2.0 * d * m * (n + 1) / (sk_ms * 1e-3) / 1e12 if sk_ms > 0 else None,
                        "householder_ms": hh_ms if variant == "qr" else None,
                        "resid_over_lstsq_resid": float((A @ x - b).norm()) / r_ls}
        if sk_ms > 0:
            out[variant]["sketch_frac_fp64_peak"] = out[variant]["sketch_TFLOPs"] / 78.6
    if cpu:
        from oracle import la
        an, bn = A.cpu().numpy(), b.cpu().numpy()
        sop = np.random.default_rng(0).standard_normal((d, m))
        t0 = time.perf_counter()
        la.sketched_least_squares_qr(an, bn, s_op=sop)
        t1 = time.perf_counter()
        la.sketched_least_squares_svd(an, bn, s_op=sop)
        t2 = time.perf_counter()
        out["cpu_baseline"] = {"kind": "port", "cores": _cpu_threads(),
                               "sample": f"oracle sketched_least_squares_qr / _svd (numpy FP64), "
                                         f"the same A {m}x{n}, S {d}x{m} given (not generated)",
                               "qr_ms": (t1 - t0) * 1e3, "svd_ms": (t2 - t1) * 1e3}
    return out


def bench_evd(L, torch, ctx, cpu=True):
    """rand_evd1 / rand_evd2 (src/lora_drivers.rs:87-224) on a symmetric
    positive definite A 16384 x 16384 FP64 with eigenvalues 10^(-i/64)
    floored at 1e-8 (A = B B^T, B a Hadamard test matrix with sigma =
    sqrt(lambda)), k = 64, s = 64.  rand_evd1 runs its range finder (one
    power round); rand_evd2 its positive semi-definiteness test (a Cholesky
    factorisation of A, n^3 / 3 flops), three tsog1 passes and the Nystrom
    factor.  Quality: max relative error of the top 32 eigenvalues."""
    n, k, s = 16384, 64, 64
    lam_true = np.maximum(10.0 ** (-np.arange(n) / 64.0), 1e-8)
    B = L.test_matrix(n, n, np.sqrt(lam_true), rng=L.rng(L.PHILOX4X32, (91, 3)), dtype=L.F64)
    A = B @ B.T
    del B
    A = 0.5 * (A + A.T)                 # exactly symmetric (rand_evd1 checks A == A^T)
    A = A.t().contiguous().t()          # column-major for the library
    torch.cuda.synchronize()
    out = {"config": f"symmetric PD A {n}x{n} FP64, lambda_i = 10^(-i/64) >= 1e-8, k={k}, s={s}"}
    for name, fn in (("rand_evd1", lambda: L.rand_evd1(A, k, 1e-6, s, ctx=ctx)),
                     ("rand_evd2", lambda: L.rand_evd2(A, k, s, ctx=ctx))):
        fn()
        torch.cuda.synchronize()
        ctx.timing(True)
        ctx.timing_reset()
        t0 = time.perf_counter()
        V, lam = fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rec = {"ms": dt * 1e3}
        for t in ("sketch_gemm", "pass_a", "pass_at", "pass_bt", "evd_y", "psd_check", "jacobi"):
            ms, cnt = ctx.timer(t)
            if cnt:
                rec[t + "_ms"] = ms
        if rec.get("psd_check_ms"):
            rec["psd_check_TFLOPs"] = n ** 3 / 3.0 / (rec["psd_check_ms"] * 1e-3) / 1e12
            rec["psd_check_frac_fp64_peak"] = rec["psd_check_TFLOPs"] / 78.6
        if rec.get("pass_a_ms"):
            # A X with X n x l (l = k + s): 2 n^2 l flops on FP64 MFMA
            rec["pass_a_frac_fp64_peak"] = 2.0 * n * n * (k + s) / (rec["pass_a_ms"] * 1e-3) / 78.6e12
        ctx.timing(False)
        lam = lam.cpu().numpy()
        rec["top32_max_rel_err"] = float(np.max(np.abs(lam[:32] / lam_true[:32] - 1.0)))
        rec["orth_err"] = float((V.T @ V - torch.eye(V.shape[1], dtype=V.dtype, device=V.device))
                                .abs().max())
        out[name] = rec
        del V
    del A
    torch.cuda.empty_cache()
    if cpu:
        # The oracle (numpy FP64, the reference's algebra with its faithful
        # tsog1 and eigenvalue PSD test) on a bounded sample: n = 4096.
        from oracle import la
        ns = 4096
        rs = np.random.default_rng(0)
        qm, _ = np.linalg.qr(rs.standard_normal((ns, ns)))
        a = (qm * np.maximum(10.0 ** (-np.arange(ns) / 64.0), 1e-8)) @ qm.T
        a = 0.5 * (a + a.T)
        t0 = time.perf_counter()
        la.rand_evd1(a, k, 1e-6, s)
        t1 = time.perf_counter()
        la.rand_evd2(a, k, s)
        t2 = time.perf_counter()
        out["cpu_baseline"] = {"kind": "port", "cores": _cpu_threads(),
                               "sample": f"oracle rand_evd1 / rand_evd2 (numpy FP64) n={ns}, "
                                         f"k={k}, s={s}",
                               "rand_evd1_ms": (t1 - t0) * 1e3, "rand_evd2_ms": (t2 - t1) * 1e3}
    return out


def bench_id(L, torch, ctx, cpu=True):
    """Interpolative decompositions (src/id.rs) on A 16384 x 4096 FP64 with
    sigma_i = 10^(-i/32) (a Hadamard test matrix), k = 256: cur_randomised
    (Gaussian 256 x 16384 sketch, 256 QRCP steps on the 256 x 4096 sketch, 256
    on A[:, J]^T, pinv of A[I, :]), two_sided_id_randomised, and the
    deterministic cur (256 QRCP steps on A itself).  Quality:
    ||A - A[:, J] U A[I, :]||_F / ||A||_F (and Z A[I, J] X)."""
    m, n, k = 16384, 4096, 256
    sig = 10.0 ** (-np.arange(n) / 32.0)
    A = L.test_matrix(m, n, sig, rng=L.rng(L.PHILOX4X32, (93, 5)), dtype=L.F64)
    torch.cuda.synchronize()
    an = float(A.norm())
    out = {"config": f"A {m}x{n} FP64, sigma_i = 10^(-i/32), k={k}"}
    for name, fn in (("cur_randomised", lambda: L.cur_randomised(A, k, ctx=ctx)),
                     ("two_sided_id_randomised", lambda: L.two_sided_id_randomised(A, k, ctx=ctx)),
                     ("cur", lambda: L.cur(A, k, ctx=ctx))):
        fn()
        torch.cuda.synchronize()
        ctx.timing(True)
        ctx.timing_reset()
        t0 = time.perf_counter()
        res = fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rec = {"ms": dt * 1e3}
        for t in ("sketch_apply", "qrcp", "jacobi"):
            ms, cnt = ctx.timer(t)
            if cnt:
                rec[t + "_ms"] = ms
        ctx.timing(False)
        if name.startswith("cur"):
            rec["pinv_jacobi_sweeps"] = ctx.diag(0)
            J, U, I = res
            approx = A[:, J] @ U @ A[I, :]
        else:
            Z, I, J, X = res
            approx = Z @ A[I][:, J] @ X
        rec["rel_err"] = float((A - approx).norm()) / an
        del approx, res
        out[name] = rec
    out["sigma_k_over_sigma_1"] = float(sig[k])
    sk = out["cur_randomised"].get("sketch_apply_ms")
    if sk:
        # S A with S a generated k x m Gaussian: 2 k m n flops on FP64 MFMA
        out["cur_randomised"]["sketch_frac_fp64_peak"] = 2.0 * k * m * n / (sk * 1e-3) / 78.6e12
    del A
    torch.cuda.empty_cache()
    if cpu:
        from oracle import la
        ms_, ns_, ks = 2048, 512, 64
        rs = np.random.default_rng(0)
        u, _ = np.linalg.qr(rs.standard_normal((ms_, ns_)))
        v, _ = np.linalg.qr(rs.standard_normal((ns_, ns_)))
        a = (u * 10.0 ** (-np.arange(ns_) / 32.0)) @ v.T
        t0 = time.perf_counter()
        la.cur(a, ks, osid=la.osid_randomised)
        t1 = time.perf_counter()
        la.cur(a, ks)
        t2 = time.perf_counter()
        out["cpu_baseline"] = {"kind": "port", "cores": _cpu_threads(),
                               "sample": f"oracle cur_randomised / cur (numpy FP64) "
                                         f"A {ms_}x{ns_}, k={ks}",
                               "cur_randomised_ms": (t1 - t0) * 1e3, "cur_ms": (t2 - t1) * 1e3}
    return out


def bench_batch(L, torch, ctx, count, cpu=True):
    """4096 uneven LS problems: m log-uniform in [256, 16384], n uniform in
    [16, 256], m >= 4n; A Gaussian with column scaling kappa 1e4 (FP32), b
    random; Gaussian sketch d = 2n; LSQR tol 1e-10; one persistent launch."""
    rs = np.random.default_rng(11)
    shapes = []
    while len(shapes) < count:
        m = int(256 * 64 ** rs.random())
        n = int(rs.integers(16, 257))
        if m >= 4 * n:
            shapes.append((m, n))
    plan = L.BatchLS(shapes, 2.0, 1e-10, 200, autotune=True, ctx=ctx)
    g = torch.Generator(device="cuda").manual_seed(3)
    A = torch.empty(plan.a_elems, dtype=torch.float32, device="cuda")
    for (m, n), o in zip(shapes, plan.a_offsets):
        blk = A[o:o + m * n].view(n, m)
        blk.normal_(generator=g)
        blk.mul_((10.0 ** (4.0 * torch.rand(n, 1, device="cuda", generator=g))).float())
    b = torch.randn(plan.b_elems, dtype=torch.float64, device="cuda", generator=g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, info = plan.solve(A, b)              # autotune + solve
    torch.cuda.synchronize()
    tune_s = time.perf_counter() - t0
    reps = 3
    ctx.timing(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    for _ in range(reps):
        x, info = plan.solve(A, b)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    ev_ms, ev_n = ctx.timer("batch_ls")
    ctx.timing(False)
    itn = np.array([i["itn"] for i in info])
    ok = sum(i["status"] == 0 for i in info)
    launch_ms = ev_ms / max(ev_n, 1)
    # algorithmic HBM bytes: A read once by the sketch and once per LSQR pass
    # (initial + itn), u read and written per pass, in FP32 / FP64
    mm = np.array([s[0] for s in shapes], dtype=np.float64)
    nn = np.array([s[1] for s in shapes], dtype=np.float64)
    byts = float(np.sum(mm * nn * 4 * (itn + 2) + 16 * mm * (itn + 1)))
    flops_sketch = float(np.sum(2 * np.floor(2 * nn) * mm * (nn + 1)))
    # spot check against FP64 least squares
    errs = []
    xh = plan.unpack(x)
    for i in rs.choice(count, 6, replace=False):
        m, n = shapes[i]
        Ai = A[plan.a_offsets[i]:plan.a_offsets[i] + m * n].view(n, m).t().double()
        bi = b[plan.b_offsets[i]:plan.b_offsets[i] + m]
        xl = torch.linalg.lstsq(Ai.cpu(), bi.cpu().unsqueeze(1)).solution.squeeze(1).numpy()
        errs.append(float(np.linalg.norm(xh[i] - xl) / np.linalg.norm(xl)))
    cfg = plan.config()
    plan.close()
    out = {"config": f"{count} problems, m in [256,16384] log-uniform, n in [16,256], m>=4n, "
                     "A FP32 column-scaled kappa 1e4, Gaussian sketch d=2n, LSQR tol 1e-10",
           "problems_per_s": count / dt, "ms_per_batch": dt * 1e3, "launch_ms": launch_ms,
           "solved": int(ok), "itn_mean": float(itn.mean()), "itn_max": int(itn.max()),
           "autotune_s": tune_s, "tile": cfg,
           "hbm_GBps": byts / (launch_ms * 1e-3) / 1e9, "hbm_frac": byts / (launch_ms * 1e-3) / 8e12,
           "sketch_TFLOPs_equiv": flops_sketch / (launch_ms * 1e-3) / 1e12,
           "max_rel_err_vs_lstsq_sample": max(errs)}
    if cpu:
        # A loop of single problems (the reference solves one per call) in the
        # oracle: blendenpik with a Gaussian S (given, not timed) and LSQR,
        # FP64, over a uniform sample of the batch.
        from oracle import la
        ns = 64
        pick = rs.choice(count, ns, replace=False)
        probs = []
        for i in pick:
            m, n = shapes[i]
            Ai = A[plan.a_offsets[i]:plan.a_offsets[i] + m * n].view(n, m).t().double().cpu().numpy()
            bi = b[plan.b_offsets[i]:plan.b_offsets[i] + m].cpu().numpy()
            probs.append((Ai, bi, rs.standard_normal((2 * n, m)) / np.sqrt(2 * n)))

        def loop():
            for Ai, bi, Si in probs:
                la.blendenpik_lsqr(Ai, bi, 1e-10, 200, 2.0, Si)
        t = _timed(loop)
        out["cpu_baseline"] = {"value": ns / t, "unit": "problems/s", "cores": _cpu_threads(),
                               "kind": "port", "cpu": _cpu_model(),
                               "sample": f"oracle blendenpik_lsqr (numpy FP64, Gaussian S given) "
                                         f"on {ns} problems drawn uniformly from the batch: "
                                         f"{t:.2f} s"}
        out["gpu_over_cpu"] = (count / dt) / (ns / t)
    return out


def bench_mlp(L, torch, ctx):
    """Proposer network throughput (34 -> 128 -> 128 -> 3, BF16 MFMA)."""
    B = 8192
    from randnla_amd import proposer
    P = torch.tensor(proposer.init_params(0), device="cuda")
    mo, vo = torch.zeros_like(P), torch.zeros_like(P)
    F = torch.randn(B, 34, device="cuda")
    T = torch.randn(B, 3, device="cuda")
    wo = torch.ones(3, device="cuda")
    for s in range(1, 4):
        L.mlp_train_step(P, mo, vo, F, T, wo, s, 1e-3, ctx=ctx, want_loss=False)
    torch.cuda.synchronize()
    reps = 50
    t0 = time.perf_counter()
    for s in range(4, 4 + reps):
        L.mlp_train_step(P, mo, vo, F, T, wo, s, 1e-3, ctx=ctx, want_loss=False)
    torch.cuda.synchronize()
    dt_train = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        L.mlp_forward(P, F, ctx=ctx)
    torch.cuda.synchronize()
    dt_fwd = (time.perf_counter() - t0) / reps
    fl = 2.0 * (34 * 128 + 128 * 128 + 128 * 3)
    return {"config": f"MLP 34-128-128-3 GELU, BF16 MFMA, batch {B}",
            "train_samples_per_s": B / dt_train, "infer_samples_per_s": B / dt_fwd,
            "train_ms_per_step": dt_train * 1e3, "infer_ms": dt_fwd * 1e3,
            "train_GFLOPs": 3 * fl * B / dt_train / 1e9}


def cpu_baseline(q, l):
    """The oracle (numpy/OpenBLAS, FP64) on a bounded sample of the same
    workload shape: A 16384 x 4096 with the same spectrum family."""
    from oracle import la, rng as R
    try:
        from threadpoolctl import threadpool_info
        threads = max((d.get("num_threads", 1) for d in threadpool_info()), default=1)
    except Exception:
        threads = os.cpu_count() or 1
    m, n = 16384, 4096
    rs = np.random.default_rng(0)
    sig = (1.0 + np.arange(n)) ** -2.0
    u, _ = np.linalg.qr(rs.standard_normal((m, 512)))
    v, _ = np.linalg.qr(rs.standard_normal((n, 512)))
    a = (u * sig[:512]) @ v.T + 1e-7 * rs.standard_normal((m, n))
    om = R.gaussian(R.Stream.from_seed(R.PHILOX4X32, 42), n * l).reshape(l, n).T
    dt = _timed(lambda: la.rand_svd_power(a, l - 64, 64, q, om))
    flops = (2 * q + 2) * 2.0 * m * n * l
    # one core (the reference's matrixmultiply is single-threaded) on a
    # quarter of the sample, and the reference's own computation (faithful
    # tsog1, full-pivot LU, Householder QR) on all cores
    mq, nq = m // 2, n // 2
    aq = np.ascontiguousarray(a[:mq, :nq])
    dt1 = _timed(lambda: la.rand_svd_power(aq, l - 64, 64, q, om[:nq]), 1)
    flops1 = (2 * q + 2) * 2.0 * mq * nq * l
    dtf = _timed(lambda: la.rand_svd(a, l - 64, 1e-6, 64))
    return {"value": flops / dt / 1e9, "unit": "GFLOP/s", "cores": int(threads), "kind": "port",
            "cpu": _cpu_model(),
            "sample": f"oracle rand_svd, intended method (numpy FP64) A {m}x{n}, l={l}, q={q}: "
                      f"{dt:.2f} s",
            "single_thread": {"value": flops1 / dt1 / 1e9, "unit": "GFLOP/s", "cores": 1,
                              "sample": f"the same on A {mq}x{nq}: {dt1:.2f} s"},
            "faithful_all_cores": {"sample": f"oracle rand_svd, the reference's computation, "
                                             f"A {m}x{n}, k={l - 64}, s=64", "s": dtf}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    import torch
    import randnla_amd as L

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ctx = L.Context(0)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)
    args.only = set(args.only.split(","))
    line = bench_svd(args, L, torch, ctx, dev) if args.only & {"all", "svd"} else {
        "metric": "rand_svd time-to-1e-6 rel err & GFLOP/s: mixed-prec vs FP64, 1xMI355X",
        "value": None, "note": f"--only {args.only}: headline workload not run"}
    line["workloads"] = other_workloads(args, L, torch, ctx)
    torch.cuda.synchronize()
    ctx.close()
    print(json.dumps(line))


def bench_svd(args, L, torch, ctx, dev):
    m, n, k, p, q = args.m, args.n, args.k, args.p, args.q
    l = k + p
    sig = (1.0 + np.arange(n)) ** -2.0
    A = L.test_matrix(m, n, sig, noise=1e-7, rng=L.rng(L.PHILOX4X32, (2024, 1)), dtype=L.F32)
    torch.cuda.synchronize()
    flops = (2 * q + 2) * 2.0 * m * n * l

    def run(prec):
        return L.rand_svd_factors(A, k, 1e-6, p, precision=prec, power_iters=q, ctx=ctx)

    for _ in range(args.warmup):
        run(L.BF16)
    torch.cuda.synchronize()
    ctx.timing(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = run(L.BF16)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    lp_ms, lp_n = ctx.timer("gemm_bf16_nn")
    f2_ms, f2_n = ctx.timer("gemm_f16x2")
    breakdown = {}
    for name in ("gemm_bf16_nn", "gemm_bf16_tn", "gemm_f16", "gemm_f16x2", "gemm_bt",
                 "convert_bf16", "gram", "trmm", "cholesky", "rprod", "jacobi", "svd_u", "svd_v"):
        t_ms, _ = ctx.timer(name)
        breakdown[name] = t_ms / args.steps
    ctx.timing(False)
    U, sv, Vt = out
    jacobi_sweeps = int(ctx.diag(0))
    e_mx, an = L.spectral_error(A, U, sv, Vt, iters=args.err_iters, ctx=ctx)

    # FP64 path on the same FP32 bits.
    t64 = None
    if args.fp64_steps > 0:
        run(L.FP64)
        torch.cuda.synchronize()
        ctx.timing(True)
        ctx.timing_reset()
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            o64 = run(L.FP64)
        torch.cuda.synchronize()
        t64 = (time.perf_counter() - t0) / args.fp64_steps
        # time of the FP64 path's passes over A (its own FP64 MFMA GEMM)
        gemm64_ms = sum(ctx.timer(t)[0] for t in ("sketch_gemm", "pass_a", "pass_at", "pass_bt")
                        ) / args.fp64_steps
        ctx.timing(False)
        e_64, _ = L.spectral_error(A, *o64, iters=args.err_iters, ctx=ctx)
        ds = float((o64[1] - sv).abs().max() / o64[1][0])
    else:
        e_64, ds = None, None

    # FP8-e4m3 ladder (128 x 128 block scales on A, per-column 128-K scales on
    # the panels) on the same FP32 bits.
    fp8 = None
    if args.fp8_steps > 0:
        run(L.FP8)
        torch.cuda.synchronize()
        ctx.timing(True)
        ctx.timing_reset()
        t0 = time.perf_counter()
        for _ in range(args.fp8_steps):
            o8 = run(L.FP8)
        torch.cuda.synchronize()
        t8 = (time.perf_counter() - t0) / args.fp8_steps
        br8 = {}
        for name in ("gemm_fp8_nn", "quantize_fp8", "convert_bf16", "gemm_bf16_nn", "gemm_bf16_tn",
                     "gemm_f16", "gemm_f16x2", "gemm_bt", "gram", "trmm", "cholesky", "jacobi"):
            br8[name] = ctx.timer(name)[0] / args.fp8_steps
        nn8_ms, nn8_n = ctx.timer("gemm_fp8_nn")
        ctx.timing(False)
        e_8, _ = L.spectral_error(A, *o8, iters=args.err_iters, ctx=ctx)
        a8 = 2.0 * m * n * l / (nn8_ms / max(nn8_n, 1) * 1e-3) / 1e12
        fp8 = {"ms_per_step": t8 * 1e3, "GFLOP/s": flops / t8 / 1e9,
               "rel_err": e_8 / an,
               "abs_diff_vs_fp64": abs(e_8 - e_64) / an if e_64 else None,
               "max_sigma_diff_rel_vs_fp64": float((o64[1] - o8[1]).abs().max() / o64[1][0]) if e_64 else None,
               "breakdown_ms_per_step": br8,
               "gemm_fp8_nn": {"avg_launch_ms": nn8_ms / max(nn8_n, 1), "launches": nn8_n,
                               "achieved_TFLOPs": a8, "frac_of_fp8_peak": a8 / 5000.0}}
        del o8

    # Vendor FP64 GEMM (torch.matmul -> rocBLAS/hipBLASLt) on one pass shape,
    # for context on the FP64 path's own MFMA kernel.
    vendor_ms = None
    if args.fp64_steps > 0:
        A64 = A.double()
        X = torch.randn(n, l, dtype=torch.float64, device=dev)
        torch.matmul(A64, X)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            torch.matmul(A64, X)
        torch.cuda.synchronize()
        vendor_ms = (time.perf_counter() - t0) / 3 * 1e3
        del A64, X
    fp64v = None
    if t64 and vendor_ms:
        fp64v = (2 * q + 2) * vendor_ms + (t64 * 1e3 - gemm64_ms)
    avg_launch_ms = lp_ms / max(lp_n, 1)
    # per launch: A (FP32) read once, the FP32 panel read and the FP32 product
    # written (m x l and n x l in one order or the other; the A^T Q launches'
    # split-K partial slabs are counted as one product)
    f2_bytes = 4.0 * m * n + 4.0 * (m + n) * l
    f2_gbps = f2_bytes / (f2_ms / max(f2_n, 1) * 1e-3) / 1e9 if f2_n else None
    per_launch_flops = 2.0 * m * n * l
    achieved = per_launch_flops / (avg_launch_ms * 1e-3) / 1e12
    line = {
        "metric": "rand_svd time-to-1e-6 rel err & GFLOP/s: mixed-prec vs FP64, 1xMI355X",
        "value": flops / dt / 1e9,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"mixed-precision rand_svd, A {m}x{n} FP32, sigma_i=(1+i)^-2 + 1e-7 noise",
                   "m": m, "n": n, "k": k, "p": p, "q": q,
                   "precision": "bf16 A Omega and first round, then fp32 A rounded once to fp16 "
                                "(x panel fp16), last round split-fp16 (A hi/lo x panel fp16), "
                                "B = A^T Q with A rounded once to fp16 and Q split hi/lo, "
                                "fp64 CholQR2/SVD"},
        "accuracy": {"rel_err_mixed": e_mx / an, "rel_err_fp64": (e_64 / an) if e_64 else None,
                     "abs_diff": abs(e_mx - e_64) / an if e_64 else None,
                     "max_sigma_diff_rel": ds},
        "fp64_ms_per_step": t64 * 1e3 if t64 else None,
        "fp64_vendor_gemm_ms_per_pass": vendor_ms,
        "speedup_vs_fp64": (t64 / dt) if t64 else None,
        # the FP64 path with its passes over A at the vendor GEMM's speed
        # ((2q+2) passes of hipBLASLt's FP64 GEMM) plus its own non-GEMM work
        "fp64_vendor_estimate_ms_per_step": fp64v,
        "speedup_vs_fp64_vendor_estimate": (fp64v / (dt * 1e3)) if fp64v else None,
        "fp8_path": fp8,
        "breakdown_ms_per_step": breakdown,
        "jacobi_sweeps": jacobi_sweeps,
        # the clean-round products (the largest GEMM class of the step): A
        # split into FP16 hi/lo times the FP16 panel, reading A (FP32) once
        "roofline": {"bound": "hbm",
                     "kernel": "gemm_lp_kernel<2,true,5,false> (A^T Q and A Qz, A split FP16)",
                     "achieved": f2_gbps, "peak": 8000.0, "unit": "GB/s",
                     "frac": f2_gbps / 8000.0, "traffic": TRAFFIC_F16X2,
                     "traffic_algorithmic": f2_bytes,
                     "avg_launch_ms": f2_ms / max(f2_n, 1), "launches": f2_n},
        "roofline_bf16_nn": {"bound": "mfma", "kernel": "gemm_lp_kernel<0,false,5,true> (Y = A_bf16 X)",
                             "achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s",
                             "frac": achieved / 2500.0, "traffic": TRAFFIC_BF16_NN,
                             "traffic_algorithmic": 2.0 * m * n + 2.0 * n * l + 4.0 * m * l,
                             "avg_launch_ms": avg_launch_ms, "launches": lp_n},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(q, l)
    del A
    torch.cuda.empty_cache()
    return line


def other_workloads(args, L, torch, ctx):
    out = {}
    if args.only & {"all", "c1"}:
        out["c1_rand_svd_fp64"] = bench_c1(L, torch, ctx, cpu=not args.no_cpu)
    if args.only & {"all", "cqrrpt"}:
        out["cqrrpt"] = bench_cqrrpt(L, torch, ctx, cpu=not args.no_cpu)
    if args.only & {"all", "ls"}:
        out["ls_sparse_sign_lsqr"] = bench_ls(L, torch, ctx, cpu=not args.no_cpu)
    if args.only & {"all", "sls"}:
        out["sketched_ls"] = bench_sketch_solve(L, torch, ctx, cpu=not args.no_cpu)
    if args.only & {"all", "evd"}:
        out["rand_evd"] = bench_evd(L, torch, ctx, cpu=not args.no_cpu)
    if args.only & {"all", "id"}:
        out["interpolative"] = bench_id(L, torch, ctx, cpu=not args.no_cpu)
    if args.only & {"all", "batch"}:
        out["batched_ls"] = bench_batch(L, torch, ctx, args.batch_count, cpu=not args.no_cpu)
    if args.only & {"all", "mlp"}:
        out["proposer_mlp"] = bench_mlp(L, torch, ctx)
    return out


if __name__ == "__main__":
    main()
