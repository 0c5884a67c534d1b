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
    ap.add_argument("--err-iters", type=int, default=25)
    ap.add_argument("--only", choices=["all", "svd", "ls", "sls", "evd", "id", "batch", "mlp"], default="all",
                    help="run one workload only (profiling); 'svd' is the headline one")
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


def bench_ls(L, torch, ctx):
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
    return {"config": f"blendenpik + LSQR, A {m}x{n} FP64, kappa 1e8, sparse-sign d={4 * n} zeta=8, tol 1e-12",
            "ms": dt * 1e3, "itn": it, "istop": info["istop"],
            "rel_resid_normal_eq": float((A.t() @ r).norm() / (A.norm() * r.norm())),
            "rel_err_vs_xstar": float((x - xs).norm() / xs.norm()),
            "breakdown_ms": {"sketch_apply": sk_ms, "householder": hh_ms, "lsqr": lsqr_ms},
            "lsqr_passes_over_A_per_iteration": 1 if fused else 2,
            "lsqr_hbm_GBps": lsqr_bytes / (lsqr_ms * 1e-3) / 1e9 if lsqr_ms > 0 else None,
            "lsqr_hbm_frac": lsqr_bytes / (lsqr_ms * 1e-3) / 8e12 if lsqr_ms > 0 else None}


def _cpu_threads():
    try:
        from threadpoolctl import threadpool_info
        return int(max((d.get("num_threads", 1) for d in threadpool_info()), default=1))
    except Exception:
        return os.cpu_count() or 1


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


def bench_batch(L, torch, ctx, count):
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
    return {"config": f"{count} problems, m in [256,16384] log-uniform, n in [16,256], m>=4n, "
                      "A FP32 column-scaled kappa 1e4, Gaussian sketch d=2n, LSQR tol 1e-10",
            "problems_per_s": count / dt, "ms_per_batch": dt * 1e3, "launch_ms": launch_ms,
            "solved": int(ok), "itn_mean": float(itn.mean()), "itn_max": int(itn.max()),
            "autotune_s": tune_s, "tile": cfg,
            "hbm_GBps": byts / (launch_ms * 1e-3) / 1e9, "hbm_frac": byts / (launch_ms * 1e-3) / 8e12,
            "sketch_TFLOPs_equiv": flops_sketch / (launch_ms * 1e-3) / 1e12,
            "max_rel_err_vs_lstsq_sample": max(errs)}


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
    t0 = time.perf_counter()
    la.rand_svd_power(a, l - 64, 64, q, om)
    dt = time.perf_counter() - t0
    flops = (2 * q + 2) * 2.0 * m * n * l
    return {"value": flops / dt / 1e9, "unit": "GFLOP/s", "cores": int(threads), "kind": "port",
            "sample": f"oracle rand_svd (numpy FP64) A {m}x{n}, l={l}, q={q}: {dt:.2f} s"}


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
    line = bench_svd(args, L, torch, ctx, dev) if args.only in ("all", "svd") else {
        "metric": "rand_svd time-to-1e-6 rel err & GFLOP/s: mixed-prec vs FP64, 1xMI355X",
        "value": None, "note": f"--only {args.only}: headline workload not run"}
    line["workloads"] = other_workloads(args, L, torch, ctx)
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
    breakdown = {}
    for name in ("gemm_bf16_nn", "gemm_bf16_tn", "gemm_f16x3", "convert_bf16", "gram", "trmm",
                 "cholesky", "rprod", "jacobi", "svd_u", "svd_v"):
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
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            o64 = run(L.FP64)
        torch.cuda.synchronize()
        t64 = (time.perf_counter() - t0) / args.fp64_steps
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
        for name in ("gemm_fp8_nn", "gemm_fp8_tn", "quantize_fp8", "gemm_f16x3", "gram", "trmm",
                     "cholesky", "jacobi"):
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
    avg_launch_ms = lp_ms / max(lp_n, 1)
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
                   "m": m, "n": n, "k": k, "p": p, "q": q, "precision": "bf16 range finder, split-fp16 B, fp64 CholQR2/SVD"},
        "accuracy": {"rel_err_mixed": e_mx / an, "rel_err_fp64": (e_64 / an) if e_64 else None,
                     "abs_diff": abs(e_mx - e_64) / an if e_64 else None,
                     "max_sigma_diff_rel": ds},
        "fp64_ms_per_step": t64 * 1e3 if t64 else None,
        "fp64_vendor_gemm_ms_per_pass": vendor_ms,
        "speedup_vs_fp64": (t64 / dt) if t64 else None,
        "fp8_path": fp8,
        "breakdown_ms_per_step": breakdown,
        "jacobi_sweeps": jacobi_sweeps,
        "roofline": {"bound": "mfma", "kernel": "gemm_lp_kernel<0,false,5> (Y = A_bf16 X)",
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
    if args.only in ("all", "ls"):
        out["ls_sparse_sign_lsqr"] = bench_ls(L, torch, ctx)
    if args.only in ("all", "sls"):
        out["sketched_ls"] = bench_sketch_solve(L, torch, ctx, cpu=not args.no_cpu)
    if args.only in ("all", "evd"):
        out["rand_evd"] = bench_evd(L, torch, ctx, cpu=not args.no_cpu)
    if args.only in ("all", "id"):
        out["interpolative"] = bench_id(L, torch, ctx, cpu=not args.no_cpu)
    if args.only in ("all", "batch"):
        out["batched_ls"] = bench_batch(L, torch, ctx, args.batch_count)
    if args.only in ("all", "mlp"):
        out["proposer_mlp"] = bench_mlp(L, torch, ctx)
    return out


if __name__ == "__main__":
    main()
