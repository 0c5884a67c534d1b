"""Learned (k, p, q) proposer for rand_svd.

The reference leaves the target rank k, the oversampling s and the number of
power iterations to the caller (src/lora_drivers.rs:30-69; the power
iterations are tsog1's num_passes, src/lora_helpers.rs:58-105).  This module
proposes them for a relative spectral error target:

* features: ``probe_features`` -- the top 32 singular values of A Omega0
  (Omega0 n x 48 Gaussian) on a log scale, plus log m and log n;
* network: the library's 34 -> 128 -> 128 -> 3 GELU MLP (BF16 matrix-core
  kernels, FP32 accumulation), trained on the GPU with the library's own
  forward/backward and Adam kernels;
* labels: for each training matrix, the cheapest (k, p, q) on a grid that
  reaches the error target under plain FP64 rand_svd, cost measured by the
  flop model ``cost``; matrices where no grid point reaches it within
  k + p <= min(m, n) / 2 are flagged infeasible and left out;
* use: ``Proposer.solve`` runs rand_svd at the proposed point, checks it with
  the a-posteriori estimator (block power iteration on A - U S V^T) and, when
  the estimate misses the target, falls back to the plain method (doubling k)
  from there.  The plain method alone (``plain``) is the baseline: fixed
  (k, p, q) = (32, 10, 2), k doubled until the estimate meets the target.

``python -m randnla_amd.proposer generate --count N --seed S --out F.npz``
writes a training set (features, targets, metadata).
"""
import argparse
import math
import sys
import time

import numpy as np

import randnla_amd as L

TOL = 1e-6
P_GRID = (8, 16, 32)
Q_GRID = (0, 1, 2, 3)
K_STEP = 8
ERR_ITERS = 20
PLAIN = (32, 10, 2)
OUT_WEIGHTS = np.array([1.0, 0.25, 1.0], dtype=np.float32)
SIZES = ((1024, 256), (2048, 256), (2048, 512), (4096, 512))
FAMILIES = ("geometric", "polynomial", "step", "noise")


# ---------------------------------------------------------------- encoding
def encode(k, p, q):
    return np.array([math.log2(k) / 10.0, math.log2(p) / 10.0, q / 4.0], dtype=np.float32)


def kmax(m, n, p):
    return max(K_STEP, min(m, n) // 2 - p)


def decode(y, m, n):
    """Network output -> (k, p, q) on the grid, within the feasible range."""
    p = min(P_GRID, key=lambda c: abs(math.log2(c) - 10.0 * float(y[1])))
    q = int(min(max(round(4.0 * float(y[2])), 0), max(Q_GRID)))
    k = int(round(2.0 ** (10.0 * float(y[0])) / K_STEP)) * K_STEP
    k = int(min(max(k, K_STEP), kmax(m, n, p)))
    return k, p, q


# ------------------------------------------------------------- cost model
def cost(m, n, k, p, q):
    """Flops of one FP64 rand_svd call: (2q + 2) passes over A and the
    orthonormalisation of each l-column panel."""
    l = k + p
    return (2 * q + 2) * 2.0 * m * n * l + (2 * q + 2) * 4.0 * max(m, n) * l * l


def estimate_cost(m, n, k):
    """The a-posteriori estimator: ERR_ITERS block power steps with 8 vectors
    on A - U S V^T."""
    return ERR_ITERS * 2.0 * (2.0 * m * n * 8 + 2.0 * (m + n) * k * 8)


def probe_cost(m, n):
    return 2.0 * m * n * 48 + 8.0 * m * 48 * 48


# ---------------------------------------------------------- training data
def spectrum(family, n, rng):
    """(singular values, noise level) of one matrix family: s_0 = 1,
    descending; the matrix adds noise * G with G standard normal (a floor of
    about noise * (sqrt(m) + sqrt(n)) in the spectrum)."""
    i = np.arange(n, dtype=np.float64)
    if family == "geometric":
        return rng.uniform(0.6, 0.93) ** i, 0.0
    if family == "polynomial":
        return (1.0 + i) ** -rng.uniform(2.6, 5.0), 0.0
    if family == "step":
        r = int(rng.integers(8, 100))
        return np.where(i < r, 0.98 ** i, 10.0 ** rng.uniform(-10.0, -7.0)), 0.0
    if family == "noise":
        base = rng.uniform(0.6, 0.9) ** i if rng.random() < 0.5 else (1.0 + i) ** -rng.uniform(3.0, 5.0)
        return base, 10.0 ** rng.uniform(-10.0, -8.5)
    raise ValueError(family)


def make_matrix(m, n, sigma, seed, noise=0.0):
    return L.test_matrix(m, n, sigma, noise=noise, rng=L.rng(L.PHILOX4X32, (int(seed), 17)),
                         dtype=L.F64)


def run(A, k, p, q):
    """Plain FP64 rand_svd at (k, p, q) and its a-posteriori relative error."""
    U, s, Vt = L.rand_svd_factors(A, k, 1e-6, p, precision=L.FP64, power_iters=q)
    e, an = L.spectral_error(A, U, s, Vt, iters=ERR_ITERS)
    return e / an if an > 0 else 0.0


def label(A, sigma, tol=TOL):
    """Cheapest grid point reaching tol: (cost, k, p, q) or None."""
    m, n = A.shape
    kopt = int(np.sum(sigma > tol * sigma[0]))
    k0 = max(K_STEP, -(-kopt // K_STEP) * K_STEP)
    combos = sorted(((q, p) for q in Q_GRID for p in P_GRID), key=lambda c: cost(m, n, k0, c[1], c[0]))
    best = None
    for q, p in combos:
        k = k0
        for _ in range(6):
            if k > kmax(m, n, p):
                break
            c = cost(m, n, k, p, q)
            if best is not None and c >= best[0]:
                break
            if run(A, k, p, q) <= tol:
                best = (c, k, p, q)
                break
            k += K_STEP
    return best


def generate(count, seed, tol=TOL, log=sys.stdout):
    rng = np.random.default_rng(seed)
    feats, targs, meta = [], [], []
    t0 = time.time()
    for i in range(count):
        m, n = SIZES[int(rng.integers(len(SIZES)))]
        fam = FAMILIES[i % len(FAMILIES)]
        sigma, noise = spectrum(fam, n, rng)
        A = make_matrix(m, n, sigma, seed * 1000003 + i, noise)
        f = L.probe_features(A)
        lab = label(A, sigma, tol)
        ok = lab is not None
        c, k, p, q = lab if ok else (0.0, 0, 0, 0)
        feats.append(f)
        targs.append(encode(k, p, q) if ok else np.zeros(3, np.float32))
        meta.append([m, n, FAMILIES.index(fam), int(ok), k, p, q, c])
        if log and (i + 1) % 10 == 0:
            print(f"generate {i + 1}/{count} {time.time() - t0:.0f}s", file=log, flush=True)
    return {"features": np.array(feats, np.float32), "targets": np.array(targs, np.float32),
            "meta": np.array(meta, np.float64)}


# ---------------------------------------------------------------- network
def init_params(seed):
    """He initialisation (weights N(0, 2 / fan_in), last layer scaled by 0.5
    with fan-in variance 1, zero biases), packed W1 b1 W2 b2 W3 b3."""
    r = np.random.default_rng(seed)
    parts = [r.standard_normal((128, 34)) * np.sqrt(2.0 / 34), np.zeros(128),
             r.standard_normal((128, 128)) * np.sqrt(2.0 / 128), np.zeros(128),
             r.standard_normal((3, 128)) * 0.5 * np.sqrt(1.0 / 128), np.zeros(3)]
    return np.concatenate([x.reshape(-1) for x in parts]).astype(np.float32)


class Proposer:
    def __init__(self, seed=0, params=None):
        import torch
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        assert p.size == L.mlp_param_count()
        self.P = torch.tensor(p, device="cuda")
        self.m = torch.zeros_like(self.P)
        self.v = torch.zeros_like(self.P)
        self.step = 0
        self.wo = torch.tensor(OUT_WEIGHTS, device="cuda")

    def fit(self, F, T, epochs=300, batch=64, lr=2e-3, seed=0):
        """Adam on minibatches (library kernels); returns the mean training
        loss per epoch."""
        import torch
        F = torch.as_tensor(np.asarray(F, np.float32), device="cuda")
        T = torch.as_tensor(np.asarray(T, np.float32), device="cuda")
        g = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        N = F.shape[0]
        for _ in range(epochs):
            perm = torch.randperm(N, generator=g).to("cuda")
            tot, nb = 0.0, 0
            for s in range(0, N, batch):
                idx = perm[s:s + batch]
                self.step += 1
                tot += L.mlp_train_step(self.P, self.m, self.v, F[idx].contiguous(),
                                        T[idx].contiguous(), self.wo, self.step, lr)
                nb += 1
            hist.append(tot / nb)
        return hist

    def predict(self, F):
        y = L.mlp_forward(self.P.cpu().numpy(), np.asarray(F, np.float32).reshape(-1, 34))
        return np.asarray(y)

    def params(self):
        return self.P.cpu().numpy()

    def solve(self, A, tol=TOL):
        """Proposed (k, p, q), checked by the a-posteriori estimator, with the
        plain method (doubling k) as the fallback."""
        m, n = A.shape
        f = L.probe_features(A)
        k, p, q = decode(self.predict(f[None, :])[0], m, n)
        total = probe_cost(m, n)
        trials = 0
        while True:
            e = run(A, k, p, q)
            total += cost(m, n, k, p, q) + estimate_cost(m, n, k)
            trials += 1
            if e <= tol or k >= kmax(m, n, p):
                break
            k = min(2 * k, kmax(m, n, p))
        return {"k": k, "p": p, "q": q, "err": e, "cost": total, "trials": trials,
                "fallback": trials > 1}


def plain(A, tol=TOL, start=PLAIN):
    """The method without the proposer: fixed (k, p, q), k doubled until the
    a-posteriori estimate meets tol."""
    m, n = A.shape
    k, p, q = start
    total, trials = 0.0, 0
    while True:
        k = min(k, kmax(m, n, p))
        e = run(A, k, p, q)
        total += cost(m, n, k, p, q) + estimate_cost(m, n, k)
        trials += 1
        if e <= tol or k >= kmax(m, n, p):
            break
        k *= 2
    return {"k": k, "p": p, "q": q, "err": e, "cost": total, "trials": trials}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--count", type=int, default=480)
    g.add_argument("--seed", type=int, default=1)
    g.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "generate":
        d = generate(a.count, a.seed)
        np.savez_compressed(a.out, **d)
        ok = d["meta"][:, 3].astype(bool)
        print(f"wrote {a.out}: {len(ok)} matrices, {int(ok.sum())} feasible", flush=True)


if __name__ == "__main__":
    main()
