"""Training-set generation and training of the geodesic surrogate.

Training set (the c5 recipe): rays sampled uniformly in spin a in [0, 0.998],
inclination in [5, 85] deg and (alpha, beta) in [-20, 20]^2 M, with the band
within 1 M of the critical curve oversampled x8; every ray traced by the
FP64 DP45 GPU integrator (disk disabled, so all equatorial crossings are
recorded) from r_obs = 100 M.  Targets come from the records
(bh_surrogate_targets).  Training is Adam with a cosine learning-rate
schedule on the BF16-MFMA kernels of libbhk.so.

    python -m blackhole_amd.surrogate_train --n 16777216 --steps 40000 \\
        --out blackhole_amd/weights/surrogate_c5.npz
"""
from __future__ import annotations

import argparse
import math
import os
import time

import numpy as np

from . import abi, analysis
from . import surrogate as S
from . import tracer

R_OBS = 100.0
SPIN_MAX = 0.998
INC_DEG = (5.0, 85.0)
HALF = 20.0
BAND = 1.0          # M, half width of the oversampled band around the critical curve
OVERSAMPLE = 8.0


def geometry_options():
    """DP45 FP64, disk disabled (every crossing recorded), r_escape 200."""
    return tracer.default_options(
        integrator=abi.INTEGRATOR_DP45, precision=abi.PRECISION_FP64, max_steps=20000,
        max_attempts=100000, disk_enabled=0, volume_enabled=0, rtol=1e-8, atol=1e-10,
        h_init=0.01, h_max=0.1, horizon_factor=1.01, r_escape=200.0, max_crossings=3)


def _band_points(rng, a, th):
    """One random critical-curve point per (a, theta) (vectorised rejection)."""
    n = a.size
    out_a = np.empty(n)
    out_b = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        aa = np.maximum(a[todo], 1e-6)
        tt = th[todo]
        rp = 2.0 * (1.0 + np.cos((2.0 / 3.0) * np.arccos(-aa)))
        rm = 2.0 * (1.0 + np.cos((2.0 / 3.0) * np.arccos(aa)))
        r = rp + (rm - rp) * rng.uniform(0.0, 1.0, todo.size)
        delta = r * r - 2.0 * r + aa * aa
        dp = 2.0 * r - 2.0
        xi = (r * r + aa * aa) / aa - 4.0 * r * delta / (aa * dp)
        eta = 16.0 * r * r * delta / (dp * dp) - (xi - aa) ** 2
        s, c = np.sin(tt), np.cos(tt)
        b2 = eta + aa * aa * c * c - xi * xi * (c / s) ** 2
        ok = b2 >= 0
        sign = np.where(rng.uniform(0, 1, todo.size) < 0.5, -1.0, 1.0)
        out_a[todo[ok]] = -xi[ok] / s[ok]
        out_b[todo[ok]] = sign[ok] * np.sqrt(b2[ok])
        todo = todo[~ok]
    return out_a, out_b


def sample_specs(n: int, seed: int) -> np.ndarray:
    """(alpha, beta, spin, theta_obs, r_obs) rows, float64 [n, 5]."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, SPIN_MAX, n)
    th = np.radians(rng.uniform(INC_DEG[0], INC_DEG[1], n))
    # band area ~ perimeter (~30 M) x 2 M of a 1600 M^2 field: a uniform
    # sample puts ~3.8 % there; x8 density adds ~26 % more samples in the band
    band_frac = 0.21
    nb = int(n * band_frac)
    alpha = rng.uniform(-HALF, HALF, n)
    beta = rng.uniform(-HALF, HALF, n)
    ca, cb = _band_points(rng, a[:nb], th[:nb])
    rad = BAND * np.sqrt(rng.uniform(0, 1, nb))
    ang = rng.uniform(0, 2 * math.pi, nb)
    alpha[:nb] = np.clip(ca + rad * np.cos(ang), -HALF, HALF)
    beta[:nb] = np.clip(cb + rad * np.sin(ang), -HALF, HALF)
    return np.stack([alpha, beta, a, th, np.full(n, R_OBS)], axis=1)


def rays4_from_specs(specs) -> np.ndarray:
    return np.stack([specs[:, 0], specs[:, 1], specs[:, 2], np.cos(specs[:, 3])], axis=1).astype(np.float32)


def build_dataset(n: int, seed: int, chunk: int = 1 << 22):
    """Trace n sampled rays on the GPU; returns device (features, T, W) and the specs."""
    import torch
    specs = sample_specs(n, seed)
    opts = geometry_options()
    feats = torch.empty((n, S.IN_PAD), dtype=torch.float32, device="cuda")
    T = torch.empty((n, S.OUT), dtype=torch.float32, device="cuda")
    W = torch.empty_like(T)
    for i in range(0, n, chunk):
        j = min(n, i + chunk)
        d_specs = torch.from_numpy(specs[i:j]).cuda()
        rec = tracer.trace_rays(d_specs, opts)
        feats[i:j] = S.Surrogate.features(torch.from_numpy(rays4_from_specs(specs[i:j])).cuda())
        T[i:j], W[i:j] = S.targets_from_records(rec, feats[i:j])
        del rec, d_specs
    torch.cuda.synchronize()
    return feats, T, W, specs


def train(sur: S.Surrogate, feats, T, W, steps: int, batch: int, lr_max: float = 2e-3,
          lr_min: float = 2e-5, log_every: int = 1000, seed: int = 0):
    import torch
    n = feats.shape[0]
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    perm = torch.randperm(n, device="cuda", generator=g)
    pos = 0
    hist = []
    t0 = time.time()
    for step in range(steps):
        if pos + batch > n:
            perm = torch.randperm(n, device="cuda", generator=g)
            pos = 0
        idx = perm[pos:pos + batch]
        pos += batch
        lr = lr_min + 0.5 * (lr_max - lr_min) * (1.0 + math.cos(math.pi * step / max(steps, 1)))
        lsum, wsum = sur.train_step(feats[idx], T[idx], W[idx], lr)
        if step % log_every == 0 or step == steps - 1:
            hist.append((step, lsum / batch, time.time() - t0))
            print(f"[train] step {step} loss/row {lsum / batch:.3e} lr {lr:.2e} t {time.time() - t0:.1f}s",
                  flush=True)
    return hist


def evaluate(sur: S.Surrogate, feats, T, W) -> dict:
    """Error statistics of every output column on a held-out set."""
    Y = sur.forward(feats).cpu().numpy()
    T = T.cpu().numpy()
    W = W.cpu().numpy()
    out = {}
    names = {S.LOGR1: "ln_r1", S.G1: "g1", S.LOGR2: "ln_r2", S.G2: "g2", S.NCROSS: "ncross",
             S.CAPTURE: "capture", S.ESC_U: "esc_u"}
    for c, name in names.items():
        m = W[:, c] > 0
        if m.any():
            e = np.abs(Y[m, c] - T[m, c])
            out[name] = {"median": float(np.median(e)), "p90": float(np.quantile(e, 0.9)),
                         "p99": float(np.quantile(e, 0.99)), "n": int(m.sum())}
    return out


def save(sur: S.Surrogate, path: str, meta: dict) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    np.savez(path, params=sur.get_params(), width=sur.width, depth=sur.depth,
             meta=np.array(repr(meta)))


def load(path: str, max_batch: int = 16384) -> S.Surrogate:
    z = np.load(path, allow_pickle=False)
    sur = S.Surrogate(width=int(z["width"]), depth=int(z["depth"]), max_batch=max_batch)
    sur.set_params(z["params"])
    return sur


def time_train_steps(batch: int = 16384, steps: int = 20, warmup: int = 3, seed: int = 5) -> dict:
    """Wall time per training step (forward, loss, backward, Adam) of the
    width-256, depth-6 network on a synthetic batch of `batch` rays with
    random targets, inputs resident on the device."""
    import torch
    sur = S.Surrogate(width=256, depth=6, max_batch=batch, seed=seed)
    rng = np.random.default_rng(seed)
    rays = np.stack([rng.uniform(-20, 20, batch), rng.uniform(-20, 20, batch), rng.uniform(0, 0.998, batch),
                     np.cos(np.radians(rng.uniform(5, 85, batch)))], 1).astype(np.float32)
    feats = S.Surrogate.features(torch.from_numpy(rays).cuda())
    T = torch.from_numpy(rng.normal(0, 1, (batch, S.OUT)).astype(np.float32)).cuda()
    W = torch.ones((batch, S.OUT), dtype=torch.float32, device="cuda")
    for _ in range(warmup):
        sur.train_step(feats, T, W, 1e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = [sur.train_step(feats, T, W, 1e-4)[0] for _ in range(steps)]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    return {"batch": batch, "steps": steps, "ms_per_step": ms, "rows_per_s": batch / ms * 1e3,
            "loss_first": losses[0], "loss_last": losses[-1]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--time-steps", type=int, default=0,
                    help="only time this many training steps on a synthetic batch and print them")
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--n-eval", type=int, default=1 << 18)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    import torch  # noqa: F401
    if args.time_steps > 0:
        print("[time]", time_train_steps(max(args.batch, 64), args.time_steps), flush=True)
        return 0
    t0 = time.time()
    feats, T, W, _ = build_dataset(args.n, args.seed)
    print(f"[data] {args.n} rays traced + targets in {time.time() - t0:.1f}s", flush=True)
    fe, Te, We, _ = build_dataset(args.n_eval, args.seed + 1)
    sur = S.Surrogate(width=256, depth=6, max_batch=max(args.batch, 16384), seed=args.seed)
    hist = train(sur, feats, T, W, args.steps, args.batch, lr_max=args.lr)
    ev = evaluate(sur, fe, Te, We)
    print("[eval]", ev, flush=True)
    if args.out:
        save(sur, args.out, {"n": args.n, "steps": args.steps, "batch": args.batch, "seed": args.seed,
                             "final_loss": hist[-1][1], "eval": ev})
        print("[save]", args.out, flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
