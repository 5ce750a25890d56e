"""Surrogate MLP kernels on the GPU: MFMA GEMM against numpy, training
gradients against the FP64 checker (oracle/mlp_oracle.py), and a short
training run that must reduce the loss."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))

pytestmark = pytest.mark.gpu


def test_gemm_nt_mfma_layout_and_values():
    import torch
    from blackhole_amd import surrogate
    g = torch.Generator().manual_seed(3)
    for (M, N, K) in ((64, 64, 32), (100, 70, 96), (16, 256, 256), (256, 16, 64)):
        A = torch.randn(M, K, generator=g).to(torch.bfloat16)
        B = torch.randn(N, K, generator=g).to(torch.bfloat16)
        bias = torch.randn(N, generator=g)
        got = surrogate.gemm_nt(A.cuda(), B.cuda(), bias.cuda()).cpu().double()
        ref = A.double() @ B.double().T + bias.double()
        err = (got - ref).abs().max().item()
        print("gemm", M, N, K, err, flush=True)
        assert err <= 1e-4 * (ref.abs().max().item() + 1)


def test_surrogate_gradients_match_fp64_checker():
    import torch
    import mlp_oracle
    from blackhole_amd import surrogate as S
    sur = S.Surrogate(width=64, depth=3, max_batch=128, seed=7)
    rng = np.random.default_rng(5)
    n = 128
    rays = np.stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n), rng.uniform(0, 0.998, n),
                     np.cos(np.radians(rng.uniform(5, 85, n)))], axis=1).astype(np.float32)
    feats = S.Surrogate.features(torch.from_numpy(rays).cuda())
    T = rng.normal(0, 1, (n, 16)).astype(np.float32)
    W = (rng.uniform(0, 1, (n, 16)) > 0.3).astype(np.float32)
    W[:, 12:] = 0
    gpu = sur.gradients(feats, torch.from_numpy(T).cuda(), torch.from_numpy(W).cuda())
    params = sur.get_params().astype(np.float64)
    loss, ref = mlp_oracle.loss_and_grad(params, sur.layer_shapes(), feats.cpu().numpy().astype(np.float64),
                                         T.astype(np.float64), W.astype(np.float64))
    cos = float(np.dot(gpu, ref) / (np.linalg.norm(gpu) * np.linalg.norm(ref)))
    rel = float(np.linalg.norm(gpu - ref) / np.linalg.norm(ref))
    print("grad check: cos", cos, "rel", rel, "loss", loss, flush=True)
    assert cos > 0.999 and rel < 0.03
    # per-layer check (BF16 activations bound the element-wise agreement)
    off = 0
    for (o, i) in sur.layer_shapes():
        k = o * i + o
        gl, rl = gpu[off:off + k], ref[off:off + k]
        assert np.linalg.norm(gl - rl) <= 0.05 * np.linalg.norm(rl) + 1e-7
        off += k


def test_surrogate_forward_matches_checker_and_training_reduces_loss():
    import torch
    import mlp_oracle
    from blackhole_amd import surrogate as S
    sur = S.Surrogate(width=128, depth=4, max_batch=1024, seed=11)
    rng = np.random.default_rng(6)
    n = 1024
    rays = np.stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n), rng.uniform(0, 0.998, n),
                     np.cos(np.radians(rng.uniform(5, 85, n)))], axis=1).astype(np.float32)
    feats = S.Surrogate.features(torch.from_numpy(rays).cuda())
    y = sur.forward(feats).cpu().numpy()
    yr, _, _ = mlp_oracle.forward(sur.get_params().astype(np.float64), sur.layer_shapes(),
                                  feats.cpu().numpy().astype(np.float64))
    assert np.abs(y - yr).max() <= 0.05 * (np.abs(yr).max() + 1e-3)
    # smooth target: fit r = |alpha, beta| / 20 in column 0
    T = np.zeros((n, 16), np.float32)
    T[:, 0] = np.hypot(rays[:, 0], rays[:, 1]) / 20.0
    W = np.zeros((n, 16), np.float32)
    W[:, 0] = 1.0
    Tt, Wt = torch.from_numpy(T).cuda(), torch.from_numpy(W).cuda()
    first = sur.train_step(feats, Tt, Wt, 1e-3)[0]
    for _ in range(300):
        last = sur.train_step(feats, Tt, Wt, 1e-3)[0]
    print("loss", first, "->", last, flush=True)
    assert last < 0.05 * first
