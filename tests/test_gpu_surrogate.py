"""Surrogate MLP kernels on the GPU: MFMA GEMM against numpy, training
gradients against the FP64 oracle (oracle/mlp_oracle.py), and a short
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


@pytest.mark.parametrize("n,width", [(128, 64), (1536, 64)])
def test_surrogate_gradients_match_fp64_oracle(n, width):
    """n = 128: one row chunk; n = 1536: the weight and bias gradients split
    over three 512-row chunks and summed in order."""
    import torch
    import mlp_oracle
    from blackhole_amd import surrogate as S
    sur = S.Surrogate(width=width, depth=3, max_batch=n, seed=7)
    rng = np.random.default_rng(5)
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


def test_surrogate_forward_matches_oracle_and_training_reduces_loss():
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


@pytest.mark.parametrize("depth", [6, 1])
def test_fused_render_matches_layerwise_forward_oracle_and_decode(depth):
    """The fused render kernel's raw network outputs (confidence output and the
    first ln r output, recovered from the records) against the layer-by-layer
    forward and the FP64 oracle; the decode (g, intensity, flags, lists)
    against the oracle's disk formulas on the host.  Depth 1 is the network
    with no hidden 256 x 256 layer (layer 0 straight into the output layer)."""
    import torch
    import mlp_oracle
    import oracle_lib as OL
    from blackhole_amd import abi, analysis, sur_render, tracer
    from blackhole_amd import surrogate as S
    print("fused render, depth", depth, flush=True)
    sur = S.Surrogate(width=256, depth=depth, max_batch=4096, seed=21)
    cams = [tracer.camera(0.3, 60.0, 100.0, 48, 40, 20.0), tracer.camera(0.9, 17.0, 100.0, 32, 32, 20.0)]
    cams[0].record_offset = 0
    cams[1].record_offset = 48 * 40 + 100  # a gap of 100 records
    n = cams[1].record_offset + 32 * 32
    opts = tracer.default_options(integrator=abi.INTEGRATOR_DP45, disk_enabled=1, disk_r_out=20.0,
                                  nu_obs=2.417989242e17, disk_T_max=1.0e7)
    luts = tracer.disk_luts(cams, 20.0)
    sc = tracer.upload_scene(cams, opts, luts)
    R = sur_render.SurrogateRenderer(sur, sc, sur_render.RenderConfig(ring_band=0.25, audit_every=7, refine=False))
    R.out.fill_(0xAB)
    R.render()
    torch.cuda.synchronize()
    rec = tracer.records_to_numpy(R.out)
    gap = R.out[48 * 40:48 * 40 + 100].cpu().numpy()
    assert (gap == 0xAB).all(), "records in the gap between cameras were written"
    rays, ids = [], []
    for c in cams:
        A, B = analysis.pixel_grid(c)
        m = A.size
        rays.append(np.stack([A.reshape(-1), B.reshape(-1), np.full(m, c.spin), np.full(m, np.cos(c.theta_obs))], 1))
        ids.append(c.record_offset + np.arange(m))
    rays = np.concatenate(rays).astype(np.float32)
    ids = np.concatenate(ids)
    feats = S.Surrogate.features(torch.from_numpy(rays).cuda())
    y_layer = sur.forward(feats).cpu().numpy()
    F = feats.cpu().numpy().astype(np.float64)
    y_ref, _, _ = mlp_oracle.forward(sur.get_params().astype(np.float64), sur.layer_shapes(), F)
    r = rec[ids]
    assert ((r["flags"] & abi.FLAG_SURROGATE) != 0).all()
    conf = r["confidence"].astype(np.float64)
    scale = np.abs(y_ref[:, S.CONF]).max() + 1e-3
    print("conf fused-ref", np.abs(conf - y_ref[:, S.CONF]).max(), "layer-ref",
          np.abs(y_layer[:, S.CONF] - y_ref[:, S.CONF]).max(), "scale", scale, flush=True)
    assert np.abs(conf - y_ref[:, S.CONF]).max() <= 0.05 * scale
    assert np.abs(conf - y_layer[:, S.CONF]).max() <= 0.05 * scale
    has = r["n_cross"] >= 1
    lnr = np.log(r["cross_r"][has, 0]) - 3.0 * F[has, 14]
    sc0 = np.abs(y_ref[:, S.LOGR1]).max() + 1e-3
    assert has.sum() > 0
    assert np.abs(lnr - y_ref[has, S.LOGR1]).max() <= 0.05 * sc0
    # escape direction of escaped records: u and atan2(sin phi, cos phi) outputs
    esc = r["hit_class"] == abi.HIT_ESCAPED
    assert esc.sum() > 0
    su = np.abs(y_ref[:, S.ESC_U]).max() + 1e-3
    assert np.abs(r["u_exit"][esc] - y_ref[esc, S.ESC_U]).max() <= 0.05 * su
    phi_ref = np.arctan2(y_ref[esc, S.ESC_SINPHI], y_ref[esc, S.ESC_COSPHI])
    big = np.hypot(y_ref[esc, S.ESC_SINPHI], y_ref[esc, S.ESC_COSPHI]) > 0.2  # angle defined
    dphi = np.angle(np.exp(1j * (r["phi_exit"][esc] - phi_ref)))
    assert np.abs(dphi[big]).max() <= 0.1, np.abs(dphi[big]).max()
    # decode: intensity of disk records from the oracle's disk formulas
    L = OL.lib()
    for k, cam in enumerate(cams):
        m = (ids >= cam.record_offset) & (ids < cam.record_offset + cam.width * cam.height)
        rr = r[m]
        r_isco = tracer.isco(cam.spin)
        lut = np.ascontiguousarray(luts[k])
        A, B = analysis.pixel_grid(cam)
        xi = -A.reshape(-1) * np.sin(cam.theta_obs)
        for q in np.nonzero(rr["hit_class"] == abi.HIT_DISK)[0][:200]:
            c = int(rr["n_cross"][q]) - 1
            rc = rr["cross_r"][q, c]
            assert r_isco <= rc <= 20.0
            g = L.orc_disk_g(rc, cam.spin, xi[q])
            assert abs(rr["cross_g"][q, c] - g) <= 1e-12 * abs(g)
            Fv = L.orc_lut_linear(OL.dptr(lut), lut.size, r_isco, 20.0, rc)
            I = L.orc_disk_intensity(Fv, g, opts.nu_obs, opts.disk_T_max)
            assert abs(rr["intensity"][q] - I) <= 1e-9 * abs(I) + 1e-300
    # lists: fallback ids are exactly the flagged records; audit ids are unflagged
    fb, au = R.counts_host()
    fb_ids = np.sort(R.fb_ids[:fb].cpu().numpy())
    flagged = np.sort(ids[(r["flags"] & abi.FLAG_FALLBACK) != 0])
    assert np.array_equal(fb_ids, flagged)
    au_ids = R.audit_ids[:au].cpu().numpy()
    assert au > 0 and ((rec[au_ids]["flags"] & abi.FLAG_FALLBACK) == 0).all()
    assert len(np.unique(au_ids)) == au


def test_render_pipeline_fallback_rays_match_integrator():
    """Fallback rays carry the integrator's records bit for bit; the audit
    traces its sample into a compact buffer identical to the full trace."""
    import torch
    from blackhole_amd import abi, sur_render, tracer, workloads
    from blackhole_amd import surrogate as S
    sur = S.Surrogate(width=256, depth=6, max_batch=4096, seed=5)
    W = workloads.c5(64, spins=(0.5, 0.9), inclinations=(30.0,))
    sc = tracer.upload_scene(W["cams"], W["opts"], W["luts"])
    plain = tracer.alloc_records(sc.n_records)
    tracer.launch(sc, plain)
    R = sur_render.SurrogateRenderer(sur, sc, sur_render.RenderConfig(ring_band=1.0, audit_every=5,
                                                                      solve_band=False))
    R.run()
    torch.cuda.synchronize()
    P = tracer.records_to_numpy(plain)
    Sr = tracer.records_to_numpy(R.out)
    fb, au = R.counts_host()
    assert fb > 0 and au > 0
    fb_ids = R.fb_ids[:fb].cpu().numpy()
    a = R.out.cpu().numpy()[fb_ids]
    b = plain.cpu().numpy()[fb_ids]
    assert np.array_equal(a, b), "fallback records differ from the plain integrator's"
    assert ((Sr["flags"][fb_ids] & abi.FLAG_SURROGATE) == 0).all()
    au_ids = R.audit_ids[:au].cpu().numpy()
    comp = R.audit_rec[:au].cpu().numpy()
    assert np.array_equal(comp, plain.cpu().numpy()[au_ids])
    st = R.audit_stats()
    assert st["n"] == au
    I_s, I_p = Sr["intensity"][au_ids], P["intensity"][au_ids]
    m = I_p > 0
    assert abs(st["sum_rel"] - np.sum(np.abs(I_s[m] - I_p[m]) / I_p[m])) <= 1e-9 * max(st["sum_rel"], 1.0)
    assert st["class_diff"] == int((Sr["hit_class"][au_ids] != P["hit_class"][au_ids]).sum())


def test_light_bending_feature_tracks_the_oracle_first_crossing():
    """Feature 14 = ln(r_bel)/3, the light-bending estimate of the first
    equatorial crossing for an observer at 100 M, against the oracle's traced
    first crossing (rays crossing outside r = 4).  A flipped image-plane sign
    or a missing observer-distance term raises the error above 0.1."""
    import math
    import torch
    import oracle_lib as OL
    from blackhole_amd import abi, analysis, tracer
    from blackhole_amd import surrogate as S
    for th, limit in ((20.0, 0.01), (60.0, 0.03)):
        cam = tracer.camera(0.0, th, 100.0, 17, 17, 18.0)
        opts = tracer.default_options(integrator=abi.INTEGRATOR_DP45, disk_enabled=0, max_crossings=3,
                                      max_steps=20000)
        rec = OL.trace([cam], opts)
        A, B = analysis.pixel_grid(cam)
        rays = np.stack([A.reshape(-1), B.reshape(-1), np.zeros(A.size), np.full(A.size, math.cos(cam.theta_obs))],
                        1).astype(np.float32)
        f = S.Surrogate.features(torch.from_numpy(rays).cuda()).cpu().numpy()
        m = (rec["n_cross"] >= 1) & (rec["cross_r"][:, 0] > 4.0)
        err = np.abs(3.0 * f[m, 14] - np.log(rec["cross_r"][m, 0]))
        flat = np.abs(3.0 * f[m, 15] - np.log(rec["cross_r"][m, 0]))
        print("theta", th, "light-bending", err.mean(), "flat", flat.mean(), flush=True)
        assert err.mean() < limit and err.mean() < 0.2 * flat.mean()
