"""The two fast thin-disk paths at full size (16 frames of 2048 x 2048,
67.1 M rays) against the plain FP64 DP45 images of the same frames: the
network-free Mino-time solver + integrator fallback, and the surrogate
pipeline (network, exact refinement, fallback, 1 % audit).

Size-independent properties only (the oracle takes minutes per frame):
per frame, mean relative intensity error over pixels with
I_ref > 1e-6 max(I_ref) at most 1e-6, hit classes differing on at most 1e-6
of the pixels, shadow (horizon-class) area identical to within 2 pixels;
the audit sees no class difference and a mean error below 1e-6; every ray
both images call escaped carries an escape direction: the network-free
solve's exact one (within 0.01 rad on average of the integrator's, which the
integrator records where its last step ends, beyond r_escape), and in the
surrogate pipeline the network's (flagged) within 0.05 rad on average.  The plain
images themselves equal the oracle bit for bit on sampled pixels
(tests/test_gpu_parity.py, bench.py's sample)."""
from __future__ import annotations

import os

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "blackhole_amd", "weights", "surrogate_c5.npz")


def _frame_errors(got, ref, cams):
    import torch
    from blackhole_amd import abi
    I_g = got.view(torch.float64)[:, 0]
    I_r = ref.view(torch.float64)[:, 0]
    out = []
    for cam in cams:
        a, b = cam.record_offset, cam.record_offset + cam.width * cam.height
        ir, ig = I_r[a:b], I_g[a:b]
        m = ir > 1e-6 * ir.max()
        rel = float(((ig[m] - ir[m]).abs() / ir[m]).mean())
        cls = int((got[a:b, 112] != ref[a:b, 112]).sum())
        shadow = int((got[a:b, 112] == abi.HIT_HORIZON).sum()) - int((ref[a:b, 112] == abi.HIT_HORIZON).sum())
        out.append((rel, cls, shadow, b - a))
    return out


def test_fast_paths_match_the_integrator_at_full_size():
    import torch
    from blackhole_amd import sur_render, tracer, workloads
    from blackhole_amd import surrogate_train as ST
    w = workloads.c5(2048)
    print("[full c5] 16 x 2048^2 rays", flush=True)
    sc = tracer.upload_scene(w["cams"], w["opts"], w["luts"])
    plain = tracer.alloc_records(sc.n_records)
    tracer.launch(sc, plain)
    M = sur_render.MinoRenderer(sc)
    M.run()
    torch.cuda.synchronize()
    print("[full c5] Mino solver: fallback rays", M.fallback_count(), flush=True)
    for k, (rel, cls, shadow, n) in enumerate(_frame_errors(M.out, plain, w["cams"])):
        assert rel <= 1e-6 and cls <= 1e-6 * n and abs(shadow) <= 2, (k, rel, cls, shadow)
    mino_out = M.out
    del M
    R = sur_render.SurrogateRenderer(ST.load(WEIGHTS), sc, sur_render.RenderConfig())
    R.run()
    torch.cuda.synchronize()
    fb, au = R.counts_host()
    print("[full c5] surrogate pipeline: fallback", fb, "audit", au, flush=True)
    for k, (rel, cls, shadow, n) in enumerate(_frame_errors(R.out, plain, w["cams"])):
        assert rel <= 1e-6 and cls <= 1e-6 * n and abs(shadow) <= 2, (k, rel, cls, shadow)
    st = R.audit_stats()
    assert au > 0.005 * sc.n_records and st["class_diff"] == 0 and st["mean_rel"] <= 1e-6
    # escape directions: every ray both images call escaped carries one: the
    # integrator's for fallback rays, the network's (flagged approximate) for
    # the others; the network-free solve's are exact (at r_escape)
    from blackhole_amd import abi

    def angle(rec, sel):
        d, p = rec.view(torch.float64), plain.view(torch.float64)
        u1, u2 = d[sel, 2].clamp(-1.0, 1.0), p[sel, 2].clamp(-1.0, 1.0)
        return torch.arccos((u1 * u2 + (1 - u1 * u1).sqrt() * (1 - u2 * u2).sqrt()
                             * torch.cos(d[sel, 3] - p[sel, 3])).clamp(-1.0, 1.0))

    d = R.out.view(torch.float64)
    esc = (R.out[:, 112] == abi.HIT_ESCAPED) & (plain[:, 112] == abi.HIT_ESCAPED)
    assert bool((torch.isfinite(d[esc, 2]) & torch.isfinite(d[esc, 3])).all())
    net = esc & ((R.out[:, 114] & abi.FLAG_EXIT_NETWORK) != 0)
    a_net = angle(R.out, net)
    dm = mino_out.view(torch.float64)
    esc_m = (mino_out[:, 112] == abi.HIT_ESCAPED) & (plain[:, 112] == abi.HIT_ESCAPED)
    assert bool((torch.isfinite(dm[esc_m, 2]) & torch.isfinite(dm[esc_m, 3])).all())
    a_m = angle(mino_out, esc_m)
    print("[full c5] escape directions:", int(esc.sum()), "escaped rays;", int(net.sum()), "from the network, "
          "mean angle to the integrator's", float(a_net.mean()), "; network-free solve (exact at r_escape)",
          float(a_m.mean()), flush=True)
    assert int(net.sum()) > 0.9 * int(esc.sum()) and float(a_net.mean()) <= 0.05
    assert float(a_m.mean()) <= 0.01
