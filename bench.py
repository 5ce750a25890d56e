#!/usr/bin/env python3
"""Throughput benchmark of the MI355X Kerr tracer (one JSON line on stdout).

Default workload (the surrogate comparison, ``--workload c5``): 16 frames of
2048 x 2048 pixels over alpha, beta in [-20, 20] M from r_obs = 100 M, spins
{0, 0.5, 0.9, 0.998} x inclinations {17, 40, 60, 85} deg, thin Novikov-Thorne
disk r_ISCO -> 20 M observed at 1 keV (T_max 1e7 K).  A step is one pass of a
method over all 16 frames (67.1 M rays), inputs resident in HBM:

  plain        the plain GPU integrator: FP64 Dormand-Prince 5(4), rtol 1e-8,
               atol 1e-10 (k_trace_fp64<DP45>); its images are the FP64
               reference (bitwise equal to the oracle's exit states, checked
               on a sample below);
  surrogate    the fused BF16-MFMA surrogate over every pixel (k_sur_render),
               each ray outside the photon-ring band then solved exactly from
               the Mino-time integrals of motion with the network's radii as
               starting points (k_mino_disk), the integrator over the
               fallback rays (ring band, unresolved), and the 1 % audit traced
               and compared.
Reported beside them: the network alone (no refinement; its accuracy) and the
network-free Mino-time solver + fallback (bh_trace_disk_mino).

Accuracy of the images against the plain ones: mean relative intensity error
over pixels with I_ref > 1e-6 max(I_ref), photon-ring (shadow) diameter
error, hit-class disagreements.  ``value`` is the rays/s of surrogate+fallback
when it meets the gate (mean relative intensity error <= 1e-3 and
ring-diameter error <= 0.1 %) and beats the plain integrator, else the plain
integrator's.

Other workloads (one JSON line each, same contract):
  --workload c4  the batched form: 256 cameras, spin-bucket tori, one
                 persistent launch replayed from a hipGraph, autotuned shape;
  --workload c3  the 4096^2 precision form (FP32 + FP64 islands) against FP64;
  --workload c2  the volumetric FP64 DP45 tracer on the c2 frame (1024^2,
                 a = 0.94, 17 deg, torus 256 x 128 x 64, 230 GHz);
  --workload c1  the 512^2 FP64 RK4 thin-disk frame against the CPU oracle
                 over the full frame.

The CPU baseline (``cpu_baseline``) is the FP64 oracle on min(16, affinity)
host threads: one warm-up pass and three timed passes over a bounded sample
of the workload, median reported, with the CPU model and affinity count.

Every roofline is priced against MI355X's published peak (``peak``, ``frac``);
after the workload the library's micro-kernels measure the device's
attainable FP64 / FP32 FMA, BF16 MFMA and copy rates (``measured_peaks``),
and each roofline also carries ``peak_measured`` and ``frac_measured``
(``--no-peaks`` skips them).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5|c4|c3|c2|c1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rays/s at ≤1e-3 rel intensity err: surrogate+fallback vs plain GPU DP45"
# HBM bytes per ray of the c5 kernels, measured with rocprofv3 --pmc FETCH_SIZE and
# --pmc WRITE_SIZE on this workload (bench.py --copy-calibration: FETCH_SIZE reads a 512 MiB
# copy as 256 MiB, so fetches are scaled x 2; WRITE_SIZE reads it exactly).
TRAFFIC_PER_RAY = {
    "k_sur_render": (2 * 3270.0 + 8422133.0) * 1024 / 67108864,              # 128.6 B: the record + lists
    # 223.6 B: record + crossing slots + the register spills of the three-waves-per-SIMD build
    "k_trace_fp64_dp45_disk": (2 * 670622.0 + 13315925.0) * 1024 / 67108864,
    # c2 at 1024 x 1024 (FP64): 379.3 B per ray, of which 128 B the record
    "k_trace_fp64_dp45_volume_c2": (2 * 83945.0 + 220529.0) * 1024 / 1048576,
}
TRAFFIC_NOTE = "rocprofv3 FETCH_SIZE (x2, copy-calibrated) + WRITE_SIZE per dispatch, per ray x rays"
FP64_VECTOR_PEAK_TFLOPS = 78.6
BF16_DENSE_PEAK_TFLOPS = 2500.0
WEIGHTS = os.path.join(ROOT, "blackhole_amd", "weights", "surrogate_c5.npz")


def cpu_threads() -> int:
    """Host threads the CPU baseline uses: the GPU's share of the host (16),
    never more than this process may run on."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def host_cpu() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count(), "cpu_model": model}


def time_cpu(fn, reps: int = 3):
    """One warm-up call, then `reps` timed calls; returns (result of the
    warm-up, warm-up s, list of timed s, median s)."""
    import statistics
    t0 = time.perf_counter()
    out = fn()
    warm = time.perf_counter() - t0
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return out, warm, times, statistics.median(times)


def cpu_baseline(w, gpu_rec_np_fn, ids, label: str, reps: int = 3):
    """Time the FP64 CPU oracle on the listed pixels (1 warm-up + `reps`
    passes, median) on cpu_threads() threads, and check the GPU records at
    those pixels against it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from parity_util import compare, run_oracle
    threads = cpu_threads()
    ref, warm, times, med = time_cpu(lambda: run_oracle(oracle_lib, w, ids=ids, nthreads=threads), reps)
    par = compare(gpu_rec_np_fn(ids), ref)
    return cpu_record(ids.size, threads, label, warm, times, med), par


def cpu_record(n: int, threads: int, label: str, warm: float, times, med: float) -> dict:
    return {
        "value": n / med,
        "unit": "rays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{label} ({n} rays; 1 warm-up + {len(times)} timed passes, median {med:.2f} s on {threads} "
                  f"threads; FP64 oracle, -O3 -march=native)",
        "warmup_s": warm, "timed_s": times, **host_cpu(),
    }


def drift_summary(rec) -> dict:
    """Conserved-quantity drift of a traced image (device records): the null
    condition and the polar (Carter) first integral at each ray's exit state,
    |v_r^2 - R(r)| / (r^2 + a^2)^2 and |v_u^2 - U(u)| / (|eta| + a^2 + xi^2).
    E and L are exact by construction (constants of the integrated form)."""
    import torch
    d = rec.view(torch.float64)
    m = rec[:, 112] != 0
    dn, dc = d[:, 11][m], d[:, 12][m]
    return {"rays": int(m.sum()), "max_null": float(dn.max()), "mean_null": float(dn.mean()),
            "max_carter": float(dc.max()), "mean_carter": float(dc.mean()),
            "frac_null_gt_1e-9": float((dn > 1e-9).double().mean()),
            "frac_carter_gt_1e-9": float((dc > 1e-9).double().mean())}


def time_steps(fn, steps: int, warmup: int):
    """Warm up, then time exactly `steps` calls bracketed by synchronisations;
    returns (wall ms per step, event ms per step on the current stream)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) * 1000.0 / steps, e0.elapsed_time(e1) / steps


def frame_accuracy(sur_rec, plain_rec, cams):
    """Per-frame and pooled accuracy of the surrogate records against the
    plain ones, on the device (records are uint8 [n, 128] tensors).  The
    photon-ring diameter of both images comes from the same shadow-edge
    estimator (analysis.ring_diameter_errors), which also checks the plain
    image's ring against the closed-form Bardeen curve; the older
    area-equivalent diameter of the captured region is reported beside it."""
    import math
    import torch
    from blackhole_amd import abi, analysis
    I_s = sur_rec.view(torch.float64)[:, 0]
    I_p = plain_rec.view(torch.float64)[:, 0]
    c_s = sur_rec[:, 112]
    c_p = plain_rec[:, 112]
    frames = []
    tot_rel, tot_n, tot_abs, tot_ref = 0.0, 0, 0.0, 0.0
    for cam in cams:
        a, b = cam.record_offset, cam.record_offset + cam.width * cam.height
        ip, is_ = I_p[a:b], I_s[a:b]
        m = ip > 1e-6 * ip.max()
        rel = (is_[m] - ip[m]).abs() / ip[m]
        pix = (cam.alpha_max - cam.alpha_min) / cam.width * (cam.beta_max - cam.beta_min) / cam.height
        d_s = 2.0 * math.sqrt(float((c_s[a:b] == abi.HIT_HORIZON).sum()) * pix / math.pi)
        d_p = 2.0 * math.sqrt(float((c_p[a:b] == abi.HIT_HORIZON).sum()) * pix / math.pi)
        ring = analysis.ring_diameter_errors(c_s[a:b].cpu().numpy(), c_p[a:b].cpu().numpy(), cam)
        f = {
            "spin": cam.spin, "theta_deg": round(math.degrees(cam.theta_obs), 3),
            "mean_rel": float(rel.mean()), "max_rel": float(rel.max()),
            "frac_rel_gt_1e-2": float((rel > 1e-2).double().mean()),
            "flux_weighted": float((is_ - ip).abs().sum() / ip.sum()),
            "ring_diameter_rel_err": ring["rel_err"],
            "ring_measured_on": ring["measured_on"], "ring_bins": ring["ring_bins"],
            "ring_diameter_M": ring["diameter_ref"],
            "plain_ring_vs_bardeen_rel": ring.get("ref_vs_curve_rel"),
            "shadow_area_diameter_rel_err": abs(d_s - d_p) / d_p,
            "class_disagreements": int((c_s[a:b] != c_p[a:b]).sum()),
        }
        frames.append(f)
        tot_rel += float(rel.sum())
        tot_n += int(m.sum())
        tot_abs += float((is_ - ip).abs().sum())
        tot_ref += float(ip.sum())
    return {
        "mean_rel": tot_rel / max(tot_n, 1),
        "flux_weighted": tot_abs / tot_ref,
        "max_frame_mean_rel": max(f["mean_rel"] for f in frames),
        "max_ring_diameter_rel_err": max(f["ring_diameter_rel_err"] for f in frames),
        "max_plain_ring_vs_bardeen_rel": max((f["plain_ring_vs_bardeen_rel"] for f in frames
                                              if f["plain_ring_vs_bardeen_rel"] is not None), default=None),
        "frames_ring_visible": sum(1 for f in frames if f["ring_measured_on"] == "ring"),
        "max_shadow_area_diameter_rel_err": max(f["shadow_area_diameter_rel_err"] for f in frames),
        "class_disagreements": sum(f["class_disagreements"] for f in frames),
        "frames": frames,
    }


def exit_direction_accuracy(rec, plain_rec):
    """Escape directions of a pipeline's records against the plain FP64
    records, on the device: over the rays both call escaped, the share whose
    record carries a direction (finite u_exit, phi_exit: the exact solve's at
    r_escape, the integrator's for fallback rays, or -- flagged
    BH_FLAG_EXIT_NETWORK -- the network's where the solve gives none) and the
    angle between the two directions (u = cos theta, phi) in radians.  The
    plain integrator records its direction where its last step ends (r_exit
    200 - 213 M), so the angle includes the turn between r_escape and that
    radius."""
    import torch
    from blackhole_amd import abi
    d_all = rec.view(torch.float64)
    esc = (plain_rec[:, 112] == abi.HIT_ESCAPED) & (rec[:, 112] == abi.HIT_ESCAPED)
    fin = esc & torch.isfinite(d_all[:, 2]) & torch.isfinite(d_all[:, 3])
    net = fin & ((rec[:, 114] & abi.FLAG_EXIT_NETWORK) != 0)
    n_esc, n_fin, n_net = int(esc.sum()), int(fin.sum()), int(net.sum())
    out = {"escaped_rays": n_esc, "with_direction": n_fin, "with_direction_frac": n_fin / max(n_esc, 1),
           "direction_from_network": n_net}

    def angles(sel):
        d, p = d_all[sel], plain_rec.view(torch.float64)[sel]
        u1, u2 = d[:, 2].clamp(-1.0, 1.0), p[:, 2].clamp(-1.0, 1.0)
        c = u1 * u2 + (1.0 - u1 * u1).sqrt() * (1.0 - u2 * u2).sqrt() * torch.cos(d[:, 3] - p[:, 3])
        ang = torch.arccos(c.clamp(-1.0, 1.0))
        samp = ang[:: max(1, ang.numel() // 4000000)]
        return {"mean_rad": float(ang.mean()), "median_rad": float(samp.median()),
                "p99_rad": float(torch.quantile(samp, 0.99)), "max_rad": float(ang.max())}

    if n_fin:
        out["angle_to_plain"] = angles(fin)
    if n_net:
        out["angle_to_plain_network_directions"] = angles(net)
    return out


def device_frame_flops(rec):
    """flops.frame_flops of a DP45 frame from the device records (uint8
    [n, 128]: steps, attempts and volume samples are int32 words 26, 27, 31;
    hit class and crossing count bytes 112, 113)."""
    import torch
    from blackhole_amd import abi, flops
    w = rec.view(torch.int32)
    steps = float(w[:, 26].double().sum())
    attempts = float(w[:, 27].double().sum())
    cls = rec[:, 112]
    n_valid = float((cls != 0).sum())
    ncross = float(rec[:, 113].double().sum())
    disk = float((cls == abi.HIT_DISK).sum())
    samples = float(w[:, 31].double().sum())
    f = (attempts * flops.DP45_ATTEMPT + steps * flops.PROJECTION + n_valid * flops.SETUP
         + ncross * flops.CROSSING + disk * flops.DISK_HIT)
    if samples > 0:
        f += samples * flops.VOLUME_SAMPLE + (steps - samples) * flops.VOLUME_STEP_NOSAMPLE
    return f, attempts


def bench_c5(args):
    import numpy as np
    import torch
    from blackhole_amd import abi, flops, sur_render, tracer, workloads
    from blackhole_amd import surrogate_train as ST

    w = workloads.c5(args.size)
    sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])
    n = sc.n_records
    plain = tracer.alloc_records(n)
    stream = torch.cuda.current_stream()

    # ---- plain GPU DP45 -------------------------------------------------------
    plain_ms, plain_ev_ms = time_steps(lambda: tracer.launch(sc, plain, stream), args.steps, args.warmup)
    tracer.launch(sc, plain, stream)
    torch.cuda.synchronize()
    plain_kernel_ms = tracer.last_kernel_ms()

    # ---- the network alone (no refinement): its own accuracy, and the fused
    # MFMA kernel timed by itself --------------------------------------------------
    sur = ST.load(WEIGHTS)
    floor_I = 1e-6 * float(plain.view(torch.float64)[:, 0].max())
    R0 = sur_render.SurrogateRenderer(sur, sc, sur_render.RenderConfig(refine=False))
    _, render_ev_ms = time_steps(R0.render, args.steps, 1)
    R0.run(audit=False)
    torch.cuda.synchronize()
    fb0, _ = R0.counts_host()
    acc0 = frame_accuracy(R0.out, plain, w["cams"])
    del R0
    torch.cuda.empty_cache()

    # ---- surrogate + exact refinement + fallback + 1 % audit ----------------------
    R = sur_render.SurrogateRenderer(sur, sc, sur_render.RenderConfig())
    R.floor_I = floor_I
    sur_ms, sur_ev_ms = time_steps(R.run, args.steps, args.warmup)
    _, refine_render_ev_ms = time_steps(R.render, args.steps, 1)
    # the network pass as it runs in the pipeline (refine mode: the decode leaves
    # g and the intensity to the solve); the roofline's kernel time
    _, net_ev_ms = time_steps(R._network, args.steps, 1)
    R.run()
    torch.cuda.synchronize()
    fb, au = R.counts_host()
    audit = R.audit_stats()
    acc = frame_accuracy(R.out, plain, w["cams"])
    esc_dir = exit_direction_accuracy(R.out, plain)
    refined = (R.out[:, 114] & abi.FLAG_REFINED) != 0
    n_kept = int(((R.out[:, 114] & abi.FLAG_NETWORK_KEPT) != 0).sum())  # network records kept unsolved
    iters = R.out.view(torch.int32)[:, 26]
    it_sur = float(iters[refined].double().mean())
    del R
    torch.cuda.empty_cache()

    # ---- network-free: the Mino-time solver over every pixel + fallback -------------
    M = sur_render.MinoRenderer(sc)
    mino_ms, mino_ev_ms = time_steps(M.run, args.steps, args.warmup)
    _, mino_solve_ev_ms = time_steps(M.solve, args.steps, 1)
    M0 = sur_render.MinoRenderer(sc, exit_directions=False)  # crossings and classes only
    _, mino_solve_noexit_ms = time_steps(M0.solve, args.steps, 1)
    del M0
    M.run()
    torch.cuda.synchronize()
    fbm = M.fallback_count()
    accm = frame_accuracy(M.out, plain, w["cams"])
    esc_dir_m = exit_direction_accuracy(M.out, plain)
    it_mino = float(M.out.view(torch.int32)[:, 26].double().mean())
    del M
    torch.cuda.empty_cache()

    if args.copy_calibration:
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def passes(a):
        return a["mean_rel"] <= 1e-3 and a["max_ring_diameter_rel_err"] <= 1e-3

    plain_rate = n / (plain_ms / 1000.0)
    sur_rate = n / (sur_ms / 1000.0)
    mino_rate = n / (mino_ms / 1000.0)
    gate = passes(acc)
    use_sur = gate and sur_rate > plain_rate

    # algorithmic FLOPs of the plain frame (counting oracle per step x executed attempts)
    # and of the network (dense MACs x 2 per ray)
    f_plain, attempts = device_frame_flops(plain)
    dims = [(16, 256)] + [(256, 256)] * (sur.depth - 1) + [(256, 16)]
    f_ray = 2.0 * sum(a * b for a, b in dims)
    f_sur = f_ray * n
    roof_plain = {
        "bound": "valu", "achieved": f_plain / (plain_kernel_ms / 1000.0) / 1e12,
        "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
        "traffic": TRAFFIC_PER_RAY["k_trace_fp64_dp45_disk"] * n if args.size == 2048 else None,
        "traffic_source": TRAFFIC_NOTE, "algorithmic_bytes": 128.0 * n,
        "kernel": "k_trace_fp64<DP45, no volume>", "kernel_ms": plain_kernel_ms, "flop_per_launch": f_plain,
    }
    roof_plain["frac"] = roof_plain["achieved"] / roof_plain["peak"]
    roof_sur = {
        "bound": "mfma", "achieved": f_sur / (net_ev_ms / 1000.0) / 1e12,
        "peak": BF16_DENSE_PEAK_TFLOPS, "unit": "TFLOP/s",
        "traffic": TRAFFIC_PER_RAY["k_sur_render"] * n if args.size == 2048 else None,
        "traffic_source": TRAFFIC_NOTE, "algorithmic_bytes": 128.0 * n,
        "kernel": "k_sur_render (BF16 MFMA, FP32 accumulation)", "kernel_ms": net_ev_ms,
        "kernel_timing": "HIP events around the pipeline's network pass (refine mode) on its stream, "
                         "with its two small set-up kernels",
        "flop_per_launch": f_sur, "flop_per_ray": f_ray,
    }
    roof_sur["frac"] = roof_sur["achieved"] / roof_sur["peak"]
    summ = lambda a: {k: v for k, v in a.items() if k != "frames"}  # noqa: E731

    res = {
        "metric": METRIC,
        "value": sur_rate if use_sur else plain_rate,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sur_ms if use_sur else plain_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+f64" if use_sur else "f64",
        "data": "synthetic (analytic cameras and disk; surrogate weights trained on 2^24 traced rays)",
        "config": {
            "workload": f"c5: 16 frames {args.size}x{args.size} (spin 0/0.5/0.9/0.998 x incl. 17/40/60/85 deg), "
                        f"alpha,beta in [-20,20] M, r_obs 100 M, thin NT disk r_isco-20 M at 1 keV",
            "rays": n,
            "method_of_value": "surrogate+fallback (network, exact Mino-time refinement, integrator fallback, "
                               "1 % audit)" if use_sur else "plain GPU DP45 (surrogate+fallback misses the gate)",
        },
        "roofline": roof_sur if use_sur else roof_plain,
        "plain_dp45": {
            "rays_per_s": plain_rate, "ms_per_step": plain_ms, "event_ms_per_step": plain_ev_ms,
            "attempts_mean": attempts / n, "roofline": roof_plain,
        },
        "surrogate_fallback": {
            "rays_per_s": sur_rate, "ms_per_step": sur_ms, "event_ms_per_step": sur_ev_ms,
            "speedup_vs_plain": sur_rate / plain_rate,
            "render_kernel_ms": net_ev_ms, "render_plus_refine_ms": refine_render_ev_ms,
            "network_alone_render_ms": render_ev_ms,
            "fallback_rays": fb, "fallback_frac": fb / n, "audit_rays": au, "newton_iters_per_ray": it_sur,
            "network_kept_rays": n_kept, "network_kept_frac": n_kept / n,
            "gate_pass": gate, "accuracy_vs_plain_fp64": summ(acc),
            "escape_direction_vs_plain_fp64": esc_dir,
            "audit": {k: audit[k] for k in ("n", "mean_rel", "max_rel", "class_diff", "flux_weighted")},
            "roofline": roof_sur,
        },
        "network_alone": {
            "note": "fused network + decode without the refinement (its own fallback criteria)",
            "fallback_rays": fb0, "gate_pass": passes(acc0), "accuracy_vs_plain_fp64": summ(acc0),
        },
        "mino_fallback": {
            "note": "network-free: Mino-time crossing solver over every pixel (with the exact escape direction of "
                    "escaped rays) + integrator fallback",
            "rays_per_s": mino_rate, "ms_per_step": mino_ms, "event_ms_per_step": mino_ev_ms,
            "solve_kernel_ms": mino_solve_ev_ms, "speedup_vs_plain": mino_rate / plain_rate,
            "solve_kernel_ms_without_escape_directions": mino_solve_noexit_ms,
            "fallback_rays": fbm, "newton_iters_per_ray": it_mino, "gate_pass": passes(accm),
            "accuracy_vs_plain_fp64": summ(accm),
            "escape_direction_vs_plain_fp64": esc_dir_m,
        },
        "learned_part_net_effect": {
            "note": "the surrogate pipeline (network + exact refinement from its radii + fallback + audit) "
                    "against the network-free pipeline (exact solve from a cubic Hermite start + fallback) "
                    "at the same accuracy gate; ratio < 1 is a net loss of the learned stage",
            "surrogate_pipeline_ms": sur_ms, "network_free_pipeline_ms": mino_ms,
            "ratio_network_free_ms_over_surrogate_ms": mino_ms / sur_ms,
            "newton_iters_from_network": it_sur, "newton_iters_from_hermite": it_mino,
            "refine_from_network_ms": refine_render_ev_ms - net_ev_ms,
            "escape_directions": "the surrogate pipeline's escaped rays carry the network's direction "
                                 "(flagged); the network-free pipeline computes the exact one at r_escape "
                                 "(included in network_free_pipeline_ms, not in solve_from_hermite_ms)",
            "solve_from_hermite_ms": mino_solve_noexit_ms,
            "solve_from_hermite_with_escape_directions_ms": mino_solve_ev_ms,
        },
        "ceiling_at_full_bf16_peak": {
            "network_ms_at_2500_tflops": f_sur / (BF16_DENSE_PEAK_TFLOPS * 1e12) * 1000.0,
            "speedup_vs_plain_if_network_alone_at_peak": plain_ms / (f_sur / (BF16_DENSE_PEAK_TFLOPS * 1e12) * 1000.0),
            "note": "the 6 x 256 SiLU network costs 0.672 MFLOP per ray; even at 100 % of the dense BF16 peak and "
                    "with nothing else in the pipeline it cannot reach 10x the plain integrator on this workload",
        },
        "plain_drift": drift_summary(plain),
        "frames": acc["frames"],
    }
    if not args.no_cpu_baseline:
        rng = np.random.default_rng(20261203)
        ids = np.sort(rng.choice(n, size=min(args.cpu_sample, n), replace=False)).astype(np.int64)
        ids_t = torch.from_numpy(ids).cuda()

        def gpu_rec(sel):
            return tracer.records_to_numpy(plain[ids_t])

        cb, par = cpu_baseline(w, gpu_rec, ids,
                               f"{ids.size} seeded random pixels of the 16 c5 frames, plain DP45 FP64")
        res["cpu_baseline"] = cb
        res["plain_gpu_vs_cpu"] = plain_rate / cb["value"]
        res["parity_vs_oracle_on_sample"] = par
    return res


def bench_c4(args):
    """Batched form: 256 cameras (48.7 M rays) of mixed resolution, spin,
    inclination and distance, thin disk + one torus per spin bucket, in one
    persistent launch; the launch shape autotuned beforehand (excluded), the
    frame replayed from its hipGraph in the timed region."""
    import numpy as np
    import torch
    from blackhole_amd import flops, tracer, workloads

    w = workloads.c4()
    sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])
    vset = tracer.VolumeSetDevice(w["cells_list"], w["vols"], w["cam_vol"])
    n = sc.n_records
    out = tracer.alloc_records(n)
    fr = tracer.Frame(sc, out, vset)
    t0 = time.perf_counter()
    tune = fr.autotune(sample_stride=16)
    tune_s = time.perf_counter() - t0
    stream = torch.cuda.current_stream()
    graph_ms, graph_ev = time_steps(lambda: fr.graph_launch(stream), args.steps, args.warmup)
    eager_ms, eager_ev = time_steps(lambda: fr.launch(stream), args.steps, 0)
    f, attempts = device_frame_flops(out)
    achieved = f / (graph_ev / 1000.0) / 1e12
    samples = float(out.view(torch.int32)[:, 31].double().sum())
    res = {
        "metric": METRIC, "value": n / (graph_ms / 1000.0), "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": graph_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded cameras and spin-bucket tori)",
        "config": {"workload": "c4: 256 cameras (side logU 64-1024, a U[0,0.998], incl U[5,85] deg, r_obs "
                               "U[50,1000]), thin disk + 8 spin-bucket tori 128x64x32, DP45 FP64, 230 GHz, "
                               "one persistent launch, hipGraph replay", "rays": n},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_VECTOR_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_trace_fp64<DP45, volume, volume set>", "kernel_ms": graph_ev,
                     "flop_per_launch": f},
        "autotune": {"blocks_per_cu": tune.blocks_per_cu, "refill_batch": tune.refill_batch,
                     "occupancy_limit": tune.occupancy_limit, "sample_rays": tune.sample_rays,
                     "best_sample_ms": tune.sample_ms, "tuning_s": tune_s},
        "graph_vs_eager": {"graph_ms": graph_ms, "eager_ms": eager_ms, "graph_event_ms": graph_ev,
                           "eager_event_ms": eager_ev},
        "frame_stats": {"attempts_mean": attempts / n, "samples_mean": samples / n,
                        "volume_GBps_algorithmic": samples * flops.VOLUME_SAMPLE_BYTES / (graph_ev / 1000.0) / 1e9,
                        "volume_GBps_stored_layout": samples * flops.VOLUME_SAMPLE_BYTES_STORED
                        / (graph_ev / 1000.0) / 1e9},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        from parity_util import compare
        rng = np.random.default_rng(20261204)
        # a c4 ray costs the oracle ~3x a c5 ray (every camera's rays cross the
        # torus): half the c5 sample keeps one pass near 6 s on 16 threads
        m = min(args.cpu_sample // 2, n)
        ids = np.sort(rng.choice(n, size=m, replace=False)).astype(np.int64)
        offs = np.array([c.record_offset for c in w["cams"]])
        cam_of = np.searchsorted(offs, ids, side="right") - 1
        gpu = tracer.records_to_numpy(out[torch.from_numpy(ids).cuda()])
        threads = cpu_threads()
        buckets = []
        for b in range(len(w["vols"])):
            sel = np.nonzero(np.isin(cam_of, [i for i, v in enumerate(w["cam_vol"]) if v == b]))[0]
            if sel.size:
                buckets.append((b, sel))

        def trace_sample():
            return [(sel, oracle_lib.trace(w["cams"], w["opts"], luts=w["luts"], lut_n=w["luts"].shape[1],
                                           vol=w["vols"][b], cells=w["cells_list"][b].reshape(-1), ids=ids[sel],
                                           nthreads=threads)) for b, sel in buckets]

        refs, warm, times, med = time_cpu(trace_sample)
        pars = [compare(gpu[sel], ref) for sel, ref in refs]
        res["cpu_baseline"] = cpu_record(m, threads, f"{m} seeded random pixels of the 256 c4 cameras, FP64 oracle "
                                                     f"per spin bucket", warm, times, med)
        res["plain_gpu_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
        res["parity_vs_oracle_on_sample"] = {
            "n": sum(p["n"] for p in pars), "class_mismatch": sum(p["class_mismatch"] for p in pars),
            "bitwise_exit": min(p["bitwise_exit"] for p in pars),
            "max_rel_intensity": max(p["max_rel_intensity"] for p in pars),
            "max_rel_tau": max(p["max_rel_tau"] for p in pars)}
    fr.close()
    return res


def bench_c3(args):
    """Precision form: the 4096^2 a = 0.998, 85 deg frame (thin disk + torus,
    DP45 rtol 1e-6) in FP32 with FP64 islands, against the same frame in
    FP64: rays/s of both and the FP32 image's errors."""
    import math
    import torch
    from blackhole_amd import abi, analysis, tracer, workloads

    size = args.size if args.size != 2048 else 4096
    w = workloads.c3(size, size)
    res_legs = {}
    recs = {}
    for prec in ("fp32", "fp64"):
        w["opts"].precision = abi.PRECISION_FP32 if prec == "fp32" else abi.PRECISION_FP64
        sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"], cells=w["cells"], vol=w["vol"])
        out = tracer.alloc_records(sc.n_records)
        ms, ev = time_steps(lambda: tracer.launch(sc, out), args.steps, args.warmup)
        tracer.launch(sc, out)
        torch.cuda.synchronize()
        k_ms = tracer.last_kernel_ms()
        recs[prec] = out
        res_legs[prec] = {"rays_per_s": sc.n_records / (ms / 1000.0), "ms_per_step": ms, "kernel_ms": k_ms}
    n = recs["fp64"].shape[0]
    acc = frame_accuracy(recs["fp32"], recs["fp64"], w["cams"])
    islands = float((recs["fp32"][:, 114] & abi.FLAG_FP64_ISLAND != 0).double().mean())
    w32 = recs["fp32"].view(torch.int32)
    fp64_frac = float(w32[:, 30].double().sum() / w32[:, 26].double().sum())
    f, _ = device_frame_flops(recs["fp32"])
    achieved = f / (res_legs["fp32"]["kernel_ms"] / 1000.0) / 1e12
    res = {
        "metric": METRIC, "value": res_legs["fp32"]["rays_per_s"], "unit": "rays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res_legs["fp32"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded torus volume, analytic camera)",
        "config": {"workload": f"c3: a=0.998, 85 deg, {size}x{size}, thin disk + torus 256x128x64, DP45 rtol 1e-6 "
                               f"FP32 with FP64 islands (r < 1.2 r_+, turning points within 0.05 M)", "rays": n},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": 157.3, "unit": "TFLOP/s",
                     "frac": achieved / 157.3, "traffic": None, "kernel": "k_trace_fp32<DP45, volume>",
                     "kernel_ms": res_legs["fp32"]["kernel_ms"], "flop_per_launch": f,
                     "note": "FP32 peak assumes packed FP32, which the library is compiled without; scalar FP32 issues at "
                             "the 78.6 TF FP64 rate"},
        "fp32": res_legs["fp32"], "fp64": res_legs["fp64"],
        "fp32_vs_fp64": {k: v for k, v in acc.items() if k != "frames"},
        "rays_with_fp64_islands": islands, "fp64_step_fraction": fp64_frac,
        "critical_curve_band_M": 0.1,
    }
    if not args.no_cpu_baseline:
        import numpy as np
        w["opts"].precision = abi.PRECISION_FP64
        rng = np.random.default_rng(20261205)
        ids = np.sort(rng.choice(n, size=min(n, max(1, n // 16)), replace=False)).astype(np.int64)
        rec64 = tracer.records_to_numpy(recs["fp64"][torch.from_numpy(ids).cuda()])
        cb, par = cpu_baseline(w, lambda sel: rec64, ids,
                               f"{ids.size} seeded random pixels (1/16) of the {size}x{size} c3 frame, FP64")
        res["cpu_baseline"] = cb
        res["fp64_gpu_vs_cpu"] = res_legs["fp64"]["rays_per_s"] / cb["value"]
        res["fp32_gpu_vs_cpu"] = res_legs["fp32"]["rays_per_s"] / cb["value"]
        res["fp64_parity_vs_oracle_on_sample"] = par
    return res


def bench_c1(args):
    """Plain method, small: the 512^2 fixed-step FP64 RK4 frame (a = 0.9,
    60 deg, r_obs 100 M, thin NT disk r_isco -> 20 M) on the GPU, against
    the FP64 CPU oracle over the full frame on the host's cores."""
    import numpy as np
    import torch
    from blackhole_amd import abi, flops, tracer, workloads

    size = args.size if args.size != 2048 else 512
    w = workloads.c1(size, size)
    sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])
    out = tracer.alloc_records(sc.n_records)
    stream = torch.cuda.current_stream()
    ms, ev = time_steps(lambda: tracer.launch(sc, out, stream), args.steps, args.warmup)
    tracer.launch(sc, out, stream)
    torch.cuda.synchronize()
    k_ms = tracer.last_kernel_ms()
    rec = tracer.records_to_numpy(out)
    f = flops.frame_flops(rec, abi.INTEGRATOR_RK4)
    achieved = f / (k_ms / 1000.0) / 1e12
    n = sc.n_records
    res = {
        "metric": METRIC, "value": n / (ms / 1000.0), "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (analytic camera and disk)",
        "config": {"workload": f"c1: plain GPU RK4 FP64, a=0.9, 60 deg, r_obs 100 M, {size}x{size} over "
                               f"alpha,beta in [-15,15] M, h=0.01, <=2000 steps, thin NT disk r_isco-20 M",
                   "rays": n},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_VECTOR_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_trace_fp64<RK4, no volume>", "kernel_ms": k_ms, "flop_per_launch": f},
        "frame_stats": {"steps_mean": float(rec["steps"].mean()),
                        "classes": np.bincount(rec["hit_class"], minlength=6).tolist()},
        "drift": drift_summary(out),
    }
    if not args.no_cpu_baseline:
        ids = np.arange(n, dtype=np.int64)
        cb, par = cpu_baseline(w, lambda sel: rec[sel], ids, f"the full {size}x{size} c1 frame")
        res["cpu_baseline"] = cb
        res["plain_gpu_vs_cpu"] = res["value"] / cb["value"]
        res["parity_vs_oracle_full_frame"] = par
    return res


def bench_c2(args):
    import numpy as np
    import torch
    from blackhole_amd import abi, flops, tracer, workloads

    size = args.size if args.size != 2048 else 1024
    w = workloads.c2(size, size)
    if args.precision == "fp32":
        w["opts"].precision = abi.PRECISION_FP32
    sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"], cells=w["cells"], vol=w["vol"])
    out = tracer.alloc_records(sc.n_records)
    stream = torch.cuda.current_stream()
    ms, _ = time_steps(lambda: tracer.launch(sc, out, stream), args.steps, args.warmup)
    tracer.launch(sc, out, stream)
    torch.cuda.synchronize()
    k_ms = tracer.last_kernel_ms()
    if args.copy_calibration:
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    rec = tracer.records_to_numpy(out)
    f = flops.frame_flops(rec, w["opts"].integrator)
    achieved = f / (k_ms / 1000.0) / 1e12
    res = {
        "metric": METRIC,
        "value": sc.n_records / (ms / 1000.0),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded torus volume, analytic camera)",
        "config": {
            "workload": f"c2: plain GPU DP45 {args.precision.upper()} + volumetric synchrotron, a=0.94, 17 deg, "
                        f"{size}x{size}, rtol 1e-8 atol 1e-10, torus 256x128x64, 230 GHz",
            "rays": sc.n_records,
        },
        "roofline": {
            "bound": "valu", "achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
            "traffic": (TRAFFIC_PER_RAY["k_trace_fp64_dp45_volume_c2"] * sc.n_records
                        if (size == 1024 and args.precision == "fp64") else None),
            "traffic_source": TRAFFIC_NOTE,
            "kernel": f"k_trace_{args.precision}<DP45, volume>", "kernel_ms": k_ms, "flop_per_launch": f,
        },
        "frame_stats": {
            "attempts_mean": float(rec["attempts"].mean()),
            "steps_mean": float(rec["steps"].mean()),
            "samples_mean": float(rec["n_samples"].mean()),
            "volume_GBps_algorithmic": float(rec["n_samples"].astype(np.float64).sum()
                                             * flops.VOLUME_SAMPLE_BYTES / (k_ms / 1000.0) / 1e9),
            "volume_GBps_stored_layout": float(rec["n_samples"].astype(np.float64).sum()
                                               * flops.VOLUME_SAMPLE_BYTES_STORED / (k_ms / 1000.0) / 1e9),
            "classes": np.bincount(rec["hit_class"], minlength=6).tolist(),
        },
        "drift": drift_summary(out),
    }
    if not args.no_cpu_baseline:
        n = sc.n_records
        ids = np.arange(0, n, args.cpu_stride, dtype=np.int64)
        cb, par = cpu_baseline(w, lambda sel: rec[sel], ids,
                               f"every {args.cpu_stride}th pixel of the same {size}x{size} c2 frame")
        res["cpu_baseline"] = cb
        res["parity_vs_oracle_on_sample"] = par
    return res


def measured_peaks() -> dict:
    """The device's attainable peaks from the library's micro-kernels
    (bh_measure_peaks: FMA chains, MFMA chains, a 2 GiB copy)."""
    import ctypes as C
    from blackhole_amd._lib import check, lib
    out = (C.c_double * 4)()
    check(lib().bh_measure_peaks(out, None), "bh_measure_peaks")
    return {"fp64_fma_tflops": out[0], "fp32_fma_scalar_tflops": out[1], "bf16_mfma_tflops": out[2],
            "hbm_copy_gbs": out[3],
            "note": "micro-kernels on this device: 8 FMA chains per lane (FP32 without packed "
                    "instructions), 4 v_mfma_f32_32x32x16_bf16 chains per wave, 2 GiB non-temporal copy (read + write)"}


def attach_measured_peaks(res: dict, pk: dict) -> None:
    """Adds peak_measured / frac_measured beside every roofline's published peak."""
    table = {FP64_VECTOR_PEAK_TFLOPS: pk["fp64_fma_tflops"], 157.3: pk["fp32_fma_scalar_tflops"],
             BF16_DENSE_PEAK_TFLOPS: pk["bf16_mfma_tflops"]}

    def walk(d):
        if isinstance(d, dict):
            if "achieved" in d and d.get("peak") in table:
                d["peak_measured"] = table[d["peak"]]
                d["frac_measured"] = d["achieved"] / d["peak_measured"]
            for v in d.values():
                walk(v)
        elif isinstance(d, list):
            for v in d:
                walk(v)

    walk(res)
    res["measured_peaks"] = pk


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c5", "c4", "c3", "c2", "c1"], default="c5")
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=8_000_000,
                    help="c5: oracle pixels for the CPU baseline (c4 takes half)")
    ap.add_argument("--cpu-stride", type=int, default=1, help="c2: oracle pixel stride")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64", help="c2 only")
    ap.add_argument("--no-peaks", action="store_true", help="skip the peak-measuring micro-kernels")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 512 MiB device copy (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the tracer runs on one GPU only (asked for --gpus {args.gpus}); "
              "rays are independent and would shard by rows or cameras without exchange", file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    res = {"c5": bench_c5, "c4": bench_c4, "c3": bench_c3, "c2": bench_c2, "c1": bench_c1}[args.workload](args)
    if not args.no_peaks:
        attach_measured_peaks(res, measured_peaks())
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
