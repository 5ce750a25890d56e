"""Batched form on the GPU: cameras of several spins sampling their own
spin-bucket torus in one launch (against the oracle, bucket by bucket),
prepared frames, hipGraph replays and the autotuner."""
from __future__ import annotations

import numpy as np
import pytest

from parity_util import compare

pytestmark = pytest.mark.gpu


def _small_c4(n_cams=6, side=24, grid=(32, 16, 8)):
    from blackhole_amd import tracer, workloads
    w = workloads.c4(n=n_cams, grid=grid, buckets=3)
    for c in w["cams"]:
        c.width = c.height = side
    tracer.pack_cameras(w["cams"])
    w["luts"] = tracer.disk_luts(w["cams"], 20.0)
    return w


def test_multi_volume_launch_matches_oracle_per_bucket(oracle):
    import torch
    from blackhole_amd import tracer
    w = _small_c4()
    sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])
    vset = tracer.VolumeSetDevice(w["cells_list"], w["vols"], w["cam_vol"])
    out = tracer.alloc_records(sc.n_records)
    out.fill_(0)
    tracer.launch_multi(sc, vset, out)
    torch.cuda.synchronize()
    gpu = tracer.records_to_numpy(out)
    assert len(set(w["cam_vol"])) >= 2
    for b in sorted(set(w["cam_vol"])):
        idx = [i for i, v in enumerate(w["cam_vol"]) if v == b]
        cams = [w["cams"][i] for i in idx]
        luts = np.ascontiguousarray(w["luts"][idx])
        ref = oracle.trace(cams, w["opts"], luts=luts, lut_n=luts.shape[1], vol=w["vols"][b],
                           cells=w["cells_list"][b].reshape(-1))
        for c in cams:
            sl = slice(c.record_offset, c.record_offset + c.width * c.height)
            st = compare(gpu[sl], ref[sl])
            print("bucket", b, "spin", c.spin, st, flush=True)
            assert st["class_mismatch"] == 0
            assert st["bitwise_exit"] == 1.0
            assert st["max_rel_intensity"] <= 1e-9
            assert (gpu["n_samples"][sl] == ref["n_samples"][sl]).all()


def test_frame_launch_and_graph_replays_match_one_shot():
    import torch
    from blackhole_amd import tracer
    w = _small_c4()
    sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])
    vset = tracer.VolumeSetDevice(w["cells_list"], w["vols"], w["cam_vol"])
    ref = tracer.alloc_records(sc.n_records)
    tracer.launch_multi(sc, vset, ref)
    out = tracer.alloc_records(sc.n_records)
    fr = tracer.Frame(sc, out, vset)
    out.fill_(0)
    fr.launch()
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    for rep in range(3):
        out.fill_(0)  # eager work between replays
        torch.cuda.synchronize()
        fr.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), f"graph replay {rep} differs from the eager launch"
    fr.close()


def test_autotune_returns_valid_shape_and_keeps_results():
    import torch
    from blackhole_amd import tracer
    w = _small_c4()
    sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])
    vset = tracer.VolumeSetDevice(w["cells_list"], w["vols"], w["cam_vol"])
    ref = tracer.alloc_records(sc.n_records)
    tracer.launch_multi(sc, vset, ref)
    out = tracer.alloc_records(sc.n_records)
    fr = tracer.Frame(sc, out, vset)
    r = fr.autotune(sample_stride=4)
    print("tuned", r.blocks_per_cu, r.refill_batch, r.occupancy_limit, r.sample_rays, r.sample_ms, flush=True)
    assert 1 <= r.blocks_per_cu <= r.occupancy_limit
    assert r.refill_batch in (1, 2, 4, 8, 16, 32)
    assert r.sample_rays == (sc.n_records + 3) // 4 and r.sample_ms > 0
    out.fill_(0)
    fr.launch()
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    fr.set_launch_shape(1, 1)
    out.fill_(0)
    fr.graph_launch()
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    fr.close()


def test_c4_full_shape_autotuned_graph_replay_matches_oracle_per_bucket(oracle):
    """The batched form at its real shape: 256 cameras (48.7 M rays) of
    log-uniform size, spin, inclination and distance, thin disk plus eight
    spin-bucket tori of 128 x 64 x 32, autotuned and replayed from its
    hipGraph.  Two replays (eager work between them) must equal the eager
    launch, and a seeded sample of 65 536 pixels per spin bucket must equal
    the oracle: identical classes, bitwise exit state, intensity within 1e-9."""
    import torch
    from blackhole_amd import tracer, workloads
    w = workloads.c4()
    assert len(w["cams"]) == 256 and len(w["vols"]) == 8
    sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])
    vset = tracer.VolumeSetDevice(w["cells_list"], w["vols"], w["cam_vol"])
    n = sc.n_records
    out = tracer.alloc_records(n)
    fr = tracer.Frame(sc, out, vset)
    r = fr.autotune(sample_stride=16)
    print("[c4] rays", n, "tuned", r.blocks_per_cu, r.refill_batch, r.occupancy_limit, r.sample_ms, flush=True)
    out.fill_(0)
    fr.launch()
    torch.cuda.synchronize()
    eager = out.clone()
    assert bool((eager[:, 112] != 0).all())
    for rep in range(2):
        out.fill_(0)
        torch.cuda.synchronize()
        fr.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), f"graph replay {rep} differs from the eager launch"
    print("[c4] graph replays equal the eager launch", flush=True)
    offs = np.array([c.record_offset for c in w["cams"]])
    rng = np.random.default_rng(0xC4)
    for b in range(8):
        cams_b = [i for i, v in enumerate(w["cam_vol"]) if v == b]
        if not cams_b:
            continue
        pool = np.concatenate([np.arange(offs[i], offs[i] + w["cams"][i].width * w["cams"][i].height)
                               for i in cams_b])
        ids = np.sort(rng.choice(pool, size=min(65536, pool.size), replace=False)).astype(np.int64)
        gpu = tracer.records_to_numpy(eager[torch.from_numpy(ids).cuda()])
        ref = oracle.trace(w["cams"], w["opts"], luts=w["luts"], lut_n=w["luts"].shape[1], vol=w["vols"][b],
                           cells=w["cells_list"][b].reshape(-1), ids=ids)
        st = compare(gpu, ref)
        print("[c4] bucket", b, "cameras", len(cams_b), st, flush=True)
        assert st["class_mismatch"] == 0 and st["bitwise_exit"] == 1.0
        assert st["max_rel_intensity"] <= 1e-9 and st["max_rel_tau"] <= 1e-9
        assert (gpu["n_samples"] == ref["n_samples"]).all()
    fr.close()
