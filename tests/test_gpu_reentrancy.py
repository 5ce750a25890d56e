"""Reentrancy of the launch entry points: launches from two host threads on
two streams at once must give the records of the same launches made one
after the other.

The reference keeps its launch parameters in __constant__ globals filled by a
separate upload (src/cuda/kernel_launch.cu:25-94), so two concurrent launches
there overwrite each other's parameters.  Here parameters travel in per
(device, stream) blocks (bh_trace_image_plane and relatives) and the
surrogate keeps its render scratch per (device, stream).
"""
from __future__ import annotations

import os
import threading

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run_threads(fns):
    errs = []

    def wrap(f):
        try:
            f()
        except BaseException as e:  # re-raised in the main thread
            errs.append(e)

    ts = [threading.Thread(target=wrap, args=(f,)) for f in fns]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    if errs:
        raise errs[0]


def test_trace_image_plane_two_threads_two_streams():
    import torch
    from blackhole_amd import tracer, workloads
    w_disk = workloads.c1(96, 80)
    w_vol = workloads.c2(64, 64, grid=(64, 32, 16))
    scenes = [tracer.upload_scene(w["cams"], w["opts"], luts=w.get("luts"), cells=w.get("cells"), vol=w.get("vol"))
              for w in (w_disk, w_vol)]
    serial = []
    for sc in scenes:
        out = tracer.alloc_records(sc.n_records)
        out.fill_(0)
        tracer.launch(sc, out)
        torch.cuda.synchronize()
        serial.append(out.clone())
    print("[reentrancy] serial records done", flush=True)
    streams = [torch.cuda.Stream() for _ in scenes]  # default priority
    outs = [[tracer.alloc_records(sc.n_records) for _ in range(3)] for sc in scenes]
    for row in outs:
        for o in row:
            o.fill_(0)
    torch.cuda.synchronize()

    def worker(k):
        def f():
            with torch.cuda.stream(streams[k]):
                for o in outs[k]:
                    tracer.launch(scenes[k], o, stream=streams[k])
            streams[k].synchronize()
        return f

    _run_threads([worker(0), worker(1)])
    torch.cuda.synchronize()
    print("[reentrancy] concurrent launches done", flush=True)
    for k in range(2):
        for o in outs[k]:
            assert torch.equal(o, serial[k]), f"scene {k}: concurrent records differ from the serial launch"


def test_surrogate_render_two_threads_two_streams():
    import torch
    from blackhole_amd import sur_render, tracer, workloads
    from blackhole_amd import surrogate_train as ST
    sur = ST.load(os.path.join(ROOT, "blackhole_amd", "weights", "surrogate_c5.npz"))
    ws = [workloads.c5(48, spins=(0.5,), inclinations=(60.0,)),
          workloads.c5(64, spins=(0.9, 0.998), inclinations=(17.0, 85.0))]
    scenes = [tracer.upload_scene(w["cams"], w["opts"], w["luts"]) for w in ws]
    cfg = sur_render.RenderConfig(overlap_audit=False)

    def render(R):
        R.run(audit=True)

    serial = []
    for sc in scenes:
        R = sur_render.SurrogateRenderer(sur, sc, cfg)
        render(R)
        torch.cuda.synchronize()
        serial.append((R.out.clone(), R.counts.clone(), R.stats.clone()))
    print("[reentrancy] serial surrogate renders done", flush=True)
    streams = [torch.cuda.Stream() for _ in scenes]
    renderers = []
    for k, sc in enumerate(scenes):
        with torch.cuda.stream(streams[k]):
            renderers.append([sur_render.SurrogateRenderer(sur, sc, cfg) for _ in range(2)])
    torch.cuda.synchronize()

    def worker(k):
        def f():
            with torch.cuda.stream(streams[k]):
                for R in renderers[k]:
                    render(R)
            streams[k].synchronize()
        return f

    _run_threads([worker(0), worker(1)])
    torch.cuda.synchronize()
    print("[reentrancy] concurrent surrogate renders done", flush=True)
    for k in range(2):
        out, counts, stats = serial[k]
        for R in renderers[k]:
            assert torch.equal(R.out, out), f"scene {k}: concurrent surrogate records differ"
            assert torch.equal(R.counts, counts)
            # audit sums are accumulated with atomics: equal up to summation order
            assert torch.allclose(R.stats, stats, rtol=1e-12, atol=0.0)
