"""bh_trace_image_plane_noalloc, the allocation-free variant of
bh_trace_image_plane (launch block and queue head in a static device array),
called through the C ABI: its records must equal bh_trace_image_plane's bit
for bit, on calls in a row whose inputs shrink, and launches beyond its
scratch must be refused before anything runs."""
from __future__ import annotations

import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def _call(fn, sc, out, stream):
    from blackhole_amd import abi
    cam_arr = (abi.BH_Camera * len(sc.cams))(*sc.cams)
    return fn(cam_arr, len(sc.cams), C.byref(sc.opts),
              C.c_void_p(sc.luts.data_ptr()) if sc.luts is not None else None, sc.lut_n,
              C.byref(sc.vol) if sc.vol is not None else None,
              C.c_void_p(out.data_ptr()), C.c_void_p(stream.cuda_stream))


def _both(sc):
    """Records of bh_trace_image_plane and of the allocation-free entry point."""
    import torch
    from blackhole_amd import tracer
    from blackhole_amd._lib import check, lib
    L = lib()
    s = torch.cuda.current_stream()
    ref = tracer.alloc_records(sc.n_records)
    got = tracer.alloc_records(sc.n_records)
    ref.fill_(0)
    got.fill_(0xAB)
    check(_call(L.bh_trace_image_plane, sc, ref, s), "bh_trace_image_plane")
    check(_call(L.bh_trace_image_plane_noalloc, sc, got, s), "bh_trace_image_plane_noalloc")
    torch.cuda.synchronize()
    return ref, got


def test_noalloc_equals_trace_image_plane_on_shrinking_inputs():
    import torch
    from blackhole_amd import tracer, workloads
    from blackhole_amd import abi
    cases = []
    w = workloads.c1(96, 64)
    cases.append(("c1 RK4 disk 96x64", tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])))
    w = workloads.c2(40, 32, grid=(64, 32, 16))
    cases.append(("c2 DP45 volume 40x32", tracer.upload_scene(w["cams"], w["opts"], cells=w["cells"], vol=w["vol"])))
    w = workloads.c5(16, spins=(0.0, 0.9), inclinations=(17.0, 85.0))
    cases.append(("c5 DP45 disk 4 x 16x16", tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])))
    w = workloads.c1(12, 10)
    cases.append(("c1 RK4 disk 12x10", tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])))
    for name, sc in cases:
        print("[noalloc]", name, flush=True)
        ref, got = _both(sc)
        assert torch.equal(got, ref), name
        recs = tracer.records_to_numpy(got)
        assert (recs["hit_class"] > 0).all(), name
    # the largest launch the scratch holds: 540 cameras of 2 x 2
    cams = [tracer.camera(0.1 + 0.8 * (i % 7) / 7.0, 20.0 + (i % 11) * 5.0, 100.0, 2, 2, 15.0) for i in range(540)]
    tracer.pack_cameras(cams)
    opts = workloads.c5_options()
    sc = tracer.upload_scene(cams, opts, luts=tracer.disk_luts(cams, 20.0))
    print("[noalloc] 540 cameras of 2x2", flush=True)
    ref, got = _both(sc)
    assert torch.equal(got, ref)
    assert abi.RECORD_DTYPE.itemsize == 128


def test_noalloc_refuses_launches_beyond_its_scratch():
    import torch
    from blackhole_amd import tracer, workloads
    from blackhole_amd._lib import lib
    cams = [tracer.camera(0.5, 60.0, 100.0, 2, 2, 15.0) for _ in range(541)]
    tracer.pack_cameras(cams)
    sc = tracer.upload_scene(cams, workloads.c5_options(), luts=tracer.disk_luts(cams, 20.0))
    out = tracer.alloc_records(sc.n_records)
    out.fill_(7)
    rc = _call(lib().bh_trace_image_plane_noalloc, sc, out, torch.cuda.current_stream())
    torch.cuda.synchronize()
    assert rc == -1
    msg = lib().bh_last_error().decode()
    assert "exceeds" in msg and "540 cameras" in msg, msg
    assert bool((out == 7).all()), "a refused launch must not write records"
