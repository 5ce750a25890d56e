"""GRMHD volumes on the GPU: the software-filtered sample against a numpy
trilinear sampler with the same address modes (r clamped, theta mirrored,
phi wrapped), and the render launch's disk modulation with the semantics the
reference's GRMHD tests fix (tests/cuda_grmhd_lut_test.cu): zero volume ->
black disk, unit volume -> unchanged, half strength -> 0.25x, alpha = 0 ->
the left frame only, alpha = 1 -> the right frame, use_luts off -> no effect;
plus a non-uniform volume checked pixel by pixel against the sample at each
record's hit point."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from blackhole_amd import grmhd as G
from blackhole_amd import tracer
from blackhole_amd._lib import lib

pytestmark = pytest.mark.gpu


def np_sample(vol, r_min, r_max, xyz):
    """Reference: normalised coordinates, texel centres at (i + 0.5)/n,
    trilinear, clamp (r) / mirror with edge repeat (theta) / wrap (phi)."""
    d, h, w, _ = vol.shape
    out = np.zeros((len(xyz), 4))
    for k, (x, y, z) in enumerate(xyz.astype(np.float64)):
        r = math.sqrt(x * x + y * y + z * z)
        th = math.acos(max(-1.0, min(1.0, z / r))) if r > 1e-6 else 0.0
        ph = math.atan2(y, x) % (2 * math.pi)
        crd = [(r - r_min) / (r_max - r_min) * w - 0.5, th / math.pi * h - 0.5, ph / (2 * math.pi) * d - 0.5]
        base = [math.floor(c) for c in crd]
        frac = [c - b for c, b in zip(crd, base)]

        def idx(i, n, mode):
            if mode == "clamp":
                return min(max(i, 0), n - 1)
            if mode == "wrap":
                return i % n
            m = i % (2 * n)
            return m if m < n else 2 * n - 1 - m
        acc = np.zeros(4)
        for bx in (0, 1):
            for by in (0, 1):
                for bz in (0, 1):
                    wt = (frac[0] if bx else 1 - frac[0]) * (frac[1] if by else 1 - frac[1]) * \
                         (frac[2] if bz else 1 - frac[2])
                    acc += wt * vol[idx(base[2] + bz, d, "wrap"), idx(base[1] + by, h, "mirror"),
                                    idx(base[0] + bx, w, "clamp")]
        out[k] = acc
    return out


def test_sample_matches_trilinear_reference():
    rng = np.random.default_rng(5)
    vol = rng.random((10, 6, 8, 4), dtype=np.float32)  # phi 10, theta 6, r 8
    V = G.Volume(vol)
    try:
        n = 400
        r = rng.uniform(0.5, 60.0, n)
        th = rng.uniform(0.0, math.pi, n)
        th[:20] = rng.uniform(0, 0.05, 20)           # near the pole (mirror)
        ph = rng.uniform(-math.pi, math.pi, n)
        ph[20:40] = rng.uniform(-0.05, 0.05, 20)     # across phi = 0 (wrap)
        xyz = np.stack([r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph), r * np.cos(th)], 1)
        got = G.sample(V.handle, 0, 0.0, 2.0, 40.0, xyz)
        want = np_sample(vol, 2.0, 40.0, xyz.astype(np.float32))
        assert np.abs(got - want).max() <= 2e-5, np.abs(got - want).max()
        # texel centres are exact, a constant volume is constant
        c = np.full((3, 2, 4, 4), 0.375, dtype=np.float32)
        K = G.Volume(c)
        try:
            s = G.sample(K.handle, 0, 0.0, 1.0, 50.0, xyz)
            assert np.abs(s - 0.375).max() <= 1e-6
        finally:
            K.close()
        # continuity across phi = 0 / 2 pi (wrap): symmetric points agree with their mean texel blend
        pts = np.array([[10.0, 1e-4, 0.0], [10.0, -1e-4, 0.0]], dtype=np.float32)
        s = G.sample(V.handle, 0, 0.0, 2.0, 40.0, pts)
        assert np.abs(s[0] - s[1]).max() <= 1e-4
        # blend: alpha = 0 -> left, alpha = 1 -> right, 0.25 in between
        rvol = rng.random((10, 6, 8, 4), dtype=np.float32)
        R = G.Volume(rvol)
        try:
            a0 = G.sample(V.handle, R.handle, 0.0, 2.0, 40.0, xyz)
            a1 = G.sample(V.handle, R.handle, 1.0, 2.0, 40.0, xyz)
            aq = G.sample(V.handle, R.handle, 0.25, 2.0, 40.0, xyz)
            rr = G.sample(R.handle, 0, 0.0, 2.0, 40.0, xyz)
            assert np.array_equal(a0, got) and np.abs(a1 - rr).max() <= 1e-6
            assert np.abs(aq - (got + 0.25 * (rr - got))).max() <= 1e-6
        finally:
            R.close()
    finally:
        V.close()


def _render(p):
    import torch
    fb = tracer.render_launch(p, variant=2)
    torch.cuda.synchronize()
    return fb.cpu().numpy().reshape(-1, 4)


def _upload(left, right):
    lib().bh_upload_lut_textures(0, 0, 0, 0, 0, left, 0, right)


def test_render_modulation_semantics():
    p = tracer.launch_params(width=48, height=32, cam_pos=(0.0, -30.0, 8.0), fov_scale=0.5)
    p.use_luts = 1
    p.grmhd_r_min, p.grmhd_r_max = 1.0, 60.0
    vols = {}
    try:
        _upload(0, 0)
        base = _render(p)
        disk = base[:, 3] == 1.0
        disk &= base[:, 0] > 0.0
        assert disk.sum() > 50
        for name, (rho, u) in {"zero": (0.0, 0.0), "unit": (1.0, 1.0), "half": (0.5, 0.5),
                               "left": (0.8, 0.6)}.items():
            v = np.zeros((4, 4, 4, 4), dtype=np.float32)
            v[..., 0], v[..., 1] = rho, u
            vols[name] = G.Volume(v)
        print("[grmhd] render 48x32", flush=True)
        _upload(vols["zero"].handle, 0)
        out = _render(p)
        assert np.all(out[disk, :3] == 0.0) and np.array_equal(out[~disk], base[~disk])
        _upload(vols["unit"].handle, 0)
        assert np.allclose(_render(p), base, rtol=1e-6, atol=0)
        _upload(vols["half"].handle, 0)
        out = _render(p)
        assert np.allclose(out[disk, :3], 0.25 * base[disk, :3], rtol=1e-6, atol=0)
        # temporal blend: left (0.8, 0.6), right (1, 1)
        _upload(vols["left"].handle, 0)
        left_only = _render(p)
        _upload(vols["left"].handle, vols["unit"].handle)
        p.grmhd_alpha = 0.0
        assert np.array_equal(_render(p), left_only)
        p.grmhd_alpha = 1.0
        assert np.allclose(_render(p), base, rtol=1e-6, atol=0)
        p.grmhd_alpha = 0.5
        out = _render(p)
        assert np.allclose(out[disk, :3], 0.9 * 0.8 * base[disk, :3], rtol=1e-5, atol=0)
        p.use_luts = 0
        assert np.array_equal(_render(p), base)
    finally:
        _upload(0, 0)
        for v in vols.values():
            v.close()


def test_render_samples_at_the_hit_point():
    """A volume varying in r, theta and phi: each disk pixel's scale equals the
    sample at (r_cross, theta = pi/2, phi at the crossing step) of its record."""
    from blackhole_amd import abi
    rng = np.random.default_rng(9)
    vol = rng.uniform(0.2, 1.0, (16, 8, 12, 4)).astype(np.float32)
    p = tracer.launch_params(width=40, height=30, cam_pos=(4.0, -28.0, 9.0), fov_scale=0.5)
    p.use_luts = 1
    p.grmhd_r_min, p.grmhd_r_max = 2.0, 40.0
    V = G.Volume(vol)
    try:
        _upload(0, 0)
        base = _render(p)
        _upload(V.handle, 0)
        out = _render(p)
        rec = tracer.records_to_numpy(tracer.trace_launch_params(p, variant=2))
        disk = rec["hit_class"] == abi.HIT_DISK
        k = np.minimum(rec["n_cross"].astype(np.int64), 3) - 1
        rc = rec["cross_r"][np.arange(len(rec)), np.maximum(k, 0)] * (p.rs / 2)
        ph = rec["phi_exit"]
        xyz = np.stack([rc * np.cos(ph), rc * np.sin(ph), np.zeros_like(rc)], 1)[disk].astype(np.float32)
        s = G.sample(V.handle, 0, 0.0, 2.0, 40.0, xyz)
        scale = np.maximum(s[:, 0], 0) * np.maximum(s[:, 1], 0)
        lit = base[disk, 0] > 0
        ratio = out[disk, 0][lit] / base[disk, 0][lit]
        assert disk.sum() > 50 and np.allclose(ratio, scale[lit], rtol=2e-5)
    finally:
        _upload(0, 0)
        V.close()
