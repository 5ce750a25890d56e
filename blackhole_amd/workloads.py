"""The five tracer workloads (cameras, options and input tables).

c1  plain RK4, a = 0.9, 60 deg, r_obs = 100, 512^2 over alpha, beta in [-15, 15],
    h = 0.01 in s, <= 2000 steps, stop at r < 1.01 r_+ or r > 200, thin NT disk
    r_isco -> 20 M; per pixel hit class, crossing radius, g and blackbody
    intensity (1 keV, T_max 1e7 K); FP64.
c2  DP45 (rtol 1e-8, atol 1e-10) plus volumetric thermal synchrotron at
    230 GHz through a Fishbone-Moncrief torus (r_in 6, r_max 12) x lognormal
    turbulence (sigma 0.5, k^-11/3) on 256 x 128 x 64 over r in [r_+, 50];
    a = 0.94, 17 deg, 1024^2 over alpha, beta in [-20, 20]; FP64.
c3  precision form: a = 0.998, 85 deg, 4096^2, disk + the c2 torus recipe.
c4  batched: 256 cameras (N ~ logU(64, 1024), a ~ U[0, 0.998],
    theta ~ U[5, 85] deg, r_obs ~ U[50, 1000]) from a fixed seed; thin disk
    plus the c2 torus recipe on 128 x 64 x 32 per spin bucket (8 buckets of
    width 0.998/8, torus at the bucket's mid spin), all buckets on one grid
    (r from the horizon of a = 0.998 to 50 M) and one support; DP45 FP64
    at the c2 tolerances, 230 GHz.
c5  surrogate evaluation frames: 2048^2 over alpha, beta in [-20, 20] from
    r_obs = 100 on a 4 x 4 sweep a in {0, 0.5, 0.9, 0.998} x theta in
    {17, 40, 60, 85} deg; the c1 disk (NT, r_isco -> 20 M, 1 keV, T_max
    1e7 K) traced by the plain DP45 integrator at the c2 tolerances.
"""
from __future__ import annotations

import math

import numpy as np

from . import abi
from .tracer import camera, default_options, disk_luts, generate_torus, outer_horizon, pack_cameras, torus_params

KEV_HZ = 2.417989242e17


def c1(width: int = 512, height: int = 512):
    cams = [camera(0.9, 60.0, 100.0, width, height, 15.0)]
    opts = default_options(
        integrator=abi.INTEGRATOR_RK4, precision=abi.PRECISION_FP64, max_steps=2000,
        disk_enabled=1, volume_enabled=0, h_fixed=0.01, horizon_factor=1.01, r_escape=200.0,
        disk_r_out=20.0, nu_obs=KEV_HZ, disk_T_max=1.0e7)
    return dict(cams=cams, opts=opts, luts=disk_luts(cams, 20.0), cells=None, vol=None)


def c2(width: int = 1024, height: int = 1024, grid=(256, 128, 64), seed: int = 0x5EED0002):
    cams = [camera(0.94, 17.0, 100.0, width, height, 20.0)]
    opts = default_options(
        integrator=abi.INTEGRATOR_DP45, precision=abi.PRECISION_FP64, max_steps=20000,
        max_attempts=100000, disk_enabled=0, volume_enabled=1, rtol=1e-8, atol=1e-10,
        h_init=0.01, h_max=0.1, h_vol=0.01, horizon_factor=1.01, r_escape=200.0,
        nu_obs=230.0e9)
    tp = torus_params(0.94, n_r=grid[0], n_th=grid[1], n_ph=grid[2], seed=seed)
    cells, vol = generate_torus(tp)
    return dict(cams=cams, opts=opts, luts=None, cells=cells, vol=vol, torus=tp)


def c4_cameras(n: int = 256, seed: int = 0x5EED0004):
    rng = np.random.default_rng(seed)
    cams = []
    for _ in range(n):
        side = int(round(math.exp(rng.uniform(math.log(64), math.log(1024)))))
        spin = float(rng.uniform(0.0, 0.998))
        inc = float(rng.uniform(5.0, 85.0))
        r_obs = float(rng.uniform(50.0, 1000.0))
        cams.append(camera(spin, inc, r_obs, side, side, 15.0))
    pack_cameras(cams)
    return cams


def c3(width: int = 4096, height: int = 4096, grid=(256, 128, 64), seed: int = 0x5EED0003,
       precision: int = abi.PRECISION_FP32):
    """Precision form: a = 0.998, 85 deg, thin disk plus the c2 torus recipe,
    FP32 DP45 stepping (rtol 1e-6, atol 1e-8) with FP64 islands at 230 GHz."""
    cams = [camera(0.998, 85.0, 100.0, width, height, 20.0)]
    opts = default_options(
        integrator=abi.INTEGRATOR_DP45, precision=precision, max_steps=20000,
        max_attempts=100000, disk_enabled=1, volume_enabled=1, rtol=1e-6, atol=1e-8,
        h_init=0.01, h_max=0.1, h_vol=0.01, horizon_factor=1.01, r_escape=200.0,
        disk_r_out=20.0, nu_obs=230.0e9, disk_T_max=1.0e7,
        island_r_factor=1.2, island_turn_dist=0.05)
    tp = torus_params(0.998, n_r=grid[0], n_th=grid[1], n_ph=grid[2], seed=seed)
    cells, vol = generate_torus(tp)
    return dict(cams=cams, opts=opts, luts=disk_luts(cams, 20.0), cells=cells, vol=vol, torus=tp)


C5_SPINS = (0.0, 0.5, 0.9, 0.998)
C5_INCLINATIONS = (17.0, 40.0, 60.0, 85.0)


def c5_options():
    """Plain DP45 FP64 thin-disk tracing (the surrogate's reference and fallback)."""
    return default_options(
        integrator=abi.INTEGRATOR_DP45, precision=abi.PRECISION_FP64, max_steps=20000,
        max_attempts=100000, disk_enabled=1, volume_enabled=0, rtol=1e-8, atol=1e-10,
        h_init=0.01, h_max=0.1, horizon_factor=1.01, r_escape=200.0, disk_r_out=20.0,
        nu_obs=KEV_HZ, disk_T_max=1.0e7, max_crossings=3)


def c5(size: int = 2048, spins=C5_SPINS, inclinations=C5_INCLINATIONS):
    cams = [camera(a, th, 100.0, size, size, 20.0) for a in spins for th in inclinations]
    pack_cameras(cams)
    return dict(cams=cams, opts=c5_options(), luts=disk_luts(cams, 20.0), cells=None, vol=None)


C4_BUCKETS = 8


def c4(n: int = 256, seed: int = 0x5EED0004, grid=(128, 64, 32), buckets: int = C4_BUCKETS):
    """The batched form: cameras, per-spin-bucket tori sharing one geometry,
    the camera -> bucket map, disk LUTs and options."""
    cams = c4_cameras(n, seed)
    width = 0.998 / buckets
    cam_vol = [min(int(c.spin / width), buckets - 1) for c in cams]
    r_lo = outer_horizon(0.998)
    cells, vols = [], []
    for b in range(buckets):
        tp = torus_params((b + 0.5) * width, n_r=grid[0], n_th=grid[1], n_ph=grid[2], seed=seed + 17 * b)
        tp.r_lo = r_lo
        c, v = generate_torus(tp)
        cells.append(c)
        vols.append(v)
    lo = min(v.sup_r_lo for v in vols)
    hi = max(v.sup_r_hi for v in vols)
    um = max(v.sup_u_max for v in vols)
    for v in vols:
        v.sup_r_lo, v.sup_r_hi, v.sup_u_max = lo, hi, um
    opts = default_options(
        integrator=abi.INTEGRATOR_DP45, precision=abi.PRECISION_FP64, max_steps=20000,
        max_attempts=100000, disk_enabled=1, volume_enabled=1, rtol=1e-8, atol=1e-10,
        h_init=0.01, h_max=0.1, h_vol=0.01, horizon_factor=1.01, r_escape=200.0,
        disk_r_out=20.0, nu_obs=230.0e9, disk_T_max=1.0e7)
    return dict(cams=cams, opts=opts, luts=disk_luts(cams, 20.0), cells_list=cells, vols=vols, cam_vol=cam_vol)
