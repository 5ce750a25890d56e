"""Algorithmic FLOP per unit of tracer work (FP64 path).

Counted by instantiating the oracle's templates with a counting scalar
(oracle/flopcount.cpp: + - * / sqrt = 1, FMA = 2, transcendentals tallied
apart); tests/test_oracle_pinning.py checks these constants against that
program's output.  The GPU kernel performs the same stepping operations;
per volume sample it evaluates the same formulas with fewer divisions and
its own acos and fractional-power evaluations, which are not counted (the
count is of the algorithm, not of an implementation).
"""
RHS = 31
RK4_STEP = 200          # 3 stage RHS + combine + RHS at the new point + projection
DP45_ATTEMPT = 521      # 6 RHS (FSAL) + stage sums + error norm + controller
PROJECTION = 14         # radial null-constraint projection (accepted DP45 steps)
CROSSING = 161          # Hermite root of u(s) = 0 + disk g
DISK_HIT = 14           # LUT lookup + observed blackbody (+1 transcendental)
VOLUME_SAMPLE = 186     # trilinear fetch (4 fields) + fluid g + Planck + j + a, e + trapezoid segment (+5 transcendentals)
VOLUME_STEP_NOSAMPLE = 7  # affine rate + segment bookkeeping on accepted steps without a sample
SETUP = 65              # image plane -> constants + first RHS

# Algorithmic bytes per trilinear sample: 8 corners x 3 FP32 fields (n_e, Theta_e, |B|).
VOLUME_SAMPLE_BYTES = 96
# Bytes the kernel reads per sample: each corner is one 16-B float4 whose fourth
# channel holds the precomputed K_2(1/Theta_e) / (2 Theta_e^2) ratio.
VOLUME_SAMPLE_BYTES_STORED = 128


def frame_flops(rec, integrator: int) -> float:
    """Algorithmic FLOP of one traced frame from its records (numpy RECORD_DTYPE)."""
    import numpy as np
    steps = rec["steps"].astype(np.float64).sum()
    attempts = rec["attempts"].astype(np.float64).sum()
    n = float((rec["hit_class"] != 0).sum())
    ncross = np.minimum(rec["n_cross"].astype(np.float64), 255).sum()
    disk = float((rec["hit_class"] == 2).sum())
    samples = rec["n_samples"].astype(np.float64).sum()
    if integrator == 1:
        f = attempts * DP45_ATTEMPT + steps * PROJECTION
    else:
        f = steps * RK4_STEP
    vol = 0.0
    if samples > 0:
        vol = samples * VOLUME_SAMPLE + (steps - samples) * VOLUME_STEP_NOSAMPLE
    return f + n * SETUP + ncross * CROSSING + disk * DISK_HIT + vol
