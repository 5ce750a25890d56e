"""Shared helpers for GPU-vs-oracle parity tests."""
from __future__ import annotations

import math

import numpy as np

from blackhole_amd import abi, tracer


def run_gpu(w: dict) -> np.ndarray:
    """Trace workload dict w on cuda:0 through the C ABI; returns records (numpy)."""
    import torch
    sc = tracer.upload_scene(w["cams"], w["opts"], luts=w.get("luts"), cells=w.get("cells"),
                             vol=w.get("vol"))
    out = tracer.alloc_records(sc.n_records)
    out.fill_(0)
    tracer.launch(sc, out)
    torch.cuda.synchronize()
    return tracer.records_to_numpy(out)


def run_oracle(oracle, w: dict, ids=None, nthreads: int = 0) -> np.ndarray:
    """Oracle records of workload w (all rays, or the listed ids); nthreads 0
    lets the oracle use every hardware thread."""
    cells = w.get("cells")
    return oracle.trace(
        w["cams"], w["opts"], luts=w.get("luts"), lut_n=(w["luts"].shape[-1] if w.get("luts") is not None else 0),
        vol=w.get("vol"), cells=cells.reshape(-1) if cells is not None else None, ids=ids, nthreads=nthreads)


# Radiative values below this compare absolutely: a transfer sum this small
# comes from samples whose emissivity is subnormal (j ~ 1e-310 at
# nu / nu_s ~ 5e7), where neither the oracle nor the GPU keeps relative
# precision (their results differ by up to ~1e-7 relative there).
RADIATIVE_FLOOR = 1e-280


def rel(a, b, floor: float = 1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    s = np.maximum(np.abs(b), floor)
    return np.where(d == 0, 0.0, d / s)


def compare(gpu: np.ndarray, ref: np.ndarray, tol: float = 1e-9) -> dict:
    """Field-by-field comparison; returns summary stats (also used by bench)."""
    same_class = gpu["hit_class"] == ref["hit_class"]
    out = {
        "n": int(ref.size),
        "class_mismatch": int((~same_class).sum()),
        "bitwise_exit": float(np.mean((gpu["r_exit"] == ref["r_exit"]) & (gpu["u_exit"] == ref["u_exit"])
                                      & (gpu["phi_exit"] == ref["phi_exit"]))),
    }
    m = same_class
    for f in ("r_exit", "u_exit", "phi_exit", "intensity", "tau"):
        r = rel(gpu[f][m], ref[f][m], RADIATIVE_FLOOR if f in ("intensity", "tau") else 1e-300)
        out[f"max_rel_{f}"] = float(r.max()) if r.size else 0.0
    for f in ("drift_null", "drift_carter"):
        d = np.abs(gpu[f][m] - ref[f][m])
        out[f"max_abs_{f}_diff"] = float(d.max()) if d.size else 0.0
    out["steps_mismatch"] = int((gpu["steps"][m] != ref["steps"][m]).sum())
    out["attempts_mismatch"] = int((gpu["attempts"][m] != ref["attempts"][m]).sum())
    return out
