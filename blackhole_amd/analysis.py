"""Image-level accuracy measures of the tracer (host-side, numpy).

* The critical curve (photon-ring edge) of a distant observer, from the
  spherical photon orbits of radius r in [r_pro, r_retro]
  (criticalImpactParams, src/physics/analytic_kerr_geodesic.h:410-448):
      xi(r)  = (r^2 + a^2)/a - 4 r Delta / (a Delta'),
      eta(r) = 16 r^2 Delta / Delta'^2 - (xi - a)^2,
      alpha = -xi / sin(theta_o),
      beta  = +-sqrt(eta + a^2 cos^2 theta_o - xi^2 cot^2 theta_o).
* The distance of every pixel from that curve (the epsilon band inside which
  hit / escape classification is chaotic and may differ between precisions).
* The shadow diameter, measured the same way on any hit-class image: the
  diameter of the disk with the area of the captured (horizon) region.
* Relative intensity errors against a reference image, over the pixels with
  I_ref > 1e-6 max(I_ref), plus the flux-weighted error sum|dI| / sum I.
"""
from __future__ import annotations

import math

import numpy as np

from . import abi


def photon_orbits(a: float) -> tuple[float, float]:
    pro = 2.0 * (1.0 + math.cos((2.0 / 3.0) * math.acos(-abs(a))))
    retro = 2.0 * (1.0 + math.cos((2.0 / 3.0) * math.acos(abs(a))))
    return pro, retro


def critical_curve(a: float, theta_o: float, n: int = 4000) -> tuple[np.ndarray, np.ndarray]:
    """Closed critical curve (alpha, beta) in M, both beta signs, ordered."""
    if a < 1e-9:
        b = 3.0 * math.sqrt(3.0)
        t = np.linspace(0.0, 2.0 * math.pi, 2 * n, endpoint=False)
        return b * np.cos(t), b * np.sin(t)
    rp, rm = photon_orbits(a)
    # Chebyshev-spaced radii resolve the curve's ends
    r = rp + (rm - rp) * 0.5 * (1.0 - np.cos(np.linspace(0.0, math.pi, n)))
    delta = r * r - 2.0 * r + a * a
    dp = 2.0 * r - 2.0
    xi = (r * r + a * a) / a - 4.0 * r * delta / (a * dp)
    eta = 16.0 * r * r * delta / (dp * dp) - (xi - a) ** 2
    s, c = math.sin(theta_o), math.cos(theta_o)
    b2 = eta + a * a * c * c - xi * xi * (c / s) ** 2
    ok = b2 >= 0.0
    alpha = -xi[ok] / s
    beta = np.sqrt(b2[ok])
    return np.concatenate([alpha, alpha[::-1]]), np.concatenate([beta, -beta[::-1]])


def pixel_grid(cam: abi.BH_Camera) -> tuple[np.ndarray, np.ndarray]:
    """(alpha, beta) of every pixel centre, shape (height, width)."""
    da = (cam.alpha_max - cam.alpha_min) / cam.width
    db = (cam.beta_max - cam.beta_min) / cam.height
    al = cam.alpha_min + (np.arange(cam.width) + 0.5) * da
    be = cam.beta_max - (np.arange(cam.height) + 0.5) * db
    return np.meshgrid(al, be)


def curve_distance(cam: abi.BH_Camera, chunk: int = 1 << 16) -> np.ndarray:
    """Distance (M) of each pixel centre from the critical curve, (h, w)."""
    ca, cb = critical_curve(cam.spin, cam.theta_obs, 2000)
    A, B = pixel_grid(cam)
    pa, pb = A.reshape(-1), B.reshape(-1)
    out = np.empty(pa.size)
    for i in range(0, pa.size, chunk):
        d2 = (pa[i:i + chunk, None] - ca[None, :]) ** 2 + (pb[i:i + chunk, None] - cb[None, :]) ** 2
        out[i:i + chunk] = np.sqrt(d2.min(axis=1))
    return out.reshape(A.shape)


def shadow_diameter_pixels(hit_class: np.ndarray, cam: abi.BH_Camera) -> float:
    """Area-equivalent diameter (M) of the captured region of a hit-class image."""
    pix = (cam.alpha_max - cam.alpha_min) / cam.width * (cam.beta_max - cam.beta_min) / cam.height
    area = float((hit_class == abi.HIT_HORIZON).sum()) * pix
    return 2.0 * math.sqrt(area / math.pi)


def shadow_diameter_closed_form(a: float, theta_o: float) -> float:
    """Area-equivalent diameter (M) of the region inside the critical curve."""
    ca, cb = critical_curve(a, theta_o, 20000)
    area = 0.5 * abs(np.dot(ca, np.roll(cb, -1)) - np.dot(cb, np.roll(ca, -1)))
    return 2.0 * math.sqrt(area / math.pi)


def intensity_errors(I: np.ndarray, I_ref: np.ndarray, floor: float = 1e-6) -> dict:
    I = np.asarray(I, dtype=np.float64).reshape(-1)
    I_ref = np.asarray(I_ref, dtype=np.float64).reshape(-1)
    mx = float(I_ref.max()) if I_ref.size else 0.0
    m = I_ref > floor * mx
    rel = np.abs(I[m] - I_ref[m]) / I_ref[m] if m.any() else np.zeros(1)
    return {
        "n_pixels": int(m.sum()),
        "mean_rel": float(rel.mean()),
        "max_rel": float(rel.max()),
        "p99_rel": float(np.quantile(rel, 0.99)),
        "flux_weighted": float(np.abs(I - I_ref).sum() / max(I_ref.sum(), 1e-300)),
    }
