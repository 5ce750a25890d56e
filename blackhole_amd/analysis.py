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


def curve_distance_points(cam: abi.BH_Camera, alpha: np.ndarray, beta: np.ndarray) -> np.ndarray:
    """Distance (M) of the image-plane points (alpha, beta) from the camera's
    critical curve."""
    ca, cb = critical_curve(cam.spin, cam.theta_obs, 2000)
    pa, pb = np.asarray(alpha, dtype=np.float64).reshape(-1), np.asarray(beta, dtype=np.float64).reshape(-1)
    out = np.empty(pa.size)
    for i in range(0, pa.size, 1 << 14):
        d2 = (pa[i:i + (1 << 14), None] - ca[None, :]) ** 2 + (pb[i:i + (1 << 14), None] - cb[None, :]) ** 2
        out[i:i + (1 << 14)] = np.sqrt(d2.min(axis=1)) if d2.size else np.zeros(0)
    return out


def pixel_alpha_beta(cam: abi.BH_Camera, idx: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(alpha, beta) of the pixel centres with row-major indices idx within the camera."""
    idx = np.asarray(idx, dtype=np.int64)
    px, py = idx % cam.width, idx // cam.width
    da = (cam.alpha_max - cam.alpha_min) / cam.width
    db = (cam.beta_max - cam.beta_min) / cam.height
    return cam.alpha_min + (px + 0.5) * da, cam.beta_max - (py + 0.5) * db


def near_disk_edge(cross_r: np.ndarray, n_cross: np.ndarray, r_in: float, r_out: float, rel: float) -> np.ndarray:
    """True where one of a record's recorded crossing radii lies within a
    relative distance rel of the disk's inner or outer edge (where a
    crossing's radius error decides between a disk hit and a pass)."""
    cr = np.asarray(cross_r, dtype=np.float64).reshape(-1, 3)
    k = np.arange(3)[None, :] < np.asarray(n_cross).reshape(-1, 1)
    near = (np.abs(cr - r_in) <= rel * r_in) | (np.abs(cr - r_out) <= rel * r_out)
    return (near & k).any(axis=1)


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


def curve_centre(a: float, theta_o: float) -> tuple[float, float]:
    """Centroid (alpha, beta) of the closed critical curve's points."""
    ca, cb = critical_curve(a, theta_o, 4000)
    return float(ca.mean()), float(cb.mean())


def _polar(pa: np.ndarray, pb: np.ndarray, centre, nbins: int):
    da, db = pa - centre[0], pb - centre[1]
    rho = np.hypot(da, db)
    psi = np.mod(np.arctan2(db, da), 2.0 * math.pi)
    return rho, np.minimum((psi * (nbins / (2.0 * math.pi))).astype(np.int64), nbins - 1)


def shadow_edge_radii(hit_class: np.ndarray, cam: abi.BH_Camera, nbins: int = 64) -> np.ndarray:
    """Outer edge of the captured region of a hit-class image: in each of
    nbins polar angle bins about the critical curve's centroid, the largest
    distance (M) of a pixel centre whose ray fell into the horizon (NaN where
    a bin has none).  Where no disk lies in front of the curve, rays just
    inside it orbit near the photon sphere and fall in while rays just
    outside leave, so the edge is the photon ring (the critical curve) to
    within a pixel; where the disk covers it, the edge is the inner shadow's."""
    A, B = pixel_grid(cam)
    m = np.asarray(hit_class).reshape(-1) == abi.HIT_HORIZON
    rho, b = _polar(A.reshape(-1)[m], B.reshape(-1)[m], curve_centre(cam.spin, cam.theta_obs), nbins)
    out = np.full(nbins, -np.inf)
    np.maximum.at(out, b, rho)
    out[~np.isfinite(out)] = np.nan
    return out


def curve_radii(cam: abi.BH_Camera, nbins: int = 64) -> np.ndarray:
    """The critical curve's largest distance from its centroid in each polar bin."""
    ca, cb = critical_curve(cam.spin, cam.theta_obs, 20000)
    rho, b = _polar(ca, cb, curve_centre(cam.spin, cam.theta_obs), nbins)
    out = np.full(nbins, -np.inf)
    np.maximum.at(out, b, rho)
    out[~np.isfinite(out)] = np.nan
    return out


def diameter_from_radii(radii: np.ndarray, mask: np.ndarray | None = None) -> float:
    """Area-equivalent diameter 2 sqrt(A / pi) of the star-shaped curve with
    the given per-bin radii, A = sum of r^2 dpsi / 2 over the (masked) bins,
    scaled to the full circle."""
    r = np.asarray(radii, dtype=np.float64)
    ok = np.isfinite(r) if mask is None else (mask & np.isfinite(r))
    if not ok.any():
        return float("nan")
    return 2.0 * math.sqrt(float(np.mean(r[ok] ** 2)))


def ring_diameter_errors(cls_img: np.ndarray, cls_ref: np.ndarray, cam: abi.BH_Camera, nbins: int = 64,
                         visible_px: float = 2.0) -> dict:
    """Photon-ring diameter of an image against a reference image of the
    same camera (both hit-class images), with the same estimator.

    The ring's bins are those where the reference's shadow edge lies within
    visible_px pixels of the closed-form critical curve (the curve is
    visible there); on them the ring diameter of both images is compared,
    and the reference's is cross-checked against the curve's.  Frames where
    the disk hides the whole curve have no ring bins; their shadow outlines
    (every bin with captured pixels in both images) are compared instead."""
    ei = shadow_edge_radii(cls_img, cam, nbins)
    er = shadow_edge_radii(cls_ref, cam, nbins)
    rc = curve_radii(cam, nbins)
    px = max((cam.alpha_max - cam.alpha_min) / cam.width, (cam.beta_max - cam.beta_min) / cam.height)
    vis = np.isfinite(er) & np.isfinite(rc) & (np.abs(er - rc) <= visible_px * px) & np.isfinite(ei)
    both = np.isfinite(ei) & np.isfinite(er)
    sel = vis if vis.any() else both
    d_i, d_r = diameter_from_radii(ei, sel), diameter_from_radii(er, sel)
    out = {"ring_bins": int(vis.sum()), "bins": nbins, "measured_on": "ring" if vis.any() else "shadow",
           "diameter": d_i, "diameter_ref": d_r,
           "rel_err": abs(d_i - d_r) / d_r if sel.any() else 0.0}
    if vis.any():
        d_c = diameter_from_radii(rc, vis)
        out["curve_diameter"] = d_c
        out["ref_vs_curve_rel"] = abs(d_r - d_c) / d_c
    return out
