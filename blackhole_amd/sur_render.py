"""Image rendering with the surrogate plus integrator fallback and audit.

One frame set (any number of cameras at r_obs = 100 M) goes through:
  1. bh_surrogate_render: every pixel through the fused BF16-MFMA network
     and decode (records, fallback id list, audit id list); with refine set
     (the default) each ray then has its crossings and hit class solved
     exactly from the Mino-time integrals of motion, starting from the
     network's radii; the photon-ring band goes to the integrator (or, with
     solve_band, to the same solver from its own cubic Hermite start), and
     the rays that solver cannot resolve join the fallback list;
  2. bh_trace_pixel_list: the fallback rays traced by the plain integrator
     into the same image (records overwrite the surrogate's);
  3. bh_trace_pixel_list_compact + bh_surrogate_audit: the audit sample
     traced into a compact buffer and compared with the surrogate records.
All steps are asynchronous; list lengths stay on the device.  With
overlap_audit (the default), the network pass leaves the solve to
bh_trace_disk_mino_refine, and after the solve the audit sample is traced
on a second stream beside the integrator fallback: two short persistent
launches of the same kernel, each filling the other's tail.  The stream is
joined back before the comparison.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi, tracer
from . import surrogate as S
from ._lib import check, lib

STAT_NAMES = ("n", "n_cmp", "sum_rel", "max_rel", "class_diff", "sum_abs_dI", "sum_I_ref", "sum_dlnr")


@dataclass
class RenderConfig:
    ring_band: float = 0.25     # M
    edge_tol: float = 2e-3      # relative
    conf_max: float = 0.0       # log10 of the predicted squared error (0: effectively off)
    ambiguity: float = 0.3
    audit_every: int = 100      # 1 % audit
    refine: bool = True         # exact Mino-time crossing solve after the network
    # the solve also takes the ring band (from its own starting points) instead of
    # listing it for the integrator: c5 then has no integrator rays, but near the
    # critical curve the quadrature needs many panels and those waves set the solve's
    # tail (refinement 10.4 against 7.3 ms; pipeline 55.8 against 54.8 ms)
    solve_band: bool = False
    # trace the audit sample on a second stream beside the integrator fallback (c5: 56.4 ms
    # against 57.0 - 57.1 ms in sequence; beside the exact solve instead it was slower)
    overlap_audit: bool = True
    intensity_floor_rel: float = 1e-6  # audit compares pixels with I_ref > floor x max I
    # escape directions of escaped rays: the network's (flagged BH_FLAG_EXIT_NETWORK) by
    # default; True: exact from the refinement's Mino-time solve (BH_TraceOptions.mino_exit = 1;
    # every ray is then solved, keep_factor does not apply)
    exact_exit: bool = False
    # with refine: rays the network places escaped or captured with every predicted crossing
    # beyond keep_factor x disk_r_out keep its record without the exact solve (their intensity
    # is 0 whatever the radii; flagged BH_FLAG_NETWORK_KEPT); 0 solves every ray
    keep_factor: float = 1.2


def ring_tables(cams) -> np.ndarray:
    arr = (abi.BH_Camera * len(cams))(*cams)
    out = np.zeros(len(cams) * S.RING_FLOATS, dtype=np.float32)
    L = S._sigs(lib())
    check(L.bh_surrogate_ring_tables(arr, len(cams), out.ctypes.data_as(C.POINTER(C.c_float))), "ring_tables")
    return out.reshape(len(cams), S.RING_FLOATS)


class SurrogateRenderer:
    """Device buffers and launches of the surrogate + fallback pipeline.  The
    scene's trace options are copied at construction (with this renderer's
    escape-direction mode), so later changes to scene.opts do not reach it."""

    def __init__(self, sur: S.Surrogate, scene: tracer.DeviceScene, cfg: RenderConfig = RenderConfig(),
                 fallback_frac: float = 1.0):
        import torch
        self.torch = torch
        self.sur, self.sc, self.cfg = sur, scene, cfg
        # the scene's options with this renderer's escape-direction mode
        self.opts = abi.BH_TraceOptions.from_buffer_copy(scene.opts)
        self.opts.mino_exit = 1 if cfg.exact_exit else 0
        n = scene.n_records
        self.n = n
        dev = scene.luts.device
        self.out = tracer.alloc_records(n, device=dev)
        # the fallback list holds every ray by default (8 B per pixel), so it
        # cannot overflow; the audit sample is 1 in audit_every of the rest
        self.fb_cap = max(1024, int(n * fallback_frac))
        self.audit_cap = max(4096, 2 * n // max(cfg.audit_every, 1) + 4096) if cfg.audit_every > 0 else 1024
        self.fb_ids = torch.empty(self.fb_cap, dtype=torch.int64, device=dev)
        self.audit_ids = torch.empty(self.audit_cap, dtype=torch.int64, device=dev)
        self.counts = torch.zeros(2, dtype=torch.int64, device=dev)  # fallback, audit
        self.audit_rec = tracer.alloc_records(self.audit_cap, device=dev)
        self.stats = torch.zeros(8, dtype=torch.float64, device=dev)
        self.ring = torch.from_numpy(ring_tables(scene.cams).reshape(-1)).to(dev)
        split = cfg.refine and cfg.overlap_audit and cfg.audit_every > 0
        self.side = torch.cuda.Stream(device=dev) if split else None  # default priority
        mode = 0 if not cfg.refine else (3 if cfg.solve_band else (2 if split else 1))
        self.params = S.RenderParams(cfg.ring_band, cfg.edge_tol, cfg.conf_max, cfg.ambiguity,
                                     cfg.audit_every, mode, cfg.keep_factor, 0)
        self.floor_I = 0.0

    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def render(self) -> None:
        """Surrogate pass (records + lists), with the exact solve when refine is set."""
        self._network()
        if self.params.refine >= 2:
            self.refine()

    def refine(self) -> None:
        """The Mino-time solve of the network's records (refine = 2 or 3 pass)."""
        cam_arr = (abi.BH_Camera * len(self.sc.cams))(*self.sc.cams)
        p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        check(lib().bh_trace_disk_mino_refine(
            cam_arr, len(self.sc.cams), C.byref(self.opts), p(self.sc.luts), self.sc.lut_n, p(self.out),
            p(self.fb_ids), p(self.counts[0:1]), self.fb_cap, self._stream()), "bh_trace_disk_mino_refine")

    def _network(self) -> None:
        cam_arr = (abi.BH_Camera * len(self.sc.cams))(*self.sc.cams)
        L = S._sigs(lib())
        p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        check(L.bh_surrogate_render(
            self.sur.h, cam_arr, len(self.sc.cams), C.byref(self.opts), p(self.sc.luts), self.sc.lut_n,
            p(self.ring), C.byref(self.params), p(self.out), p(self.fb_ids), p(self.counts[0:1]),
            self.fb_cap, p(self.audit_ids), p(self.counts[1:2]), self.audit_cap, self._stream()),
            "bh_surrogate_render")

    def fallback(self) -> None:
        tracer.launch_ids(self.sc, self.fb_ids, self.counts[0:1], self.out)

    def audit(self) -> None:
        self.audit_trace()
        self.audit_compare()

    def audit_trace(self, stream=None) -> None:
        tracer.launch_ids(self.sc, self.audit_ids, self.counts[1:2], self.audit_rec, stream=stream, compact=True)

    def audit_compare(self) -> None:
        L = S._sigs(lib())
        p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        check(L.bh_surrogate_audit(p(self.out), p(self.audit_ids), p(self.counts[1:2]), self.audit_cap,
                                   p(self.audit_rec), self.floor_I, p(self.stats), self._stream()),
              "bh_surrogate_audit")

    def run(self, audit: bool = True) -> None:
        if self.side is not None and audit:
            torch = self.torch
            main = torch.cuda.current_stream()
            self._network()
            self.refine()
            solved = torch.cuda.Event()
            solved.record(main)
            self.side.wait_event(solved)
            self.audit_trace(stream=self.side)
            traced = torch.cuda.Event()
            traced.record(self.side)
            self.fallback()
            main.wait_event(traced)
            self.audit_compare()
            return
        self.render()
        self.fallback()
        if audit:
            self.audit()

    def counts_host(self) -> tuple[int, int]:
        """(fallback, audit) list lengths; raises if a list overflowed (its
        rays would then be missing from the fallback or the audit)."""
        c = self.counts.cpu().numpy()
        fb, au = int(c[0]), int(c[1])
        if fb > self.fb_cap or au > self.audit_cap:
            raise RuntimeError(f"surrogate render: list overflow (fallback {fb} > {self.fb_cap} "
                               f"or audit {au} > {self.audit_cap})")
        return fb, au

    def audit_stats(self) -> dict:
        s = self.stats.cpu().numpy()
        d = dict(zip(STAT_NAMES, (float(x) for x in s)))
        d["mean_rel"] = d["sum_rel"] / max(d["n_cmp"], 1.0)
        d["flux_weighted"] = d["sum_abs_dI"] / max(d["sum_I_ref"], 1e-300)
        return d


class MinoRenderer:
    """Network-free thin-disk images: the Mino-time crossing solver over every
    pixel (bh_trace_disk_mino, with the exact escape direction of escaped rays
    unless exit_directions is False), then the plain integrator over the rays
    the solver leaves (bh_trace_pixel_list into the same records)."""

    def __init__(self, scene: tracer.DeviceScene, exit_directions: bool = True):
        import torch
        self.torch = torch
        self.sc = scene
        # exact escape directions of escaped rays (BH_TraceOptions.mino_exit 0), or none (2)
        self.opts = abi.BH_TraceOptions.from_buffer_copy(scene.opts)
        self.opts.mino_exit = 0 if exit_directions else 2
        n = scene.n_records
        dev = scene.luts.device
        self.out = tracer.alloc_records(n, device=dev)
        self.fb_ids = torch.empty(max(n, 1024), dtype=torch.int64, device=dev)
        self.count = torch.zeros(1, dtype=torch.int64, device=dev)

    def solve(self) -> None:
        tracer.trace_disk_mino(self.sc, self.out, self.fb_ids, self.count, opts=self.opts)

    def fallback(self) -> None:
        tracer.launch_ids(self.sc, self.fb_ids, self.count, self.out)

    def run(self) -> None:
        self.solve()
        self.fallback()

    def fallback_count(self) -> int:
        return int(self.count.cpu()[0])
