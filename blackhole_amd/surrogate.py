"""Python binding of the surrogate C ABI (include/bhk_surrogate.h).

The network, its training and its inference all run in libbhk.so
(BF16 MFMA kernels); this module only moves device pointers and host
parameter vectors across the ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

FEATURES, IN_PAD, OUT = 16, 32, 16
(LOGR1, G1, LOGR2, G2, LOGR3, G3, NCROSS, CAPTURE, ESC_U, ESC_SINPHI, ESC_COSPHI, CONF) = range(12)


RING_BINS = 256
RING_FLOATS = 2 + 3 * RING_BINS


class RenderParams(C.Structure):
    """BH_SurrogateRenderParams (fallback criteria and audit rate)."""
    _fields_ = [("ring_band", C.c_float), ("edge_tol", C.c_float), ("conf_max", C.c_float),
                ("ambiguity", C.c_float), ("audit_every", C.c_int32), ("refine", C.c_int32),
                ("keep_factor", C.c_float), ("pad0", C.c_int32)]


class SurrogateConfig(C.Structure):
    _fields_ = [("width", C.c_int32), ("depth", C.c_int32), ("max_batch", C.c_int32),
                ("pad0", C.c_int32), ("seed", C.c_uint64)]


def _sigs(L):
    if getattr(L, "_bhk_sur", False):
        return L
    vp, fp = C.c_void_p, C.POINTER(C.c_float)
    L.bh_surrogate_create.argtypes = [C.POINTER(SurrogateConfig)]
    L.bh_surrogate_create.restype = vp
    L.bh_surrogate_destroy.argtypes = [vp]
    L.bh_surrogate_param_count.argtypes = [vp]
    L.bh_surrogate_param_count.restype = C.c_int64
    L.bh_surrogate_get_params.argtypes = [vp, fp]
    L.bh_surrogate_set_params.argtypes = [vp, fp]
    L.bh_surrogate_features.argtypes = [vp, C.c_int64, vp, vp]
    L.bh_surrogate_forward.argtypes = [vp, vp, C.c_int, vp, vp]
    L.bh_surrogate_train_step.argtypes = [vp, vp, vp, vp, C.c_int, C.c_float, fp, vp]
    L.bh_surrogate_gradients.argtypes = [vp, vp, vp, vp, C.c_int, fp, vp]
    L.bh_surrogate_targets.argtypes = [vp, vp, C.c_int64, vp, vp, fp, vp]
    L.bh_mlp_gemm_nt.argtypes = [C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, vp]
    from . import abi
    L.bh_surrogate_ring_tables.argtypes = [C.POINTER(abi.BH_Camera), C.c_int, fp]
    L.bh_surrogate_render.argtypes = [vp, C.POINTER(abi.BH_Camera), C.c_int, C.POINTER(abi.BH_TraceOptions),
                                      vp, C.c_int, vp, C.POINTER(RenderParams), vp, vp, vp, C.c_int64,
                                      vp, vp, C.c_int64, vp]
    L.bh_surrogate_audit.argtypes = [vp, vp, vp, C.c_int64, vp, C.c_double, vp, vp]
    for n in ("bh_surrogate_ring_tables", "bh_surrogate_render", "bh_surrogate_audit",
              "bh_surrogate_get_params", "bh_surrogate_set_params", "bh_surrogate_features",
              "bh_surrogate_forward", "bh_surrogate_train_step", "bh_surrogate_gradients", "bh_mlp_gemm_nt",
              "bh_surrogate_targets"):
        getattr(L, n).restype = C.c_int
    L._bhk_sur = True
    return L


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class Surrogate:
    """Device-resident MLP: 16 features -> depth x width SiLU -> 16 outputs."""

    def __init__(self, width: int = 256, depth: int = 6, max_batch: int = 16384, seed: int = 1):
        self.L = _sigs(lib())
        cfg = SurrogateConfig(width, depth, max_batch, 0, seed)
        self.h = self.L.bh_surrogate_create(C.byref(cfg))
        if not self.h:
            check(-1, "bh_surrogate_create")
        self.width, self.depth, self.max_batch = width, depth, max_batch

    def close(self):
        if self.h:
            self.L.bh_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_params(self) -> int:
        return int(self.L.bh_surrogate_param_count(self.h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        check(self.L.bh_surrogate_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), "get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        check(self.L.bh_surrogate_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), "set_params")

    def layer_shapes(self):
        """[(out, in_pad)] per layer, in parameter-vector order (W then b)."""
        shapes = []
        for l in range(self.depth + 1):
            shapes.append((OUT if l == self.depth else self.width, IN_PAD if l == 0 else self.width))
        return shapes

    @staticmethod
    def features(rays4):
        """rays4: torch float32 cuda [n, 4] = (alpha, beta, spin, cos theta_obs)."""
        import torch
        n = rays4.shape[0]
        out = torch.empty((n, IN_PAD), dtype=torch.float32, device=rays4.device)
        L = _sigs(lib())
        check(L.bh_surrogate_features(_ptr(rays4.contiguous()), n, _ptr(out), _stream()), "features")
        return out

    def forward(self, feats):
        import torch
        n = feats.shape[0]
        out = torch.empty((n, OUT), dtype=torch.float32, device=feats.device)
        for i in range(0, n, self.max_batch):
            j = min(n, i + self.max_batch)
            check(self.L.bh_surrogate_forward(self.h, _ptr(feats[i:j]), j - i, _ptr(out[i:j]), _stream()),
                  "forward")
        return out

    def train_step(self, feats, targets, weights, lr: float) -> tuple[float, float]:
        loss = np.zeros(2, dtype=np.float32)
        check(self.L.bh_surrogate_train_step(self.h, _ptr(feats), _ptr(targets), _ptr(weights),
                                             feats.shape[0], lr, loss.ctypes.data_as(C.POINTER(C.c_float)),
                                             _stream()), "train_step")
        return float(loss[0]), float(loss[1])

    def gradients(self, feats, targets, weights) -> np.ndarray:
        g = np.zeros(self.n_params, dtype=np.float32)
        check(self.L.bh_surrogate_gradients(self.h, _ptr(feats), _ptr(targets), _ptr(weights), feats.shape[0],
                                            g.ctypes.data_as(C.POINTER(C.c_float)), _stream()), "gradients")
        return g


LOSS_WEIGHTS = (1.0, 0.1, 0.1, 0.1, 0.05)  # ln r, g, count, capture, escape


def targets_from_records(records_u8, feats, weights=LOSS_WEIGHTS):
    """Device targets / weights [n, 16] from a uint8 [n, 128] record tensor and
    the rays' features [n, 32] (ln r targets are residuals against 3 x feature 14,
    the logarithm of the light-bending estimate of the first crossing radius)."""
    import torch
    n = records_u8.shape[0]
    T = torch.empty((n, OUT), dtype=torch.float32, device=records_u8.device)
    W = torch.empty_like(T)
    w5 = (C.c_float * 5)(*weights)
    L = _sigs(lib())
    check(L.bh_surrogate_targets(_ptr(records_u8), _ptr(feats), n, _ptr(T), _ptr(W), w5, _stream()), "targets")
    return T, W


def gemm_nt(A_bf16, Bt_bf16, bias=None):
    """C = A Bt^T (+ bias) on MFMA; A [M, K], Bt [N, K] torch bfloat16 cuda (K % 32 == 0)."""
    import torch
    M, K = A_bf16.shape
    N = Bt_bf16.shape[0]
    C_ = torch.empty((M, N), dtype=torch.float32, device=A_bf16.device)
    L = _sigs(lib())
    check(L.bh_mlp_gemm_nt(M, N, K, _ptr(A_bf16), K, _ptr(Bt_bf16), K,
                           _ptr(bias) if bias is not None else None, _ptr(C_), N, _stream()), "gemm_nt")
    return C_
