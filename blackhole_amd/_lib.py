"""Loader for the in-tree HIP library ``libbhk.so`` (C ABI of include/*.h).

There is no CPU fallback: if the library is missing or cannot be loaded, every
entry point raises.  Building happens in ``blackhole_amd.build`` (called by
``__graft_entry__.build()``), never implicitly at import time.
"""
from __future__ import annotations

import ctypes as C
import os

from . import abi

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libbhk.so")

_lib = None

_D = C.c_double
_PD = C.POINTER(C.c_double)


class NativeLibraryMissing(RuntimeError):
    pass


def lib():
    """Load libbhk.so (once) and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 (same
    # SONAME, libamdhip64.so.7).  Loading torch first makes the dynamic loader
    # resolve libbhk.so's dependency to that already-loaded copy instead of
    # mapping a second runtime from /opt/rocm, which would not see the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.bh_version.restype = C.c_int
    L.bh_last_error.restype = C.c_char_p
    L.bh_trace_options_default.argtypes = [C.POINTER(abi.BH_TraceOptions)]
    L.bh_outer_horizon.argtypes = [_D]
    L.bh_outer_horizon.restype = _D
    L.bh_isco_prograde.argtypes = [_D]
    L.bh_isco_prograde.restype = _D
    L.bh_disk_lut_build.argtypes = [_D, _D, _D, C.c_int, _PD]
    L.bh_volume_generate_torus.argtypes = [
        C.POINTER(abi.BH_TorusParams), C.POINTER(C.c_float), C.POINTER(abi.BH_Volume)]
    L.bh_trace_image_plane.argtypes = [
        C.POINTER(abi.BH_Camera), C.c_int, C.POINTER(abi.BH_TraceOptions),
        C.c_void_p, C.c_int, C.POINTER(abi.BH_Volume), C.c_void_p, C.c_void_p,
    ]
    L.bh_trace_image_plane_noalloc.argtypes = L.bh_trace_image_plane.argtypes
    L.bh_trace_image_plane_noalloc.restype = C.c_int
    L.bh_symplectic_integrate.argtypes = [C.c_void_p, C.c_int64, _D, C.c_int32, _D, _D, C.c_void_p]
    L.bh_symplectic_integrate.restype = C.c_int
    L.bh_symplectic_hamiltonian.argtypes = [C.c_void_p, C.c_int64, _D, _D, C.c_void_p, C.c_void_p]
    L.bh_symplectic_hamiltonian.restype = C.c_int
    L.bh_symplectic_integrate_host.argtypes = [_PD, C.c_int64, _D, C.c_int32, _D, _D]
    L.bh_symplectic_integrate_host.restype = C.c_int
    L.bh_trace_pixel_list.argtypes = [
        C.POINTER(abi.BH_Camera), C.c_int, C.POINTER(abi.BH_TraceOptions), C.c_void_p, C.c_int,
        C.POINTER(abi.BH_Volume), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    L.bh_trace_pixel_list.restype = C.c_int
    L.bh_trace_pixel_list_compact.argtypes = L.bh_trace_pixel_list.argtypes
    L.bh_trace_pixel_list_compact.restype = C.c_int
    L.bh_trace_image_plane_multi.argtypes = [
        C.POINTER(abi.BH_Camera), C.c_int, C.POINTER(abi.BH_TraceOptions), C.c_void_p, C.c_int,
        C.POINTER(abi.BH_VolumeSet), C.c_void_p, C.c_void_p]
    L.bh_trace_image_plane_multi.restype = C.c_int
    L.bh_frame_create.argtypes = [
        C.POINTER(abi.BH_Camera), C.c_int, C.POINTER(abi.BH_TraceOptions), C.c_void_p, C.c_int,
        C.POINTER(abi.BH_Volume), C.POINTER(abi.BH_VolumeSet), C.c_void_p]
    L.bh_frame_create.restype = C.c_void_p
    L.bh_frame_launch.argtypes = [C.c_void_p, C.c_void_p]
    L.bh_frame_graph_launch.argtypes = [C.c_void_p, C.c_void_p]
    L.bh_frame_destroy.argtypes = [C.c_void_p]
    L.bh_frame_destroy.restype = None
    L.bh_frame_set_launch_shape.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.bh_frame_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.BH_TuneResult), C.c_void_p]
    for name in ("bh_frame_launch", "bh_frame_graph_launch", "bh_frame_set_launch_shape", "bh_frame_autotune"):
        getattr(L, name).restype = C.c_int
    L.bh_launch_geodesic_kernel.argtypes = [C.c_void_p, C.POINTER(abi.BH_LaunchParams), C.c_int, C.c_void_p]
    L.bh_launch_geodesic_kernel.restype = C.c_int
    L.bh_trace_launch_params.argtypes = [C.POINTER(abi.BH_LaunchParams), C.c_int, C.c_void_p, C.c_void_p]
    L.bh_trace_launch_params.restype = C.c_int
    L.bh_select_kernel_variant.argtypes = []
    L.bh_select_kernel_variant.restype = C.c_int
    L.bh_upload_lut_textures.argtypes = [C.c_ulonglong] * 8
    L.bh_upload_lut_textures.restype = None
    L.bh_upload_bridge_background_texture.argtypes = [C.c_ulonglong]
    L.bh_upload_bridge_background_texture.restype = None
    L.bh_lut_texture_handle.argtypes = [C.c_int]
    L.bh_lut_texture_handle.restype = C.c_ulonglong
    L.bh_trace_rays.argtypes = [C.c_void_p, C.c_int64, C.POINTER(abi.BH_TraceOptions),
                                C.POINTER(abi.BH_Volume), C.c_void_p, C.c_void_p]
    L.bh_trace_rays.restype = C.c_int
    L.bh_last_kernel_ms.restype = C.c_float
    L.bh_trace_disk_mino.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.BH_TraceOptions), C.c_void_p, C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    L.bh_measure_peaks.restype = C.c_int
    L.bh_measure_peaks.argtypes = [C.POINTER(_D), C.c_void_p]
    L.bh_trace_disk_mino.restype = C.c_int
    L.bh_trace_disk_mino_refine.argtypes = L.bh_trace_disk_mino.argtypes
    L.bh_trace_disk_mino_refine.restype = C.c_int
    L.bh_trace_image_plane_pol.argtypes = [
        C.POINTER(abi.BH_Camera), C.c_int, C.POINTER(abi.BH_TraceOptions), C.c_void_p, C.c_int,
        C.POINTER(abi.BH_Volume), C.POINTER(abi.BH_PolParams), C.c_void_p, C.c_void_p, C.c_void_p]
    L.bh_trace_image_plane_pol.restype = C.c_int
    # renderer backend (bh_backend.h)
    L.bhCudaInit.argtypes = [C.c_uint, C.c_int, C.c_int]
    L.bhCudaInit.restype = C.c_void_p
    L.bhCudaRenderFrame.argtypes = [C.c_void_p, C.POINTER(abi.BH_LaunchParams)]
    L.bhCudaResize.argtypes = [C.c_void_p, C.c_uint, C.c_int, C.c_int]
    L.bhCudaRegisterLut.argtypes = [C.c_void_p, C.c_int, C.c_uint, C.c_uint]
    L.bhCudaRegisterBackgroundTexture.argtypes = [C.c_void_p, C.c_uint, C.c_uint]
    L.bhCudaSetVariant.argtypes = [C.c_void_p, C.c_int]
    L.bhCudaGetVariant.argtypes = [C.c_void_p]
    L.bhCudaSynchronize.argtypes = [C.c_void_p]
    for name in ("bhCudaRenderFrame", "bhCudaResize", "bhCudaRegisterLut", "bhCudaRegisterBackgroundTexture",
                 "bhCudaSetVariant", "bhCudaGetVariant", "bhCudaSynchronize"):
        getattr(L, name).restype = C.c_int
    L.bhCudaShutdown.argtypes = [C.c_void_p]
    L.bhCudaShutdown.restype = None
    L.bhCudaFramebuffer.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.bhCudaFramebuffer.restype = C.c_void_p
    L.bhCudaGlTexture.argtypes = [C.c_void_p, C.c_int]
    L.bhCudaGlTexture.restype = C.c_uint
    for name in (
        "bh_version", "bh_trace_options_default", "bh_disk_lut_build",
        "bh_volume_generate_torus", "bh_trace_image_plane",
    ):
        if getattr(L, name).restype is None or name != "bh_version":
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().bh_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed: {msg}")
