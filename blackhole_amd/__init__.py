"""blackhole_amd - MI355X (gfx950) native per-pixel Kerr null-geodesic tracer.

The compute path is the HIP library ``libbhk.so`` (C ABI in include/*.h);
this package holds its Python host interface.
"""
from . import abi  # noqa: F401
from ._lib import LIB_PATH, NativeLibraryMissing, lib  # noqa: F401

__all__ = ["abi", "lib", "LIB_PATH", "NativeLibraryMissing"]
