"""MI355X-native rigid-body step library: HIP kernels for gfx950 behind the C ABI of
include/dynamis.h, with host mirrors of the reference's interfaces for the same path."""

from ._native import DynamisError, device_count, lib, sort_faults  # noqa: F401

__all__ = ["DynamisError", "device_count", "lib", "sort_faults"]
