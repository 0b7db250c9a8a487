"""ctypes binding of libdynamis_amd.so, the HIP library behind the C ABI of include/dynamis.h.

The library is built in-tree (dynamis_amd/build.py). Loading fails loudly when it is missing:
there is no CPU fallback anywhere in the product.
"""

from __future__ import annotations

import ctypes
import pathlib
import threading

LIBRARY = pathlib.Path(__file__).resolve().parent / "libdynamis_amd.so"

c_u32 = ctypes.c_uint32
c_u32_p = ctypes.POINTER(ctypes.c_uint32)
c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t
c_int = ctypes.c_int

# name -> (restype, argtypes); every symbol include/dynamis.h declares.
SIGNATURES: dict[str, tuple[object, list[object]]] = {
    "dyn_version": (c_int, []),
    "dyn_last_error": (ctypes.c_char_p, []),
    "dyn_device_count": (c_int, [ctypes.POINTER(c_int)]),
    "dyn_set_device": (c_int, [c_int]),
    "dyn_malloc": (c_int, [ctypes.POINTER(c_void_p), c_size_t]),
    "dyn_free": (c_int, [c_void_p]),
    "dyn_memcpy_htod": (c_int, [c_void_p, c_void_p, c_size_t]),
    "dyn_memcpy_dtoh": (c_int, [c_void_p, c_void_p, c_size_t]),
    "dyn_memcpy_dtod": (c_int, [c_void_p, c_void_p, c_size_t]),
    "dyn_memset": (c_int, [c_void_p, c_int, c_size_t]),
    "dyn_stream_create": (c_int, [ctypes.POINTER(c_void_p)]),
    "dyn_stream_destroy": (c_int, [c_void_p]),
    "dyn_stream_synchronize": (c_int, [c_void_p]),
    "dyn_event_create": (c_int, [ctypes.POINTER(c_void_p)]),
    "dyn_event_destroy": (c_int, [c_void_p]),
    "dyn_event_record": (c_int, [c_void_p, c_void_p]),
    "dyn_event_elapsed_ms": (c_int, [ctypes.POINTER(ctypes.c_float), c_void_p, c_void_p]),
    "dyn_profile_enable": (c_int, [c_int]),
    "dyn_profile_watch": (c_int, [ctypes.c_char_p]),
    "dyn_profile_collect": (c_int, [ctypes.POINTER(c_int)]),
    "dyn_profile_entry": (c_int, [c_int, ctypes.c_char_p, c_size_t, ctypes.POINTER(c_u32),
                                  ctypes.POINTER(ctypes.c_double)]),
    "dyn_sort_units_for": (c_u32, [c_u32]),
    "dyn_sort_key_words": (c_u32, [c_u32]),
    "dyn_sort_scratch_bytes": (c_size_t, [c_u32, c_u32]),
    "dyn_sort_pairs": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_u32, c_void_p,
                               c_void_p, c_u32, c_u32, c_u32, c_void_p, c_size_t]),
    "dyn_sort_static_scratch_bytes": (c_size_t, []),
    "dyn_sort_pairs_static": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_u32, c_void_p,
                                      c_void_p, c_u32, c_u32, c_u32]),
    "dyn_sort_faults": (c_int, [ctypes.POINTER(c_u32), c_int]),
    "dyn_broadphase_scratch_bytes": (c_size_t, [c_u32, c_u32, c_u32]),
    "dyn_broadphase": (c_int, [c_void_p, c_void_p]),
    "dyn_narrowphase_scratch_bytes": (c_size_t, [c_u32]),
    "dyn_narrowphase": (c_int, [c_void_p, c_void_p]),
    "dyn_set_epa_small_vertices": (c_int, [c_u32]),
    "dyn_islands_scratch_bytes": (c_size_t, [c_u32]),
    "dyn_islands": (c_int, [c_void_p, c_void_p]),
    "dyn_solve_scratch_bytes": (c_size_t, [c_u32, c_u32]),
    "dyn_solve": (c_int, [c_void_p, c_void_p, ctypes.c_float]),
    "dyn_config_default": (None, [c_void_p]),
    "dyn_step_params_build": (c_int, [c_void_p, c_u32, c_u32, c_u32, c_void_p]),
    "dyn_inertia_inverse": (c_int, [c_void_p, c_void_p]),
    "dyn_hull_support_cells": (c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_u32, c_void_p]),
    "dyn_hull_build": (c_int, [c_void_p, c_u32, c_void_p, ctypes.POINTER(c_u32), c_void_p,
                               ctypes.POINTER(c_u32), c_u32]),
    "dyn_hull_solid": (c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_void_p]),
    "dyn_world_upload_shapes": (c_int, [c_void_p, c_void_p, c_u32, c_void_p, c_u32]),
    "dyn_encode_body": (c_int, [c_void_p, c_void_p, c_u32, c_u32, c_void_p, c_void_p, c_void_p]),
    "dyn_encode_bodies": (c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_void_p]),
    "dyn_world_create": (c_int, [ctypes.POINTER(c_void_p), c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_u32, c_u32, c_void_p, c_void_p, c_u32, c_u32, c_u32]),
    "dyn_world_destroy": (c_int, [c_void_p]),
    "dyn_batch_create": (c_int, [ctypes.POINTER(c_void_p), c_void_p, c_void_p, c_void_p, c_void_p, c_u32,
                                 c_u32, c_void_p, c_void_p, c_void_p, c_u32, c_u32, c_u32, c_u32]),
    "dyn_world_step": (c_int, [c_void_p, ctypes.c_float]),
    "dyn_world_edit": (c_int, [c_void_p, c_void_p, c_void_p, c_u32]),
    "dyn_world_set_graphs": (c_int, [c_void_p, c_int]),
    "dyn_world_read_work": (c_int, [c_void_p, ctypes.POINTER(c_u32)]),
    "dyn_fixed_grid_scratch_bytes": (c_size_t, [c_u32, c_u32, c_u32]),
    "dyn_fixed_grid": (c_int, [c_void_p, c_void_p, c_u32, ctypes.c_float, c_void_p, c_void_p, c_void_p, c_u32,
                               c_void_p, c_void_p, c_u32, c_void_p, c_void_p, c_size_t]),
    "dyn_sort_autotune": (c_int, [c_void_p, c_u32, c_u32, c_u32_p, c_u32_p]),
    "dyn_sort_tuning_export": (c_int, [c_void_p, c_u32, c_u32_p]),
    "dyn_sort_tuning_import": (c_int, [c_void_p, c_u32]),
    "dyn_world_autotune": (c_int, [c_void_p, c_void_p]),
    "dyn_device_name": (c_int, [ctypes.c_char_p, c_u32]),
    "dyn_world_snapshot_bytes": (c_int, [c_void_p, ctypes.POINTER(c_size_t)]),
    "dyn_world_snapshot": (c_int, [c_void_p, c_void_p]),
    "dyn_world_restore": (c_int, [c_void_p, c_void_p]),
    "dyn_adjoint_create": (c_int, [ctypes.POINTER(c_void_p), c_void_p, c_u32, c_u32, c_void_p, c_u32]),
    "dyn_adjoint_forward": (c_int, [c_void_p, ctypes.c_float]),
    "dyn_adjoint_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dyn_adjoint_read_tape": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    "dyn_adjoint_read_bodies": (c_int, [c_void_p, c_void_p]),
    "dyn_adjoint_destroy": (c_int, [c_void_p]),
    "dyn_world_wait": (c_int, [c_void_p]),
    "dyn_world_capacity": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "dyn_world_set_solver_threads": (c_int, [c_void_p, c_u32]),
    "dyn_world_tune_solver": (c_int, [c_void_p, c_u32, c_void_p, c_void_p]),
    "dyn_world_reserve": (c_int, [c_void_p, c_u32, c_u32]),
    "dyn_world_stream": (c_void_p, [c_void_p]),
    "dyn_world_read_states": (c_int, [c_void_p, c_void_p]),
    "dyn_world_read_counters": (c_int, [c_void_p, c_void_p]),
    "dyn_world_read_contacts": (c_int, [c_void_p, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    "dyn_world_read_resting": (c_int, [c_void_p, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    "dyn_world_read_pairs": (c_int, [c_void_p, c_void_p, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
}


class DynamisError(RuntimeError):
    """A C-ABI entry point returned a negative status."""

    def __init__(self, name: str, code: int, message: str):
        super().__init__(f"{name} failed ({code}): {message}")
        self.code = code


_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIBRARY.exists():
                raise RuntimeError(
                    f"{LIBRARY} is missing: build the HIP library first "
                    "(python -c 'import __graft_entry__ as g; g.build()')")
            handle = ctypes.CDLL(str(LIBRARY))
            for name, (restype, argtypes) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = handle
        return _lib


def check(name: str, code: int) -> None:
    if code != 0:
        message = lib().dyn_last_error().decode(errors="replace")
        raise DynamisError(name, code, message)


def call(name: str, *args) -> None:
    """Calls an int-status entry point and raises DynamisError on failure."""
    check(name, getattr(lib(), name)(*args))


def device_count() -> int:
    n = c_int(0)
    call("dyn_device_count", ctypes.byref(n))
    return n.value


def sort_faults(clear: bool = False) -> int:
    """Look-back waits of any sort on the current device that gave up since the last clear
    (dyn_sort_faults); 0 unless some sort's order is wrong."""
    out = c_u32(0)
    call("dyn_sort_faults", ctypes.byref(out), 1 if clear else 0)
    return int(out.value)
