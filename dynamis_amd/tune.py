"""Autotuning of the sort passes' tiling, kept per GPU.

The library chooses each sort pass's tiling (keys per lane, persistent workgroups) from a
process-wide table keyed by device and capacity bucket (ceil(log2(capacity))); `dyn_sort_autotune`
fills an entry by timing every tiling on the device. `autotune_world` tunes the capacities of a
world's sorts and keeps the table in a JSON file per GPU name, so later runs on the same kind of
GPU load it instead of timing again. Every tiling sorts to the same stable order: tuning changes
time, never results.
"""

from __future__ import annotations

import ctypes
import json
import os
import pathlib
import re

import numpy as np

from ._native import call

MAX_CELLS_PER_COLLIDER = 8  # entries per collider: the world's entry-sort capacity factor


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    call("dyn_device_name", buf, len(buf))
    return buf.value.decode()


def table() -> np.ndarray:
    """The process-wide table: rows of (device, bucket, keys per lane, groups)."""
    count = ctypes.c_uint32()
    call("dyn_sort_tuning_export", None, 0, ctypes.byref(count))
    out = np.zeros((max(count.value, 1), 4), np.uint32)
    call("dyn_sort_tuning_export", out.ctypes.data, count.value, ctypes.byref(count))
    return out[:count.value].copy()


def load(entries: np.ndarray) -> None:
    entries = np.ascontiguousarray(entries, np.uint32).reshape(-1, 4)
    call("dyn_sort_tuning_import", entries.ctypes.data if len(entries) else None, len(entries))


def bucket(capacity: int) -> int:
    return 0 if capacity <= 1 else int(capacity - 1).bit_length()


def cache_path(directory: str | os.PathLike | None = None) -> pathlib.Path:
    base = pathlib.Path(directory or os.environ.get("DYNAMIS_TUNE_DIR",
                                                      pathlib.Path.home() / ".cache" / "dynamis_amd"))
    return base / ("sort-tilings-" + re.sub(r"[^A-Za-z0-9_.-]+", "_", device_name()) + ".json")


def autotune_solver(world, directory: str | os.PathLike | None = None, retune: bool = False,
                    steps: int = 3) -> dict:
    """Tunes (or loads from the per-GPU cache, keyed by the contact capacity's bucket) the
    workgroup size of a world's solver substeps; the world's state is left as it was."""
    if hasattr(world, "_ensure"):
        world._ensure()
    path = cache_path(directory).with_name(cache_path(directory).name.replace("sort-tilings-", "solver-threads-"))
    key = str(bucket(world.capacity))
    data = {}
    if path.exists() and not retune:
        data = json.loads(path.read_text())
        if data.get("device") != device_name():
            data = {}
    sizes = data.get("threads", {})
    if key in sizes:
        world.set_solver_threads(int(sizes[key]))
        return {"threads": int(sizes[key]), "from_cache": True, "cache": str(path)}
    threads, ms = world.tune_solver(steps)
    sizes[key] = threads
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps({"device": device_name(), "threads": sizes}))
    return {"threads": threads, "ms_per_step": {"64": ms[0], "128": ms[1], "256": ms[2]}, "from_cache": False,
            "cache": str(path)}


def autotune_world(world, directory: str | os.PathLike | None = None, retune: bool = False) -> dict:
    """Tunes (or loads from the per-GPU cache) the tilings of a World's or Batch's sorts and
    makes its next step re-capture. Returns {capacity: (keys per lane, groups)} for its sorts and
    whether the table came from the cache."""
    if hasattr(world, "_ensure"):
        world._ensure()
    caps = [MAX_CELLS_PER_COLLIDER * max(len(world.scene.colliders), 1), world.pair_capacity, world.capacity]
    path = cache_path(directory)
    cached = None
    if path.exists() and not retune:
        data = json.loads(path.read_text())
        if data.get("device") == device_name():
            cached = np.array(data["entries"], np.uint32).reshape(-1, 4)
    if cached is not None and {bucket(c) for c in caps} <= set(int(b) for b in cached[:, 1]):
        rows = {(int(r[0]), int(r[1])): r for r in table()}
        for r in cached:
            rows[(int(r[0]), int(r[1]))] = r
        load(np.array(list(rows.values()), np.uint32))
        world.set_graphs(True)  # drops the captured step
        chosen = {c: tuple(int(v) for v in cached[cached[:, 1] == bucket(c)][0, 2:4]) for c in caps}
        return {"tilings": chosen, "from_cache": True, "cache": str(path)}
    out = np.zeros(9, np.uint32)
    call("dyn_world_autotune", world._handle, out.ctypes.data)
    chosen = {int(out[3 * i]): (int(out[3 * i + 1]), int(out[3 * i + 2])) for i in range(3)}
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps({"device": device_name(), "entries": table().tolist()}))
    return {"tilings": chosen, "from_cache": False, "cache": str(path)}
