"""Synthetic scenes for the benchmark and tests (no datasets exist for this path)."""

from __future__ import annotations

import numpy as np

from . import abi


def uniform_boxes(k: int, extent: float, half_lo: float, half_hi: float,
                  seed: int) -> np.ndarray:
    """K boxes with centres uniform in [0, extent)^3 and per-axis half-extents log-uniform in
    [half_lo, half_hi]. At k=2^20, extent=1000, 0.2..2 m this is the isolated sort + broadphase
    workload."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, extent, size=(k, 3)).astype(np.float32)
    halves = np.exp(rng.uniform(np.log(half_lo), np.log(half_hi), size=(k, 3))).astype(np.float32)
    return abi.aabbs(centres - halves, centres + halves)
