"""The sort-tiling table of the autotuner on the host: round trip and validation (no device)."""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd import tune
from dynamis_amd._native import DynamisError


def test_bucket():
    assert [tune.bucket(c) for c in (0, 1, 2, 3, 4, 5, 1 << 20, (1 << 20) + 1)] == [0, 0, 1, 2, 2, 3, 20, 21]


def test_table_round_trip_and_refusals():
    entries = np.array([[0, 19, 8, 512], [0, 21, 16, 2048], [1, 10, 8, 1024]], np.uint32)
    try:
        tune.load(entries)
        got = tune.table()
        assert sorted(map(tuple, got.tolist())) == sorted(map(tuple, entries.tolist()))
        with pytest.raises(DynamisError, match="not a tiling"):
            tune.load(np.array([[0, 19, 12, 512]], np.uint32))
        with pytest.raises(DynamisError, match="not a tiling"):
            tune.load(np.array([[0, 19, 8, 4096]], np.uint32))
    finally:
        tune.load(np.zeros((0, 4), np.uint32))
    assert len(tune.table()) == 0
