"""The host restatement of the fixed-cell keys against the oracle's cell hash (no device)."""

from __future__ import annotations

import numpy as np

from .fixed_grid_keys import cell_hash


def test_cell_hash_matches_the_oracle(oracle):
    rng = np.random.default_rng(0)
    pts = rng.integers(-2**20, 2**20, size=(500, 3))
    pts[:10] = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, -1, -1], [2**31 - 1, 0, -2**31],
                [123, 456, 789], [-5, 7, -9], [250, 250, 250], [0, -1, 0]]
    got = cell_hash(pts[:, 0], pts[:, 1], pts[:, 2])
    want = np.array([oracle.cell_hash(int(x), int(y), int(z)) for x, y, z in pts], np.uint32)
    assert np.array_equal(got, want)
