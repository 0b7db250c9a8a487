"""The literal fixed-cell broadphase (dyn_fixed_grid, 4 m cells, keys (cell hash, box index)):
sorted entries against the host restatement, pairs against the brute-force overlap set and
against the hierarchical grid's pair set on the isolated workload's boxes."""

from __future__ import annotations

import numpy as np
import pytest

from dynamis_amd import abi
from dynamis_amd.broadphase import Broadphase, FixedGrid
from dynamis_amd.scenes import uniform_boxes

from .fixed_grid_keys import sorted_entries
from .scenes import brute_force_pairs, canonical_pairs

pytestmark = pytest.mark.gpu


def fixed(boxes, cell=4.0, **kw):
    g = FixedGrid(len(boxes), cell, **kw)
    g.upload(boxes)
    g.run()
    return g.read()


@pytest.mark.parametrize("seed", [1, 2])
def test_pairs_match_brute_force(device, seed):
    boxes = uniform_boxes(4000, 60.0, 0.2, 2.0, seed)
    r = fixed(boxes)
    want = brute_force_pairs(boxes)
    assert r.refused_boxes == 0 and r.refused_pairs == 0
    assert np.array_equal(r.pairs, want)


def test_sorted_entries_match_host_keys(device):
    boxes = uniform_boxes(20000, 200.0, 0.2, 2.0, 3)
    r = fixed(boxes)
    h, b, o = sorted_entries(boxes, 4.0)
    assert np.array_equal(r.entry_hash, h)
    assert np.array_equal(r.entry_box, b)
    assert np.array_equal(r.entry_ordinal, o)


@pytest.mark.parametrize("k", [1 << 16, 1 << 20])
def test_pairs_equal_hierarchical_grid(device, k):
    """The isolated sort + broadphase workload's boxes (centres U[0,1000)^3, half-extents
    log-U[0.2, 2]): the fixed 4 m grid and the reference's hierarchical grid find the same set."""
    boxes = uniform_boxes(k, 1000.0, 0.2, 2.0, 2)
    r = fixed(boxes)
    bp = Broadphase(k)
    bp.upload(boxes)
    bp.run()
    h = bp.read()
    assert r.refused_boxes == 0 and r.refused_pairs == 0
    assert len(r.pairs) == abi.counter(h.counters, "PAIRS")
    assert np.array_equal(r.pairs, canonical_pairs(h.pairs))
    # every sorted key run is ordered by box index, the stable sort's tie order
    same = r.entry_hash[1:] == r.entry_hash[:-1]
    assert np.all(r.entry_box[1:][same] >= r.entry_box[:-1][same])


def test_refusals_are_counted(device):
    boxes = uniform_boxes(3000, 30.0, 0.2, 2.0, 4)
    big = boxes.copy()
    big["min"][0] = (0.0, 0.0, 0.0)
    big["max"][0] = (300.0, 4.0, 4.0)  # 76 cells along x: refused
    r = fixed(big)
    assert r.refused_boxes == 1
    want = brute_force_pairs(big)
    keep = (want[:, 0] != 0) & (want[:, 1] != 0)
    assert np.array_equal(r.pairs, want[keep])
    few = fixed(boxes, pair_capacity=10)
    assert len(few.pairs) == 10 and few.refused_pairs == len(brute_force_pairs(boxes)) - 10
