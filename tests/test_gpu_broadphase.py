"""dyn_broadphase on the device against the oracle (bit-exact keys, entries and pair sets) and
the brute-force overlap set, from the reference's scenes up to 2^20 boxes."""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from dynamis_amd import abi
from dynamis_amd.broadphase import Broadphase
from tests.scenes import brute_force_pairs, canonical_pairs, entry_tuples, uniform_boxes
from tests.test_oracle_grid import scene_arrays

pytestmark = pytest.mark.gpu

GRID = json.loads((pathlib.Path(__file__).resolve().parent / "golden" / "grid_cases.json").read_text())

SAME_COUNTERS = ["ENTRIES", "PAIRS", "REFUSED_PAIRS", "ENTRY_FAULTS", "GRID_LEVELS",
                 "COARSE_ACTIVE", "GRID_SCALE", "GRID_EXTENT"]


def gpu_broadphase(boxes, groups=None, flags=None, resolution=None, entry_capacity=None,
                   pair_capacity=None):
    bp = Broadphase(len(boxes), entry_capacity, pair_capacity)
    bp.upload(boxes, groups, flags, resolution)
    bp.run(sort_pairs=True)
    return bp.read()


def assert_matches_oracle(gpu, ref):
    for name in SAME_COUNTERS:
        assert abi.counter(gpu.counters, name) == abi.counter(ref["counters"], name), name
    assert np.array_equal(gpu.entry_keys, ref["entry_keys"])
    # Both lay entries out in collider order, so even the within-key order agrees.
    assert np.array_equal(gpu.entry_order, ref["entry_order"])
    assert np.array_equal(gpu.entries.view(np.uint32), ref["entries"].view(np.uint32))
    assert np.array_equal(gpu.pairs, ref["pairs"])


@pytest.mark.parametrize("scene", GRID["scenes"], ids=lambda s: s["name"])
def test_reference_scenes(device, oracle, scene):
    boxes, res, flags = scene_arrays(scene)
    gpu = gpu_broadphase(boxes, flags=flags, resolution=res)
    assert abi.counter(gpu.counters, "PAIRS") == scene["pairs"]
    assert_matches_oracle(gpu, oracle.broadphase(boxes, flags=flags, resolution=res))


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_scenes_match_oracle_and_brute_force(device, oracle, seed):
    boxes = uniform_boxes(3000, 50.0, 0.2, 2.0, seed)
    gpu = gpu_broadphase(boxes)
    assert_matches_oracle(gpu, oracle.broadphase(boxes))
    assert np.array_equal(gpu.pairs, brute_force_pairs(boxes))


def test_mixed_flags_groups_and_sizes(device, oracle):
    rng = np.random.default_rng(17)
    boxes = uniform_boxes(2500, 30.0, 0.05, 4.0, 17)
    groups = rng.integers(0, 2000, size=len(boxes)).astype(np.uint32)
    flags = rng.choice([0, abi.BP_MOBILE, abi.BP_AWAKE, abi.BP_MOBILE | abi.BP_AWAKE],
                       size=len(boxes)).astype(np.uint32)
    gpu = gpu_broadphase(boxes, groups, flags)
    assert_matches_oracle(gpu, oracle.broadphase(boxes, groups, flags))
    assert np.array_equal(gpu.pairs, brute_force_pairs(boxes, groups, flags))


def test_sparse_coarse_levels(device, oracle):
    """A few large boxes over many small ones: the coarse levels hold a handful of entries each
    and are read whole by the links pass."""
    rng = np.random.default_rng(23)
    small = uniform_boxes(3000, 40.0, 0.1, 0.6, 23)
    centres = rng.uniform(-5.0, 45.0, size=(14, 3)).astype(np.float32)
    halves = rng.uniform(2.0, 30.0, size=(14, 3)).astype(np.float32)
    halves[0] = (1.0e4, 0.5, 1.0e4)  # a floor slab far larger than the rest
    big = abi.aabbs(centres - halves, centres + halves)
    boxes = np.concatenate([big[:7], small, big[7:]])
    groups = rng.integers(0, 2500, size=len(boxes)).astype(np.uint32)
    flags = rng.choice([abi.BP_MOBILE, abi.BP_AWAKE, abi.BP_MOBILE | abi.BP_AWAKE],
                       size=len(boxes)).astype(np.uint32)
    gpu = gpu_broadphase(boxes, groups, flags)
    assert_matches_oracle(gpu, oracle.broadphase(boxes, groups, flags))
    assert np.array_equal(gpu.pairs, brute_force_pairs(boxes, groups, flags))


def test_empty_and_single(device, oracle):
    for k in (0, 1):
        boxes = uniform_boxes(k, 10.0, 0.2, 2.0, 1)
        gpu = gpu_broadphase(boxes)
        assert abi.counter(gpu.counters, "PAIRS") == 0
        assert abi.counter(gpu.counters, "ENTRIES") == abi.counter(
            oracle.broadphase(boxes)["counters"], "ENTRIES")


def test_capacity_overflows_are_counted(device, oracle):
    boxes = uniform_boxes(2000, 15.0, 0.5, 2.0, 5)
    ref = oracle.broadphase(boxes, entry_capacity=3000, pair_capacity=50)
    gpu = gpu_broadphase(boxes, entry_capacity=3000, pair_capacity=50)
    for name in SAME_COUNTERS:
        assert abi.counter(gpu.counters, name) == abi.counter(ref["counters"], name), name
    assert abi.counter(gpu.counters, "REFUSED_PAIRS") > 0
    assert abi.counter(gpu.counters, "ENTRY_FAULTS") > 0
    assert np.array_equal(gpu.entry_keys, ref["entry_keys"])


def test_repeated_runs_are_bit_identical(device):
    boxes = uniform_boxes(20_000, 100.0, 0.2, 2.0, 9)
    bp = Broadphase(len(boxes))
    bp.upload(boxes)
    bp.run()
    first = bp.read()
    for _ in range(3):
        bp.run()
        again = bp.read()
        assert np.array_equal(first.entry_keys, again.entry_keys)
        assert np.array_equal(first.entry_order, again.entry_order)
        assert np.array_equal(first.pairs, again.pairs)


def test_full_size_isolated_workload(device, oracle):
    # 2^20 boxes, centres uniform in a 1000 m cube, half-extents log-uniform in [0.2, 2] m.
    boxes = uniform_boxes(1 << 20, 1000.0, 0.2, 2.0, 2)
    gpu = gpu_broadphase(boxes)
    ref = oracle.broadphase(boxes)
    assert_matches_oracle(gpu, ref)
    assert np.array_equal(entry_tuples(gpu.entry_keys, gpu.entry_order, gpu.entries),
                          entry_tuples(ref["entry_keys"], ref["entry_order"], ref["entries"]))
    assert np.array_equal(canonical_pairs(gpu.pairs), gpu.pairs)
