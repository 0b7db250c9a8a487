"""The oracle's grid broadphase against hand-worked hashes, the reference's own pair-count
scenes and the brute-force overlap set. CPU only."""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from dynamis_amd import abi
from tests.scenes import brute_force_pairs, canonical_pairs, uniform_boxes

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "grid_cases.json"
GRID = json.loads(GOLDEN.read_text())


@pytest.mark.parametrize("case", GRID["hashes"], ids=lambda c: str(c["coord"]))
def test_cell_hash_known_answers(oracle, case):
    assert oracle.cell_hash(*case["coord"]) == case["hash"]


@pytest.mark.parametrize("case", GRID["keys"], ids=lambda c: f"{c['level']}{c['coord']}")
def test_cell_key_known_answers(oracle, case):
    assert oracle.cell_key(case["level"], *case["coord"]) == case["key"]


def scene_arrays(scene):
    bodies = scene["bodies"]
    boxes = abi.aabbs(np.array([b["min"] for b in bodies], np.float32),
                      np.array([b["max"] for b in bodies], np.float32))
    res = abi.aabbs(np.array([b["res_min"] for b in bodies], np.float32),
                    np.array([b["res_max"] for b in bodies], np.float32))
    flags = np.array([(abi.BP_MOBILE | abi.BP_AWAKE) if b["mobile"] else 0 for b in bodies],
                     np.uint32)
    return boxes, res, flags


@pytest.mark.parametrize("scene", GRID["scenes"], ids=lambda s: s["name"])
def test_reference_pair_counts(oracle, scene):
    boxes, res, flags = scene_arrays(scene)
    out = oracle.broadphase(boxes, flags=flags, resolution=res)
    assert abi.counter(out["counters"], "PAIRS") == scene["pairs"]
    assert abi.counter(out["counters"], "REFUSED_PAIRS") == 0
    assert abi.counter(out["counters"], "ENTRY_FAULTS") == 0
    assert np.array_equal(canonical_pairs(out["pairs"]), brute_force_pairs(boxes, flags=flags))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_scene_matches_brute_force(oracle, seed):
    boxes = uniform_boxes(1500, 40.0, 0.2, 2.0, seed)
    out = oracle.broadphase(boxes)
    assert abi.counter(out["counters"], "ENTRY_FAULTS") == 0
    assert np.array_equal(out["pairs"], brute_force_pairs(boxes))


def test_mixed_flags_and_groups_match_brute_force(oracle):
    rng = np.random.default_rng(7)
    boxes = uniform_boxes(1200, 30.0, 0.1, 3.0, 7)
    groups = rng.integers(0, 900, size=len(boxes)).astype(np.uint32)
    flags = rng.choice([0, abi.BP_MOBILE, abi.BP_AWAKE, abi.BP_MOBILE | abi.BP_AWAKE],
                       size=len(boxes)).astype(np.uint32)
    out = oracle.broadphase(boxes, groups=groups, flags=flags)
    assert np.array_equal(out["pairs"], brute_force_pairs(boxes, groups, flags))


def test_entries_are_the_cells_of_each_level(oracle):
    boxes = uniform_boxes(500, 50.0, 0.2, 2.0, 11)
    out = oracle.broadphase(boxes)
    keys, order, entries = out["entry_keys"], out["entry_order"], out["entries"]
    assert np.all(np.diff(keys.astype(np.int64)) >= 0), "region keys are sorted"
    info = entries["info"]
    collider = info & abi.ENTRY_INDEX_MASK
    # every collider emits its primary entry (cell offset 0) exactly once
    primary = (info & abi.ENTRY_PRIMARY) != 0
    assert np.array_equal(np.sort(collider[primary]), np.arange(len(boxes)))
    # the slot payload is a permutation of the slots
    assert np.array_equal(np.sort(order), np.arange(len(order)))
    # levels: the finest at which the box spans at most two cells per axis
    scale = abi.counter(out["counters"], "GRID_SCALE")
    extent = abi.counter(out["counters"], "GRID_EXTENT")
    base = oracle.h.oracle_grid_cell_size(scale, extent)
    sizes = (boxes["max"] - boxes["min"]).max(axis=1)
    assert base == pytest.approx(sizes.min(), rel=1e-6)
    for slot in np.nonzero(primary)[0][:50]:
        c = collider[slot]
        level = keys[np.nonzero(order == slot)[0][0]] >> abi.LEVEL_KEY_SHIFT
        cell = np.float32(base) * np.float32(2.0 ** level)
        span = np.floor(boxes["max"][c] / cell) - np.floor(boxes["min"][c] / cell) + 1
        assert span.max() <= 2
        if level > 0:
            finer = np.float32(base) * np.float32(2.0 ** (level - 1))
            fspan = np.floor(boxes["max"][c] / finer) - np.floor(boxes["min"][c] / finer) + 1
            assert fspan.max() > 2


def test_pair_capacity_overflow_is_counted(oracle):
    boxes = uniform_boxes(800, 15.0, 0.5, 2.0, 5)
    full = oracle.broadphase(boxes)
    total = abi.counter(full["counters"], "PAIRS")
    assert total > 20
    cut = oracle.broadphase(boxes, pair_capacity=20)
    assert abi.counter(cut["counters"], "PAIRS") == total
    assert abi.counter(cut["counters"], "REFUSED_PAIRS") == total - 20


def test_a_plausible_mistake_is_caught():
    # Open-interval overlap (a wrong comparison) would drop touching pairs that the closed
    # test of core.wgsl:128-130 keeps.
    boxes = abi.aabbs(np.array([[0, 0, 0], [1, 0, 0]], np.float32),
                      np.array([[1, 1, 1], [2, 1, 1]], np.float32))
    assert len(brute_force_pairs(boxes)) == 1
    mn, mx = boxes["min"], boxes["max"]
    assert not np.all(mn[0] < mx[1]) or not np.all(mn[1] < mx[0])
