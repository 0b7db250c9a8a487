"""dyn_sort_pairs on the device against the oracle and the stable order, bit-exact on all three
lanes: the reference's own cases, its wide and reservation cases, and the 2^20-key size."""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from dynamis_amd.device import DeviceBuffer
from dynamis_amd.sort import RadixSort, SortChannels, units_for
from tests.sort_cases import expected_order, xorshift

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "sort_cases.json"
CASES = json.loads(GOLDEN.read_text())


def run(major, minor, payload, major_words, minor_words, slots=None, plan=None):
    channels = SortChannels.upload(major, minor, payload, slots)
    n = len(major)
    RadixSort().sort(channels, plan if plan is not None else units_for(n), major_words,
                     minor_words)
    return channels.download(n)


@pytest.mark.parametrize("case", CASES["cases"], ids=lambda c: c["name"])
def test_reference_cases(device, case):
    m, k, p = run(case["major"], case["minor"], case["payload"], case["major_words"],
                  case["minor_words"])
    assert m.tolist() == case["expected_major"]
    assert k.tolist() == case["expected_minor"]
    assert p.tolist() == case["expected_payload"]


def test_wide_sort_300k_matches_oracle(device, oracle):
    n = 300_000
    major, minor = xorshift(n, 1), xorshift(n, 2)
    payload = np.arange(n, dtype=np.uint32)
    m, k, p = run(major, minor, payload, 4, 4)
    om, ok, op = oracle.sort_pairs(major, minor, payload, 4, 4)
    assert np.array_equal(m, om) and np.array_equal(k, ok) and np.array_equal(p, op)
    assert np.array_equal(p, expected_order(major, minor, 4, 4))


@pytest.mark.parametrize("length", [0, 1, 7, 255, 256, 257, 1000, 5000])
def test_stream_inside_a_wider_reservation(device, length):
    major, minor = xorshift(length, 5), xorshift(length, 6)
    payload = np.arange(length, dtype=np.uint32)
    m, k, p = run(major, minor, payload, 4, 4, slots=65_536)
    assert np.array_equal(p, expected_order(major, minor, 4, 4))


@pytest.mark.parametrize("plan", [1, 2, 7, 64])
@pytest.mark.parametrize("length", [0, 1, 255, 256, 1000, 5000, 60_000])
def test_plan_never_changes_the_order(device, plan, length):
    major, minor = xorshift(length, 7), xorshift(length, 9)
    payload = np.arange(length, dtype=np.uint32)
    _, _, p = run(major, minor, payload, 4, 4, plan=plan)
    assert np.array_equal(p, expected_order(major, minor, 4, 4))


@pytest.mark.parametrize("major_words,minor_words", [(1, 0), (2, 0), (3, 3), (0, 2), (2, 1), (4, 0)])
def test_partial_words(device, oracle, major_words, minor_words):
    n = 70_000
    major, minor = xorshift(n, 11), xorshift(n, 12)
    payload = np.arange(n, dtype=np.uint32)
    m, k, p = run(major, minor, payload, major_words, minor_words)
    om, ok, op = oracle.sort_pairs(major, minor, payload, major_words, minor_words)
    assert np.array_equal(p, op) and np.array_equal(m, om) and np.array_equal(k, ok)


@pytest.mark.parametrize("n,words", [(1 << 20, (4, 0)), (1 << 20, (4, 4)), ((1 << 22) + 77, (4, 0))])
def test_full_size(device, n, words):
    rng = np.random.default_rng(n)
    major = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    minor = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    payload = np.arange(n, dtype=np.uint32)
    m, k, p = run(major, minor, payload, *words)
    order = expected_order(major, minor, *words)
    assert np.array_equal(p, order)
    assert np.array_equal(m, major[order]) and np.array_equal(k, minor[order])


def test_heavy_duplicates_stay_stable(device):
    n = 500_000
    rng = np.random.default_rng(3)
    major = rng.integers(0, 3, n).astype(np.uint32)
    minor = np.zeros(n, np.uint32)
    payload = np.arange(n, dtype=np.uint32)
    _, _, p = run(major, minor, payload, 4, 0)
    assert np.array_equal(p, expected_order(major, minor, 4, 0))


def test_base_offset_and_reuse(device, oracle):
    n, base, count = 100_000, 4321, 60_000
    major, minor = xorshift(n, 21), xorshift(n, 22)
    payload = np.arange(n, dtype=np.uint32)
    channels = SortChannels.upload(major, minor, payload)
    channels.count = DeviceBuffer.from_array(np.array([count], np.uint32))
    channels.base = DeviceBuffer.from_array(np.array([base], np.uint32))
    sorter = RadixSort()
    om, ok, op = oracle.sort_pairs(major, minor, payload, 4, 0, count=count, base=base)
    for _ in range(3):  # a sorted stream sorted again is unchanged (and the scratch is reused)
        sorter.sort(channels, units_for(count), 4, 0)
        m, k, p = channels.download(n)
        assert np.array_equal(m, om) and np.array_equal(k, ok) and np.array_equal(p, op)
