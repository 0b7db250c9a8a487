"""The oracle's radix sort against the reference's own test cases (golden fixture) and against
numpy's stable sort on the masked key bytes. CPU only."""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from tests.sort_cases import expected_order, xorshift

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "sort_cases.json"
CASES = json.loads(GOLDEN.read_text())


@pytest.mark.parametrize("case", CASES["cases"], ids=lambda c: c["name"])
def test_reference_cases(oracle, case):
    major, minor, payload = oracle.sort_pairs(case["major"], case["minor"], case["payload"],
                                              case["major_words"], case["minor_words"])
    assert major.tolist() == case["expected_major"]
    assert minor.tolist() == case["expected_minor"]
    assert payload.tolist() == case["expected_payload"]
    if "literal" in case:
        assert major.tolist() == case["literal"]["major"]
        assert payload.tolist() == case["literal"]["payload"]


def test_key_words_table(oracle):
    for elements, words in CASES["key_words"]:
        assert oracle.key_words(elements) == words


def test_xorshift_generator_matches_fixture():
    assert xorshift(8, 1).tolist() == CASES["xorshift_salt1_first8"]


def test_wide_sort_300k_matches_stable_sort(oracle):
    # crates/dynamis-sort/tests/gpu.rs:227-249
    n = 300_000
    major, minor = xorshift(n, 1), xorshift(n, 2)
    payload = np.arange(n, dtype=np.uint32)
    m, k, p = oracle.sort_pairs(major, minor, payload, 4, 4)
    order = expected_order(major, minor, 4, 4)
    assert np.array_equal(p, order)
    assert np.array_equal(m, major[order]) and np.array_equal(k, minor[order])


@pytest.mark.parametrize("length", [0, 1, 7, 255, 256, 257, 1000, 5000])
def test_stream_inside_a_wider_reservation(oracle, length):
    # gpu.rs:401-436: only `count` records of a 65536-slot reservation are sorted
    slots = 65_536
    major = np.zeros(slots, np.uint32)
    minor = np.zeros(slots, np.uint32)
    payload = np.full(slots, 0xDEADBEEF, np.uint32)
    major[:length] = xorshift(length, 5)
    minor[:length] = xorshift(length, 6)
    payload[:length] = np.arange(length)
    m, k, p = oracle.sort_pairs(major, minor, payload, 4, 4, count=length)
    order = expected_order(major[:length], minor[:length], 4, 4)
    assert np.array_equal(p[:length], order)
    assert np.all(p[length:] == 0xDEADBEEF)


@pytest.mark.parametrize("major_words,minor_words", [(1, 0), (2, 0), (3, 3), (4, 4), (0, 2), (2, 1)])
def test_partial_words_order_only_their_bytes(oracle, major_words, minor_words):
    n = 20_000
    major, minor = xorshift(n, 11), xorshift(n, 12)
    payload = np.arange(n, dtype=np.uint32)
    _, _, p = oracle.sort_pairs(major, minor, payload, major_words, minor_words)
    assert np.array_equal(p, expected_order(major, minor, major_words, minor_words))


def test_base_offset_sorts_only_its_slice(oracle):
    n, base, count = 10_000, 1234, 5000
    major = xorshift(n, 21)
    minor = xorshift(n, 22)
    payload = np.arange(n, dtype=np.uint32)
    m, k, p = oracle.sort_pairs(major, minor, payload, 4, 0, count=count, base=base)
    assert np.array_equal(m[:base], major[:base])
    assert np.array_equal(m[base + count:], major[base + count:])
    order = expected_order(major[base:base + count], minor[base:base + count], 4, 0)
    assert np.array_equal(p[base:base + count], base + order)


def test_no_digit_is_refused(oracle):
    with pytest.raises(ValueError):
        oracle.sort_pairs([1, 2], [1, 2], [0, 1], 0, 0)


def test_a_plausible_mistake_is_caught():
    # Sorting the minor word above the major word (a swapped digit order) must disagree with
    # the stable (major, minor) order the reference defines.
    n = 5000
    major, minor = xorshift(n, 31) & 0xF, xorshift(n, 32)
    good = expected_order(major, minor, 4, 4)
    swapped = np.lexsort((major, minor))
    assert not np.array_equal(good, swapped)
