"""The C-ABI library loads and exports every entry point include/*.h declares; host-only
helpers agree with the reference's tables. No device work runs here."""

from __future__ import annotations

import ctypes
import json
import pathlib
import re

import numpy as np

from dynamis_amd import _native, abi

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"


def declared_functions() -> set[str]:
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = header.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for match in re.finditer(r"\b(dyn_[a-z0-9_]+)\s*\(", text):
            names.add(match.group(1))
    return names


def test_every_declared_symbol_is_exported():
    lib = _native.lib()
    declared = declared_functions()
    assert declared, "no declarations found"
    missing = [name for name in sorted(declared) if not hasattr(lib, name)]
    assert not missing, f"declared but not exported: {missing}"


def test_every_declared_symbol_has_a_binding():
    assert declared_functions() <= set(_native.SIGNATURES)


def test_version_and_error_text():
    lib = _native.lib()
    assert lib.dyn_version() >= 1
    assert isinstance(lib.dyn_last_error(), bytes)


def test_key_words_table_matches_the_reference():
    cases = json.loads((GOLDEN / "sort_cases.json").read_text())
    lib = _native.lib()
    for elements, words in cases["key_words"]:
        assert lib.dyn_sort_key_words(elements) == words, elements


def test_units_for_matches_the_reference_formula():
    lib = _native.lib()
    # radix.rs:63-67: ceil(live/256) + 64, clamped to [1, 512]
    for live, units in [(0, 64), (1, 65), (256, 65), (257, 66), (1 << 20, 512), (200_000, 512)]:
        assert lib.dyn_sort_units_for(live) == units, live


def test_sort_rejects_a_sort_without_digits():
    lib = _native.lib()
    dummy = ctypes.c_void_p(16)
    rc = lib.dyn_sort_pairs(None, dummy, dummy, dummy, 16, None, None, 1, 0, 0, dummy, 1 << 20)
    assert rc == -1
    assert b"at least one key digit" in lib.dyn_last_error()


def test_sort_rejects_too_small_scratch():
    lib = _native.lib()
    dummy = ctypes.c_void_p(16)
    rc = lib.dyn_sort_pairs(None, dummy, dummy, dummy, 1024, None, None, 1, 4, 0, dummy, 16)
    assert rc == -3


def test_record_layouts():
    assert abi.AABB.itemsize == 32
    assert abi.GRID_ENTRY.itemsize == 32
    assert abi.COUNTER["PAIRS"] == 10 and abi.COUNTER["GRID_SCALE"] == 35
    assert abi.COUNTER_DEVICE_COUNT == 52


def test_device_count_is_never_an_error():
    assert _native.device_count() >= 0


def test_scratch_sizes_grow_with_capacity():
    lib = _native.lib()
    small = lib.dyn_sort_scratch_bytes(1000, 4)
    large = lib.dyn_sort_scratch_bytes(1_000_000, 4)
    assert 0 < small < large
    assert large >= 3 * 4 * 1_000_000
    assert lib.dyn_broadphase_scratch_bytes(1000, 8000, 8000) > 0
    assert np.uint64(lib.dyn_broadphase_scratch_bytes(0, 0, 0)) > 0
