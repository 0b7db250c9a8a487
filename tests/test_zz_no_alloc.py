"""dyn_sort_pairs_static: the radix sort with its scratch in the library's static device array
(no allocation in the call), against dyn_sort_pairs with a caller's scratch on the same lanes,
bit for bit, on two calls in a row (the second smaller, over the first's leftover scratch), and
its refusal of a capacity beyond the static scratch."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from dynamis_amd._native import DynamisError, call, lib
from dynamis_amd.sort import RadixSort, SortChannels, units_for

pytestmark = pytest.mark.gpu


def sort_static(channels: SortChannels, major_words: int, minor_words: int) -> None:
    call("dyn_sort_pairs_static", None, channels.major.handle, channels.minor.handle,
         channels.payload.handle, channels.capacity, channels.count.handle, None,
         units_for(channels.capacity), major_words, minor_words)
    call("dyn_stream_synchronize", None)


def lanes(n: int, seed: int):
    rng = np.random.default_rng(seed)
    major = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    minor = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    major[::7] = major[0]  # runs of equal keys: the order among them is the input order
    return major, minor, np.arange(n, dtype=np.uint32)


@pytest.mark.parametrize("n,major_words,minor_words,slots",
                         [(700_000, 4, 4, 1 << 20), (50_000, 2, 1, 60_000)])
def test_static_scratch_sort_equals_scratch_sort(device, n, major_words, minor_words, slots):
    major, minor, payload = lanes(n, seed=n)
    want = SortChannels.upload(major, minor, payload, slots)
    RadixSort().sort(want, units_for(n), major_words, minor_words)
    got = SortChannels.upload(major, minor, payload, slots)
    sort_static(got, major_words, minor_words)
    for g, w in zip(got.download(n), want.download(n)):
        assert g.tobytes() == w.tobytes()


def test_two_calls_in_a_row(device):
    """The second call runs over what the first left in the shared array."""
    first = lanes(900_000, seed=1)
    second = lanes(123_457, seed=2)
    results = []
    for (major, minor, payload), words in ((first, (4, 4)), (second, (1, 3))):
        want = SortChannels.upload(major, minor, payload)
        RadixSort().sort(want, units_for(len(major)), *words)
        got = SortChannels.upload(major, minor, payload)
        sort_static(got, *words)
        results.append((got.download(len(major)), want.download(len(major))))
    for got, want in results:
        for g, w in zip(got, want):
            assert g.tobytes() == w.tobytes()


def test_capacity_beyond_the_static_scratch_is_refused(device):
    lib().dyn_sort_static_scratch_bytes.restype = ctypes.c_size_t
    assert lib().dyn_sort_static_scratch_bytes() == 17 << 20
    major, minor, payload = lanes(16, seed=3)
    big = SortChannels.upload(major, minor, payload, 1_400_000)
    with pytest.raises(DynamisError, match="beyond the static scratch"):
        sort_static(big, 4, 4)
    with pytest.raises(DynamisError):
        sort_static(SortChannels.upload(major, minor, payload), 0, 0)
