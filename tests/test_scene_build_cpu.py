"""The selection step of the device BVH build (scene_build.hip select_nth,
run on the host through lucid_debug_select_nth) against std::nth_element as
the reference's build calls it (include/core/scene.hpp:138-142), through the
oracle.  The device build's BVH order is the exact-tie key of closest-hit
traversal, so the permutation must be identical, not merely a valid
selection: keys with many ties (grid meshes put whole rows of centroids on
one coordinate), sorted, reversed and constant inputs, every nth of small
ranges, and the heap-select fallback forced with small partition budgets.
"""

import ctypes as C

import numpy as np
import pytest

import lucid_capi
import oracle


def _product(keys, idx, nth, depth_limit=-1):
    L = lucid_capi.load()
    k = np.ascontiguousarray(keys, np.float32).copy()
    i = np.ascontiguousarray(idx, np.int32).copy()
    assert L.lucid_debug_select_nth(k.ctypes.data_as(C.POINTER(C.c_float)), i.ctypes.data_as(C.POINTER(C.c_int32)),
                                    len(k), int(nth), int(depth_limit)) == 0
    return k, i


def _check(keys, nth, depth_limit=-1):
    idx = np.arange(len(keys), dtype=np.int32)
    pk, pi = _product(keys, idx, nth, depth_limit)
    ok, oi = oracle.nth_element(keys, idx, nth, depth_limit)
    np.testing.assert_array_equal(pi, oi)
    np.testing.assert_array_equal(pk, ok)
    if 0 <= nth < len(keys):  # and it is a selection
        assert (pk[:nth] <= pk[nth]).all() and (pk[nth + 1:] >= pk[nth]).all()


@pytest.mark.parametrize("seed", range(6))
def test_random_keys_with_ties(seed):
    rng = np.random.default_rng(seed)
    for _ in range(60):
        n = int(rng.integers(0, 3000))
        levels = int(rng.choice([1, 2, 3, 7, 50, 1 << 20]))
        keys = (rng.integers(0, levels, n) * 0.25 - 3.0).astype(np.float32)
        _check(keys, n // 2)


def test_structured_inputs():
    for n in (5, 6, 7, 8, 9, 16, 17, 100, 1023, 4096, 4097, 20000):
        a = np.arange(n, dtype=np.float32)
        for keys in (a, a[::-1].copy(), np.zeros(n, np.float32), np.abs(a - n // 2),
                     np.concatenate([a[: n // 2], a[: n - n // 2]]), (a % 3).astype(np.float32)):
            _check(keys, n // 2)


def test_every_nth_of_small_ranges():
    rng = np.random.default_rng(9)
    for n in range(0, 40):
        keys = rng.integers(0, 4, n).astype(np.float32)
        for nth in range(0, n + 1):
            _check(keys, nth)


def test_heap_select_fallback():
    rng = np.random.default_rng(10)
    for depth in (0, 1, 2, 3):
        for _ in range(40):
            n = int(rng.integers(4, 600))
            keys = rng.integers(0, int(rng.choice([2, 5, 1000])), n).astype(np.float32)
            _check(keys, int(rng.integers(0, n)), depth)
