"""Writes tests/golden/reference_unittests.npz: the reference's own unit-test
cases as input/expected-output vectors.

The reference commits no golden files; its tests build inputs from closed
forms at run time and derive expectations from them
(test/unittest_btree.cu:77-87,138-251, test/unittest_versioning.cu:82-95,
145-376): keys[i] = 2(i+1) in insertion order, value = key, absent keys
key + 1, ranges [key, key + 2L) with L = --range-length (default 2), num_keys
default 1024. The versioned tests insert the first half, take snapshot s0,
then insert the second half. Expected outputs below are computed from those
closed forms directly (no oracle, no tree): at a snapshot holding keys
2..2m, the range starting at key k holds the keys k, k+2, ... up to
min(k + 2(L-1), 2m).

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MISS = np.uint32(0xFFFFFFFF)
EMPTY_PAIR = np.uint64(0xFFFFFFFFFFFFFFFF)


def range_expect(lower: np.ndarray, L: int, max_key: int):
    n = lower.size
    counts = np.zeros(n, dtype=np.uint32)
    pairs = np.full((n, L), EMPTY_PAIR, dtype=np.uint64)
    for i, k in enumerate(lower.astype(np.int64)):
        c = 0 if k > max_key else min(L, (max_key - k) // 2 + 1)
        counts[i] = c
        for j in range(c):
            key = k + 2 * j
            pairs[i, j] = np.uint64(key) | (np.uint64(key) << np.uint64(32))
    return pairs.reshape(-1), counts


def case(num_keys: int, L: int):
    keys = (np.arange(num_keys, dtype=np.uint64) + 1) * 2
    keys = keys.astype(np.uint32)
    values = keys.copy()
    absent = keys + 1
    lower = keys.copy()
    upper = lower + np.uint32(2 * L)
    half = num_keys // 2
    d = {
        "keys": keys, "values": values, "absent": absent, "lower": lower, "upper": upper,
        "half": np.array([half], dtype=np.int64), "L": np.array([L], dtype=np.int64),
        # unittest_btree.cu FindExist / FindNotExist / EraseAll / EraseNone
        "find_exist": values.copy(),
        "find_not_exist": np.full(num_keys, MISS, dtype=np.uint32),
        # unittest_versioning.cu at s0 (first half only) and after the second half
        "find_s0_first_half": values[:half].copy(),
        "find_s0_second_half": np.full(num_keys - half, MISS, dtype=np.uint32),
    }
    d["range_s0_pairs"], d["range_s0_counts"] = range_expect(lower[:half], L, int(keys[half - 1]))
    d["range_s1_pairs"], d["range_s1_counts"] = range_expect(lower, L, int(keys[-1]))
    return d


def main():
    out = {}
    for num_keys, L in ((1024, 2), (1024, 32), (10000, 16)):
        for k, v in case(num_keys, L).items():
            out[f"n{num_keys}_L{L}__{k}"] = v
    path = os.path.join(HERE, "reference_unittests.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({len(out)} arrays)")


if __name__ == "__main__":
    main()
