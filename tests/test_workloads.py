"""CPU tests of the seeded workload generators (small sizes)."""
from __future__ import annotations

import numpy as np

import oracle
from mvgpubtree_amd import workloads as wl


def test_value_rule_and_sentinel_exclusion():
    assert wl.SENTINEL_COLLISION ^ wl.VALUE_XOR == 0xFFFFFFFF
    rng = np.random.default_rng(0)
    k = wl.unique_keys(rng, 200000)
    assert k.size == np.unique(k).size == 200000
    assert k.min() >= 1 and k.max() <= 0xFFFFFFFE
    assert not np.any(k == wl.SENTINEL_COLLISION)
    assert not np.any(wl.value_of(k) == 0xFFFFFFFF)


def test_generators_are_deterministic():
    a = wl.config1(seed=3, n=5000)
    b = wl.config1(seed=3, n=5000)
    assert np.array_equal(a.keys, b.keys) and np.array_equal(a.queries, b.queries)


def test_config1_half_hits():
    c = wl.config1(seed=1, n=20000)
    s = oracle.SortedState(c.keys, c.values)
    hits = s.find(c.queries) != 0xFFFFFFFF
    assert hits.sum() == 10000


def test_config2_ranges_hold_about_32_keys_and_inserts_are_new():
    w = wl.config2(seed=2, n_initial=200000, num_batches=2, ops_per_batch=20000)
    s = oracle.SortedState(w.initial_keys, w.initial_values)
    for b in w.batches:
        c = s.range_counts(b.range_lo, b.range_hi)
        assert 28 < c.mean() < 36
        assert c.max() < w.range_capacity
        assert not np.any(s.find(b.insert_keys) != 0xFFFFFFFF)
        s.insert(b.insert_keys, b.insert_values)


def test_config3_erases_live_keys_and_finds_half_hit():
    w = wl.config3(seed=3, n_initial=100000, num_batches=3, ops_per_batch=10000)
    s = oracle.SortedState(w.initial_keys, w.initial_values)
    for b in w.batches:
        assert np.all(s.find(b.erase_keys) != 0xFFFFFFFF)
        hits = (s.find(b.find_keys) != 0xFFFFFFFF).sum()
        assert hits == b.find_keys.size // 2
        s.erase(b.erase_keys)


def test_config4_mix_and_sizes():
    w = wl.config4(seed=4, n_initial=50000, num_batches=64, min_batch=32, max_batch=4096)
    sizes = np.array([b.num_ops for b in w.batches])
    assert sizes.min() >= 32 and sizes.max() <= 4096
    tot = sizes.sum()
    f = sum(b.find_keys.size for b in w.batches) / tot
    i = sum(b.insert_keys.size for b in w.batches) / tot
    assert 0.55 < f < 0.62 and 0.17 < i < 0.21
