"""CPU tests pinning the oracle.

* against the reference unit tests' cases (tests/golden/reference_unittests.npz,
  expectations from closed forms, test/unittest_btree.cu:138-251,
  test/unittest_versioning.cu:145-376);
* against brute-force event replay on tiny key domains with interleaved
  snapshots, updates and erases (every read recomputed by replaying all writes
  stamped <= snapshot into a dict);
* SortedState (full-size batch oracle) against MVMap on batch workloads.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from mvgpubtree_amd import workloads as wl

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_unittests.npz")
MISS = 0xFFFFFFFF


def golden_cases():
    g = np.load(GOLDEN)
    names = sorted({k.split("__")[0] for k in g.files})
    return [(n, {k.split("__")[1]: g[k] for k in g.files if k.startswith(n + "__")}) for n in names]


@pytest.mark.parametrize("name,c", golden_cases(), ids=[n for n, _ in golden_cases()])
def test_oracle_reference_versioning_cases(name, c):
    half, L = int(c["half"][0]), int(c["L"][0])
    m = oracle.MVMap()
    m.insert(c["keys"][:half], c["values"][:half])
    s0 = m.take_snapshot()
    assert s0 == 0
    m.insert(c["keys"][half:], c["values"][half:])
    s1 = m.take_snapshot()
    assert s1 == 1
    assert np.array_equal(m.find(c["keys"][:half], s0), c["find_s0_first_half"])
    assert np.array_equal(m.find(c["keys"][half:], s0), c["find_s0_second_half"])
    assert np.array_equal(m.find(c["keys"]), c["find_exist"])
    assert np.array_equal(m.find(c["absent"]), c["find_not_exist"])
    p, n = m.range_query(c["lower"][:half], c["upper"][:half], L, s0)
    assert np.array_equal(n, c["range_s0_counts"]) and np.array_equal(p, c["range_s0_pairs"])
    p, n = m.range_query(c["lower"], c["upper"], L, s1)
    assert np.array_equal(n, c["range_s1_counts"]) and np.array_equal(p, c["range_s1_pairs"])


@pytest.mark.parametrize("name,c", golden_cases(), ids=[n for n, _ in golden_cases()])
def test_oracle_reference_btree_cases(name, c):
    m = oracle.MVMap()
    m.insert(c["keys"], c["values"])
    assert np.array_equal(m.find(c["keys"]), c["find_exist"])
    assert np.array_equal(m.find(c["absent"]), c["find_not_exist"])
    assert m.erase(c["absent"]) == 0  # EraseNone
    assert np.array_equal(m.find(c["keys"]), c["find_exist"])
    assert m.erase(c["keys"]) == c["keys"].size  # EraseAll
    assert np.array_equal(m.find(c["keys"]), c["find_not_exist"])
    m.insert(c["keys"], c["values"])  # EraseAllInsertAll
    assert np.array_equal(m.find(c["keys"]), c["find_exist"])
    assert m.size() == c["keys"].size


def replay(events, ts):
    """Brute force: apply every write stamped <= ts, in order."""
    state = {}
    for stamp, op, k, v in events:
        if ts != oracle.LATEST and stamp > ts:
            continue
        if op == "ins":
            state[k] = v
        else:
            state.pop(k, None)
    return state


@pytest.mark.parametrize("seed", range(12))
def test_oracle_brute_force_snapshots(seed):
    rng = np.random.default_rng(seed)
    domain = int(rng.integers(4, 60))
    m = oracle.MVMap()
    events, snaps, clock = [], [], 0
    for _ in range(int(rng.integers(20, 200))):
        r = rng.random()
        if r < 0.15:
            snaps.append(m.take_snapshot())
            assert snaps[-1] == clock
            clock += 1
        elif r < 0.7:
            k = int(rng.integers(0, domain + 2))  # includes reserved key 0
            v = int(rng.integers(0, 2**32 - 1))
            m.insert([k], [v])
            if 0 < k < 0xFFFFFFFF:
                events.append((clock, "ins", k, v))
        else:
            k = int(rng.integers(1, domain + 1))
            m.erase([k])
            events.append((clock, "del", k, 0))
    q = np.arange(0, domain + 3, dtype=np.uint32)
    for ts in snaps + [oracle.LATEST]:
        st = replay(events, ts)
        want = np.array([st.get(int(k), MISS) if k != 0 else MISS for k in q], dtype=np.uint32)
        assert np.array_equal(m.find(q, ts), want), f"find at {ts}"
        lo = rng.integers(0, domain + 2, size=30).astype(np.uint32)
        hi = (lo + rng.integers(0, domain, size=30)).astype(np.uint32)
        L = 5
        pairs, counts = m.range_query(lo, hi, L, ts)
        for i in range(lo.size):
            ks = sorted(k for k in st if lo[i] <= k < hi[i])
            assert counts[i] == len(ks)
            for j in range(L):
                got = int(pairs[i * L + j])
                if j < len(ks):
                    assert got == ks[j] | (st[ks[j]] << 32)
                else:
                    assert got == 0xFFFFFFFFFFFFFFFF
        assert m.size(ts) == len(st)
        exp = m.export(ts)
        assert [int(x) & 0xFFFFFFFF for x in exp] == sorted(st)


def test_sorted_state_matches_mvmap_config2_shape():
    w = wl.config2(seed=5, n_initial=20000, num_batches=3, ops_per_batch=4000, range_capacity=64)
    m = oracle.MVMap()
    m.insert(w.initial_keys, w.initial_values)
    s = oracle.SortedState(w.initial_keys, w.initial_values)
    for b in w.batches:
        ts = m.take_snapshot()
        mp, mc = m.range_query(b.range_lo, b.range_hi, w.range_capacity, ts)
        sp, sc = s.range_query(b.range_lo, b.range_hi, w.range_capacity)
        assert np.array_equal(mc, sc) and np.array_equal(mp, sp)
        m.insert(b.insert_keys, b.insert_values)
        s.insert(b.insert_keys, b.insert_values)
    assert np.array_equal(m.export(), s.keys.astype(np.uint64) | (s.values.astype(np.uint64) << np.uint64(32)))


def test_sorted_state_matches_mvmap_config3_shape():
    w = wl.config3(seed=6, n_initial=20000, num_batches=3, ops_per_batch=4000)
    m = oracle.MVMap()
    m.insert(w.initial_keys, w.initial_values)
    s = oracle.SortedState(w.initial_keys, w.initial_values)
    for b in w.batches:
        ts = m.take_snapshot()
        assert np.array_equal(m.find(b.find_keys, ts), s.find(b.find_keys))
        assert m.erase(b.erase_keys) == s.erase(b.erase_keys) == b.erase_keys.size
    assert m.size() == len(s)


def test_oracle_build_sorted_equals_insert():
    rng = np.random.default_rng(9)
    k = wl.unique_keys(rng, 5000)
    a, b = oracle.MVMap(), oracle.MVMap()
    a.insert(k, wl.value_of(k))
    ks = np.sort(k)
    b.build_sorted(ks, wl.value_of(ks))
    assert np.array_equal(a.export(), b.export())


def test_planted_mistake_is_caught():
    """A sorted-state oracle with an off-by-one range bound must fail the
    brute-force comparison (guards against checks that cannot fail)."""
    rng = np.random.default_rng(1)
    k = wl.unique_keys(rng, 3000)
    s = oracle.SortedState(k, wl.value_of(k))
    lo = np.sort(k)[::7]
    hi = lo + np.uint32(1)  # [k, k+1) holds exactly k
    counts = s.range_counts(lo, hi)
    assert np.all(counts == 1)
    a, c = s.range_index(lo, hi)
    planted = np.searchsorted(s.keys, hi, side="right") - a  # inclusive bound: wrong
    assert np.array_equal(planted, c)  # no key equals k+1 here ...
    hi2 = np.sort(k)[1::7][: lo.size]  # ... but with hi on a key the planted bound differs
    lo2 = np.sort(k)[0::7][: hi2.size]
    a2, c2 = s.range_index(lo2, hi2)
    planted2 = np.searchsorted(s.keys, hi2, side="right") - a2
    assert not np.array_equal(planted2, c2)
