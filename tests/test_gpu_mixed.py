"""GPU tests of the reference's mixed-operation calls, concurrent_insert_range
and concurrent_find_erase (one launch per call).

Versioned trees answer every read of a call from the state before the call,
so the results are compared word for word with the oracle. Plain trees read
the latest state while the call's writes land (any linearization, as in the
reference), so their reads are checked against the bounds the reference's
tolerant validators use (validation.hpp:106-149, 266-369): every range holds
at least the keys present before the call and at most those present after
it, and a find of a key the call does not touch is exact.
"""
from __future__ import annotations

import numpy as np
import pytest

import mvgpubtree_amd as mv
import oracle
from mvgpubtree_amd import workloads as wl

from conftest import to_dev, to_np

pytestmark = pytest.mark.gpu

WIDTHS = (16, 32, 64)


def setup(dev, kind, nw, seed, n_initial=60_000, n_new=20_000):
    rng = np.random.default_rng(seed)
    keys = wl.unique_keys(rng, n_initial + n_new)
    init, new = keys[:n_initial], keys[n_initial:]
    cls = mv.GpuVersionedBTree if kind == "versioned" else mv.GpuBlinkTree
    tree = cls(pool_nodes=1 << 17, node_width=nw)
    tree.insert(to_dev(init, dev), to_dev(wl.value_of(init), dev))
    return rng, init, new, tree


def ranges(rng, n, n_live, expected=12):
    lo = rng.integers(1, 0xFFFFFF00, size=n, dtype=np.uint64).astype(np.uint32)
    hi = np.minimum(lo.astype(np.uint64) + wl.span_for(expected, n_live), 0xFFFFFFFF).astype(np.uint32)
    return lo, hi


@pytest.mark.parametrize("nw", WIDTHS)
def test_versioned_insert_range_reads_state_before_call(dev, nw):
    rng, init, new, tree = setup(dev, "versioned", nw, seed=10 + nw)
    ref = oracle.MVMap()
    ref.build_sorted(np.sort(init), wl.value_of(np.sort(init)))
    s0 = tree.take_snapshot()
    assert s0 == ref.take_snapshot()
    lo, hi = ranges(rng, 15_000, init.size)
    L = 64
    pairs, counts = tree.concurrent_insert_range(to_dev(new, dev), to_dev(wl.value_of(new), dev),
                                                 to_dev(lo, dev), to_dev(hi, dev), L)
    rp, rc = ref.range_query(lo, hi, L)  # state before the call
    assert np.array_equal(to_np(counts)[: lo.size], rc)
    assert np.array_equal(to_np(pairs, np.uint64)[: lo.size * L], rp)
    ref.take_snapshot()  # the call reserved one snapshot id
    ref.insert(new, wl.value_of(new))
    assert tree.take_snapshot() == ref.take_snapshot()
    assert np.array_equal(tree.export(), ref.export())
    # the snapshot taken before the call still reads the old state
    q = np.concatenate([init[:5000], new[:5000]])
    assert np.array_equal(to_np(tree.find(to_dev(q, dev), timestamp=s0)), ref.find(q, s0))
    assert tree.error_flags() == 0
    tree.close()


@pytest.mark.parametrize("nw", WIDTHS)
def test_versioned_find_erase_reads_state_before_call(dev, nw):
    rng, init, new, tree = setup(dev, "versioned", nw, seed=20 + nw)
    ref = oracle.MVMap()
    ref.build_sorted(np.sort(init), wl.value_of(np.sort(init)))
    s0 = tree.take_snapshot()
    ref.take_snapshot()
    era = init[rng.choice(init.size, 20_000, replace=False)]
    fk = np.concatenate([era[:8000], init[rng.integers(0, init.size, 8000)], new[:4000]])
    rng.shuffle(fk)
    got = to_np(tree.concurrent_find_erase(to_dev(fk, dev), to_dev(era, dev)))
    assert np.array_equal(got[: fk.size], ref.find(fk))  # every find sees the state before the call
    ref.take_snapshot()
    ref.erase(era)
    assert np.array_equal(tree.export(), ref.export())
    assert np.array_equal(to_np(tree.find(to_dev(era, dev), timestamp=s0)), ref.find(era, s0))
    assert tree.error_flags() == 0
    tree.close()


@pytest.mark.parametrize("nw", WIDTHS)
def test_plain_insert_range_within_reference_bounds(dev, nw):
    rng, init, new, tree = setup(dev, "plain", nw, seed=30 + nw)
    pre = oracle.SortedState(init, wl.value_of(init))
    post = oracle.SortedState(np.concatenate([init, new]), wl.value_of(np.concatenate([init, new])))
    lo, hi = ranges(rng, 10_000, init.size)
    L = 96
    pairs, counts = tree.concurrent_insert_range(to_dev(new, dev), to_dev(wl.value_of(new), dev),
                                                 to_dev(lo, dev), to_dev(hi, dev), L)
    c = to_np(counts)[: lo.size]
    p = to_np(pairs, np.uint64)[: lo.size * L].reshape(lo.size, L)
    _, c_pre = pre.range_query(lo, hi, L)
    p_post, c_post = post.range_query(lo, hi, L)
    assert (c_post <= L).all()
    assert ((c_pre <= c) & (c <= c_post)).all()
    p_post = p_post.reshape(lo.size, L)
    pre_set = set(init.tolist())
    for i in range(0, lo.size, 7):  # sampled row checks: sorted, within post, covering pre
        row = p[i, : c[i]]
        keys = (row & 0xFFFFFFFF).astype(np.uint64)
        assert np.all(np.diff(keys.astype(np.int64)) > 0)
        assert set(row.tolist()) <= set(p_post[i, : c_post[i]].tolist())
        want_pre = {k for k in p_post[i, : c_post[i]] & 0xFFFFFFFF if int(k) in pre_set}
        assert want_pre <= set(keys.tolist())
    assert np.array_equal(tree.export(), post_export(post))
    assert tree.error_flags() == 0
    tree.close()


def post_export(state):
    k, v = state.keys, state.values
    return k.astype(np.uint64) | (v.astype(np.uint64) << np.uint64(32))


@pytest.mark.parametrize("nw", WIDTHS)
def test_plain_find_erase_within_reference_bounds(dev, nw):
    rng, init, new, tree = setup(dev, "plain", nw, seed=40 + nw)
    era = init[rng.choice(init.size, 20_000, replace=False)]
    keep = np.setdiff1d(init, era)
    fk = np.concatenate([era[:6000], keep[:6000], new[:3000]])
    rng.shuffle(fk)
    got = to_np(tree.concurrent_find_erase(to_dev(fk, dev), to_dev(era, dev)))[: fk.size]
    erased = np.isin(fk, era)
    want = np.where(np.isin(fk, init), wl.value_of(fk), np.uint32(mv.NOT_FOUND))
    # untouched keys are exact; keys erased by the call read either their value or a miss
    assert np.array_equal(got[~erased], want[~erased])
    assert np.all((got[erased] == want[erased]) | (got[erased] == mv.NOT_FOUND))
    final = oracle.SortedState(keep, wl.value_of(keep))
    assert np.array_equal(tree.export(), post_export(final))
    assert tree.error_flags() == 0
    tree.close()


@pytest.mark.parametrize("nw", WIDTHS)
def test_versioned_mixed_calls_on_hot_leaves(dev, nw):
    """Small trees, so that nearly every leaf a read scans is written by the
    same call (version copies published mid-scan, leaves split under the
    reads): five insert/range calls and five find/erase calls in a row,
    every result word for word against the state before its call."""
    rng = np.random.default_rng(300 + nw)
    keys = wl.unique_keys(rng, 40_000)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 16, node_width=nw)
    ref = oracle.MVMap()
    live = keys[:3000]
    tree.insert(to_dev(live, dev), to_dev(wl.value_of(live), dev))
    ref.insert(live, wl.value_of(live))
    pos = 3000
    L = 256
    for call in range(5):
        new = keys[pos:pos + 3000]
        pos += 3000
        lo = np.sort(rng.choice(np.concatenate([live, new]), 6000)).astype(np.uint32)
        hi = np.minimum(lo.astype(np.uint64) + wl.span_for(40, live.size), 0xFFFFFFFF).astype(np.uint32)
        ref.take_snapshot()
        pairs, counts = tree.concurrent_insert_range(to_dev(new, dev), to_dev(wl.value_of(new), dev),
                                                     to_dev(lo, dev), to_dev(hi, dev), L)
        rp, rc = ref.range_query(lo, hi, L)
        assert np.array_equal(to_np(counts)[: lo.size], rc), call
        assert np.array_equal(to_np(pairs, np.uint64)[: lo.size * L], rp), call
        ref.insert(new, wl.value_of(new))
        live = np.concatenate([live, new])
    for call in range(5):
        era = live[rng.choice(live.size, 2000, replace=False)]
        fk = np.concatenate([era, live[rng.integers(0, live.size, 4000)], keys[pos:pos + 1000]])
        ref.take_snapshot()
        got = to_np(tree.concurrent_find_erase(to_dev(fk, dev), to_dev(era, dev)))
        assert np.array_equal(got[: fk.size], ref.find(fk)), call
        ref.erase(era)
        live = np.setdiff1d(live, era)
        tree.reclaim()
    assert np.array_equal(tree.export(), ref.export())
    assert tree.error_flags() == 0, hex(tree.error_flags())
    tree.close()


def test_versioned_insert_range_calls_reclaiming_between_them(dev):
    """The shape of the reference's reclaimer benchmark
    (concurrent_insert_range_reclaimer_bench.cu:50-153), scaled down: a
    16-pair tree of 120K keys, then 60K inserts and 60K ranges [k, k + 16)
    with k drawn from the initial keys (rkg.hpp:50-73), L = 16, in 30 mixed
    calls with a reclamation after each. Every call's ranges match the state
    before it; each reclamation frees every copy the call retired (no
    snapshot is held), one epoch per call lands in the timeline, live nodes
    never shrink (inserts only), and nodes ever allocated = live + freed."""
    rng = np.random.default_rng(606)
    keys = wl.unique_keys(rng, 180_000)
    init, new = keys[:120_000], keys[120_000:]
    lo = init[rng.permutation(init.size)[: new.size]]
    hi = np.minimum(lo.astype(np.uint64) + 16, 0xFFFFFFFF).astype(np.uint32)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 17, node_width=16)
    tree.insert(to_dev(init, dev), to_dev(wl.value_of(init), dev))
    ref = oracle.MVMap()
    ref.build_sorted(np.sort(init), wl.value_of(np.sort(init)))
    L, calls = 16, 30
    e0 = len(tree.memory_series())
    live_prev = tree.memory_stats().live_nodes
    cut = np.linspace(0, new.size, calls + 1).astype(int)
    for c in range(calls):
        a, b = cut[c], cut[c + 1]
        ref.take_snapshot()
        pairs, counts = tree.concurrent_insert_range(to_dev(new[a:b], dev), to_dev(wl.value_of(new[a:b]), dev),
                                                     to_dev(lo[a:b], dev), to_dev(hi[a:b], dev), L)
        rp, rc = ref.range_query(lo[a:b], hi[a:b], L)
        assert np.array_equal(to_np(counts)[: b - a], rc), c
        assert np.array_equal(to_np(pairs, np.uint64)[: (b - a) * L], rp), c
        assert rc.min() >= 1  # k itself is live in every range
        ref.insert(new[a:b], wl.value_of(new[a:b]))
        tree.reclaim()
        st = tree.memory_stats()
        assert st.retired_pending == 0, c
        assert st.live_nodes >= live_prev, c
        live_prev = st.live_nodes
    ser = tree.memory_series()[e0:]
    assert len(ser) == calls
    assert np.all(ser.allocated_total == ser.live_nodes + ser.reclaimed_total)
    assert np.all(np.diff(ser.allocated_total.astype(np.int64)) > 0)
    assert int(ser.reclaimed_total[-1]) > new.size // 2  # a version copy per touched leaf, freed
    assert np.array_equal(tree.export(), ref.export())
    assert tree.error_flags() == 0, hex(tree.error_flags())
    tree.close()
