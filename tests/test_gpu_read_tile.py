"""GPU parity of read tiles narrower than the node (mvbt_set_read_tile): T
lanes read an NW-pair node, NW / T pairs per lane. Finds and ranges at every
snapshot of a random history, on versioned and plain trees, against the
oracle and against the one-pair-per-lane reader; range counts above L; the
sorted batched form's read passes, also across read-tile changes between
graph launches; empty and one-leaf trees with degenerate ranges; and the
setter's argument checks."""
from __future__ import annotations

import numpy as np
import pytest

import mvgpubtree_amd as mv
import oracle
from mvgpubtree_amd import drivers
from mvgpubtree_amd import workloads as wl

from conftest import to_dev, to_np

pytestmark = pytest.mark.gpu

TILES = [(32, 16), (64, 16), (64, 32)]  # (node width, read tile lanes)


def history(rng, tree, ref, dev, versioned, rounds=6):
    """Rounds of inserts, value updates and erases (distinct keys per round);
    a snapshot after each round of a versioned tree. Returns the keys and
    the snapshot ids."""
    universe = wl.unique_keys(rng, 40_000)
    snaps = []
    for _ in range(rounds):
        idx = rng.choice(universe.size, size=int(rng.integers(2000, 9000)), replace=False)
        ins = idx[rng.random(idx.size) < 0.75]
        era = np.setdiff1d(idx, ins)
        k = universe[ins]
        v = rng.integers(0, 0xFFFFFFFF, size=k.size, dtype=np.uint64).astype(np.uint32)
        tree.insert(to_dev(k, dev), to_dev(v, dev))
        ref.insert(k, v)
        if era.size:
            tree.erase(to_dev(universe[era], dev))
            ref.erase(universe[era])
        if versioned:
            s = tree.take_snapshot()
            assert s == ref.take_snapshot()
            snaps.append(s)
    return universe, snaps


def read_all(tree, dq, dlo, dhi, L, ts):
    f = to_np(tree.find(dq, timestamp=ts))
    p, n = tree.range_query(dlo, dhi, L, timestamp=ts)
    return f, to_np(n), to_np(p, np.uint64)


@pytest.mark.parametrize("nw,lanes", TILES)
@pytest.mark.parametrize("kind", ["versioned", "plain"])
def test_read_tile_matches_oracle_at_every_snapshot(dev, nw, lanes, kind):
    versioned = kind == "versioned"
    rng = np.random.default_rng(nw * 7 + lanes + (1 if versioned else 0))
    cls = mv.GpuVersionedBTree if versioned else mv.GpuBlinkTree
    tree = cls(pool_nodes=1 << 16, node_width=nw)
    ref = oracle.MVMap()
    universe, snaps = history(rng, tree, ref, dev, versioned)
    q = np.concatenate([universe, wl.absent_keys(rng, 4000, np.sort(universe)), np.array([1, 0xFFFFFFFE], np.uint32)])
    lo = rng.integers(1, 0xFFFFFF00, size=5000, dtype=np.uint64).astype(np.uint32)
    lo[:3] = [1, 2, 0xFFFFFFF0]  # ranges at both ends of the key space
    hi = np.minimum(lo.astype(np.uint64) + wl.span_for(20, 30_000), 0xFFFFFFFE).astype(np.uint32)
    dq, dlo, dhi = to_dev(q, dev), to_dev(lo, dev), to_dev(hi, dev)
    L = 64
    for ts in snaps + [mv.LATEST]:
        tree.set_read_tile(0)
        base = read_all(tree, dq, dlo, dhi, L, ts)
        tree.set_read_tile(lanes)
        got = read_all(tree, dq, dlo, dhi, L, ts)
        want_f = ref.find(q, ts)
        rp, rn = ref.range_query(lo, hi, L, ts)
        assert np.array_equal(got[0], want_f), f"find@{ts}"
        assert np.array_equal(got[1], rn), f"counts@{ts}"
        assert np.array_equal(got[2], rp), f"pairs@{ts}"
        for a, b in zip(base, got):
            assert np.array_equal(a, b)
    assert tree.error_flags() == 0, hex(tree.error_flags())
    tree.close()


@pytest.mark.parametrize("nw,lanes", TILES)
def test_read_tile_range_counts_beyond_capacity(dev, nw, lanes):
    """Ranges of a few hundred keys with L = 16: the true count comes back,
    the first L pairs are written, and the overflow flag is raised."""
    rng = np.random.default_rng(5 + nw + lanes)
    keys = wl.unique_keys(rng, 60_000)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 16, node_width=nw)
    tree.insert(to_dev(keys, dev), to_dev(wl.value_of(keys), dev))
    ref = oracle.MVMap()
    ref.insert(keys, wl.value_of(keys))
    s = tree.take_snapshot()
    assert s == ref.take_snapshot()
    lo = rng.integers(1, 0xF0000000, size=3000, dtype=np.uint64).astype(np.uint32)
    hi = (lo.astype(np.uint64) + wl.span_for(300, keys.size)).astype(np.uint32)
    tree.set_read_tile(lanes)
    p, n = tree.range_query(to_dev(lo, dev), to_dev(hi, dev), 16, timestamp=s)
    rp, rn = ref.range_query(lo, hi, 16, s)
    assert rn.max() > 16
    assert np.array_equal(to_np(n), rn)
    assert np.array_equal(to_np(p, np.uint64), rp)
    assert tree.error_flags() == mv.FLAG_RANGE_OVERFLOW
    tree.clear_error_flags()
    tree.close()


@pytest.mark.parametrize("nw,lanes", TILES)
def test_sorted_form_read_passes_with_read_tile(dev, nw, lanes):
    w = wl.config4(seed=60 + nw + lanes, n_initial=200_000, num_batches=80, min_batch=32, max_batch=8192,
                   range_capacity=256)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 19, node_width=nw)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    tree.set_worker_mode(2)
    tree.set_read_tile(lanes)
    r = drivers.BatchedRunner(w.batches, w.range_capacity, dev, seed=lanes)
    res = to_np(r.run(tree, use_graph=True))
    want, wpairs, final = oracle.batch_expected(w.initial_keys, w.initial_values, r.ops_host, r.offsets_host,
                                                w.range_capacity, r.slots)
    bad = np.nonzero(res != want)[0]
    assert bad.size == 0, (bad[:10], r.ops_host[bad[:10]], res[bad[:10]], want[bad[:10]])
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * w.range_capacity], wpairs)
    assert tree.error_flags() == 0, hex(tree.error_flags())
    tree.close()


def test_read_tile_setter_checks_its_argument(dev):
    t16 = mv.GpuVersionedBTree(pool_nodes=1 << 12, node_width=16)
    t32 = mv.GpuVersionedBTree(pool_nodes=1 << 12, node_width=32)
    for bad in (8, 24, 32, 64):
        with pytest.raises(mv.MvbtError):
            t16.set_read_tile(bad)
    t16.set_read_tile(16)  # the node width itself
    t16.set_read_tile(0)
    for bad in (8, 64):
        with pytest.raises(mv.MvbtError):
            t32.set_read_tile(bad)
    t32.set_read_tile(16)
    t16.close()
    t32.close()


def test_read_tile_change_rebuilds_the_sorted_graphs(dev):
    """Calls through the launch graphs with one read tile, then another and
    back: each change rebuilds the graphs (their read passes are
    width-specific), and every call stays bit-exact."""
    w = wl.config4(seed=71, n_initial=150_000, num_batches=90, min_batch=32, max_batch=4096, range_capacity=256)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 19, node_width=32)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    tree.set_worker_mode(2)
    keys, vals = w.initial_keys, w.initial_values
    for i, (lanes, a, b) in enumerate([(0, 0, 30), (16, 30, 60), (0, 60, 90)]):
        tree.set_read_tile(lanes)
        r = drivers.BatchedRunner(w.batches[a:b], w.range_capacity, dev, seed=30 + i)
        tree.reserve_batches(r.nops, b - a)
        res = to_np(r.run(tree, use_graph=True))
        want, wpairs, final = oracle.batch_expected(keys, vals, r.ops_host, r.offsets_host, w.range_capacity, r.slots)
        assert np.array_equal(res, want), f"call {i} (tile {lanes})"
        assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * w.range_capacity], wpairs)
        keys, vals = final.keys, final.values
    assert tree.error_flags() == 0, hex(tree.error_flags())
    tree.close()


@pytest.mark.parametrize("nw,lanes", TILES)
def test_read_tile_edge_cases(dev, nw, lanes):
    """An empty tree, then a one-leaf tree: finds miss or hit, ranges with
    hi <= lo, lo = 0 and reserved bounds return what the oracle returns,
    and chunks whose length is not a multiple of the tile width end cleanly."""
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 12, node_width=nw)
    tree.set_read_tile(lanes)
    ref = oracle.MVMap()
    q = np.array([1, 2, 77, 0xFFFFFFFE], dtype=np.uint32)
    lo = np.array([0, 1, 50, 100, 100, 0xFFFFFFF0, 5], dtype=np.uint32)
    hi = np.array([10, 0xFFFFFFFF, 50, 99, 101, 0xFFFFFFFF, 0xFFFFFFFE], dtype=np.uint32)
    for step in range(2):
        s = tree.take_snapshot()
        assert s == ref.take_snapshot()
        for ts in (s, mv.LATEST):
            assert np.array_equal(to_np(tree.find(to_dev(q, dev), timestamp=ts)), ref.find(q, ts))
            p, n = tree.range_query(to_dev(lo, dev), to_dev(hi, dev), 8, timestamp=ts)
            rp, rn = ref.range_query(lo, hi, 8, ts)
            assert np.array_equal(to_np(n), rn), (step, ts)
            assert np.array_equal(to_np(p, np.uint64), rp), (step, ts)
        k = np.array([77, 100, 3, 0xFFFFFFFE, 1], dtype=np.uint32)  # fits one leaf
        tree.insert(to_dev(k, dev), to_dev(wl.value_of(k), dev))
        ref.insert(k, wl.value_of(k))
    rng = np.random.default_rng(9)
    odd = rng.integers(1, 200, size=37, dtype=np.uint64).astype(np.uint32)  # 37: not a multiple of any tile
    assert np.array_equal(to_np(tree.find(to_dev(odd, dev))), ref.find(odd))
    assert tree.error_flags() == 0, hex(tree.error_flags())
    tree.close()
