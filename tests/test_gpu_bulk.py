"""GPU parity tests of the sorted bulk build (mvbt_bulk_build) against the oracle.

A bulk-built tree must answer exactly like a tree the same pairs were inserted
into, keep the structural invariants, and carry on correctly under later
inserts, erases and snapshots (the first writes split full nodes, including
the root's children and the root itself). Bad input is rejected and leaves the
tree new.
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


def data_lanes(kind, nw):
    return nw - 2 if kind == "versioned" else nw - 1


def expected_height(n, fill, root_cap):
    c, levels = -(-n // fill), 1
    while c > root_cap:
        c, levels = -(-c // fill), levels + 1
    return levels + 1


def new_tree(kind, nw, pool=1 << 16):
    if kind == "plain":
        return mv.GpuBlinkTree(pool_nodes=pool, node_width=nw)
    return mv.GpuVersionedBTree(pool_nodes=pool, node_width=nw)


def check_against(tree, ref, rng, dev, keys, ts=mv.LATEST, L=40):
    q = np.concatenate([keys, wl.absent_keys(rng, max(keys.size // 4, 16), np.sort(keys))]) if keys.size \
        else wl.absent_keys(rng, 16, np.empty(0, np.uint32))
    assert np.array_equal(to_np(tree.find(to_dev(q, dev), timestamp=ts)), ref.find(q, ts)), f"find@{ts}"
    lo = rng.integers(1, 0xFFFFFF00, size=2000, dtype=np.uint64).astype(np.uint32)
    lo[:8] = [1, 2, 3, 0x7FFFFFFF, 0x80000000, 0xFFFFFFF0, 0xFFFFFFFE, 5]
    hi = np.minimum(lo.astype(np.uint64) + wl.span_for(10, max(keys.size, 1)), 0xFFFFFFFF).astype(np.uint32)
    p, n = tree.range_query(to_dev(lo, dev), to_dev(hi, dev), L, timestamp=ts)
    rp, rn = ref.range_query(lo, hi, L, ts)
    assert np.array_equal(to_np(n), rn), f"counts@{ts}"
    assert np.array_equal(to_np(p, np.uint64), rp), f"pairs@{ts}"


@pytest.mark.parametrize("nw", WIDTHS)
@pytest.mark.parametrize("kind", ["plain", "versioned"])
def test_bulk_build_shapes(dev, kind, nw):
    """Every boundary of the level plan: one key, one full leaf, one more than
    a leaf, a root exactly full of leaves and one past it (a new level), and a
    multi-level tree; default fill, minimum fill and full nodes."""
    R = data_lanes(kind, nw)
    rng = np.random.default_rng(7 + nw)
    for fill in (0, 2, R):
        f = fill or (2 * R + 2) // 3
        for n in (1, f, f + 1, R * f, R * f + 1, 3 * R * f * f + 5, 20000):
            keys = np.sort(wl.unique_keys(rng, n))
            vals = wl.value_of(keys)
            tree = new_tree(kind, nw, pool=max(1 << 12, 3 * n))
            tree.bulk_build(to_dev(keys, dev), to_dev(vals, dev), fill=fill)
            nk, h = tree.validate_tree_structure()
            assert nk == n, (fill, n)
            assert h == expected_height(n, f, R), (fill, n, h)
            exp = tree.export()
            assert np.array_equal(exp, keys.astype(np.uint64) | (vals.astype(np.uint64) << 32))
            ref = oracle.MVMap()
            ref.build_sorted(keys, vals)
            check_against(tree, ref, rng, dev, keys)
            assert tree.error_flags() == 0
            tree.close()


@pytest.mark.parametrize("nw", WIDTHS)
@pytest.mark.parametrize("fill", [0, -1])  # -1: full nodes, so the first writes split
def test_bulk_then_versioned_history(dev, nw, fill):
    """Bulk build, snapshot, then rounds of inserts, updates and erases with a
    snapshot after each; every snapshot (including the one of the bulk-built
    state) is read back and compared word-for-word with the oracle."""
    fill = (nw - 2) if fill == -1 else fill
    rng = np.random.default_rng(40 + nw + fill)
    universe = wl.unique_keys(rng, 40000)
    base = np.sort(universe[:15000])
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 17, node_width=nw)
    ref = oracle.MVMap()
    tree.bulk_build(to_dev(base, dev), to_dev(wl.value_of(base), dev), fill=fill)
    ref.build_sorted(base, wl.value_of(base))
    snaps = [tree.take_snapshot()]
    assert snaps[0] == ref.take_snapshot()
    for rnd in range(6):
        idx = rng.choice(universe.size, size=int(rng.integers(1000, 8000)), replace=False)
        ins = idx[rng.random(idx.size) < 0.7]
        era = np.setdiff1d(idx, ins)
        k = universe[ins]
        v = rng.integers(0, 0xFFFFFFFF, size=k.size, dtype=np.uint64).astype(np.uint32)
        tree.insert(to_dev(k, dev), to_dev(v, dev))
        ref.insert(k, v)
        if era.size:
            tree.erase(to_dev(universe[era], dev))
            ref.erase(universe[era])
        s = tree.take_snapshot()
        assert s == ref.take_snapshot()
        snaps.append(s)
    for ts in snaps + [mv.LATEST]:
        check_against(tree, ref, rng, dev, universe, ts)
    nk, _ = tree.validate_tree_structure()
    assert nk == ref.size()
    assert np.array_equal(tree.export(), ref.export())
    assert tree.error_flags() == 0
    tree.close()


@pytest.mark.parametrize("nw", WIDTHS)
def test_bulk_then_plain_writes(dev, nw):
    """Plain tree: bulk build with full nodes, then inserts (new keys and
    updates) and erases; latest state against the oracle."""
    rng = np.random.default_rng(90 + nw)
    universe = wl.unique_keys(rng, 60000)
    base = np.sort(universe[:30000])
    tree = mv.GpuBlinkTree(pool_nodes=1 << 17, node_width=nw)
    ref = oracle.MVMap()
    tree.bulk_build(to_dev(base, dev), to_dev(wl.value_of(base), dev), fill=nw - 1)
    ref.build_sorted(base, wl.value_of(base))
    k = universe[rng.choice(universe.size, 30000, replace=False)]
    v = rng.integers(0, 0xFFFFFFFF, size=k.size, dtype=np.uint64).astype(np.uint32)
    tree.insert(to_dev(k, dev), to_dev(v, dev))
    ref.insert(k, v)
    e = universe[rng.choice(universe.size, 20000, replace=False)]
    tree.erase(to_dev(e, dev))
    ref.erase(e)
    check_against(tree, ref, rng, dev, universe)
    nk, _ = tree.validate_tree_structure()
    assert nk == ref.size()
    assert np.array_equal(tree.export(), ref.export())
    assert tree.error_flags() == 0
    tree.close()


def test_bulk_build_unsorted_input_keeps_last_value(dev):
    """sorted_input=False: device sort, repeated keys keep their last value,
    the same as inserting the pairs in order."""
    rng = np.random.default_rng(5)
    keys = wl.unique_keys(rng, 50000)
    dup = keys[rng.integers(0, keys.size, 20000)]
    all_k = np.concatenate([keys, dup])
    all_v = rng.integers(0, 0xFFFFFFFF, size=all_k.size, dtype=np.uint64).astype(np.uint32)
    perm = rng.permutation(all_k.size)
    all_k, all_v = all_k[perm], all_v[perm]
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 16)
    tree.bulk_build(to_dev(all_k, dev), to_dev(all_v, dev), sorted_input=False)
    ref = oracle.MVMap()
    ref.insert(all_k, all_v)
    assert np.array_equal(tree.export(), ref.export())
    check_against(tree, ref, rng, dev, keys)
    tree.close()


def test_bulk_build_config1_size(dev):
    """1M keys: finds half present / half absent, ranges, against the oracle."""
    rng = np.random.default_rng(11)
    keys = np.sort(wl.unique_keys(rng, 1 << 20))
    tree = mv.GpuBlinkTree(pool_nodes=1 << 18)
    tree.bulk_build(to_dev(keys, dev), to_dev(wl.value_of(keys), dev))
    ref = oracle.MVMap()
    ref.build_sorted(keys, wl.value_of(keys))
    check_against(tree, ref, rng, dev, keys)
    assert tree.validate_tree_structure()[0] == keys.size
    tree.close()


def test_bulk_build_rejects_bad_input(dev):
    rng = np.random.default_rng(3)
    keys = np.sort(wl.unique_keys(rng, 5000))
    vals = wl.value_of(keys)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 14)
    bad_cases = {
        "unsorted": keys[::-1].copy(),
        "duplicate": np.concatenate([keys[:10], keys[9:]]),
        "zero key": np.concatenate([[0], keys[1:]]).astype(np.uint32),
        "max key": np.concatenate([keys[:-1], [0xFFFFFFFF]]).astype(np.uint32),
    }
    for name, k in bad_cases.items():
        with pytest.raises(mv.MvbtError):
            tree.bulk_build(to_dev(k, dev), to_dev(wl.value_of(k), dev))
        assert tree.validate_tree_structure()[0] == 0, name
    with pytest.raises(mv.MvbtError):
        tree.bulk_build(to_dev(keys, dev), to_dev(vals, dev), fill=1)
    with pytest.raises(mv.MvbtError):
        tree.bulk_build(to_dev(keys, dev), to_dev(vals, dev), fill=15)
    small = mv.GpuVersionedBTree(pool_nodes=64)
    with pytest.raises(mv.MvbtError):
        small.bulk_build(to_dev(keys, dev), to_dev(vals, dev))
    small.close()
    # rejected calls left the tree new: a valid build still works
    tree.bulk_build(to_dev(keys, dev), to_dev(vals, dev))
    assert tree.validate_tree_structure()[0] == keys.size
    # and a second build (the tree is no longer new) is refused
    with pytest.raises(mv.MvbtError):
        tree.bulk_build(to_dev(keys, dev), to_dev(vals, dev))
    tree.close()
    used = mv.GpuBlinkTree(pool_nodes=1 << 14)
    used.insert(to_dev(keys[:5], dev), to_dev(vals[:5], dev))
    with pytest.raises(mv.MvbtError):
        used.bulk_build(to_dev(keys, dev), to_dev(vals, dev))
    used.close()
