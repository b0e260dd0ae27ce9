"""GPU parity tests: the HIP tree through the C ABI against the oracle.

Bit-exact comparison for every result word (values, counts, pairs). Cases:
the reference's own unit tests (golden fixtures), both tree kinds and all node
widths, random interleavings of snapshots/inserts/updates/erases read at
every live snapshot, concurrent snapshot readers beside writers on another
stream, reclamation, and error paths (reserved keys, pool exhaustion).
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import mvgpubtree_amd as mv
import oracle
from mvgpubtree_amd import workloads as wl

from conftest import to_dev, to_np

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_unittests.npz")
MISS = 0xFFFFFFFF
WIDTHS = (16, 32, 64)


def golden_cases():
    g = np.load(GOLDEN)
    names = sorted({k.split("__")[0] for k in g.files})
    return [(n, {k.split("__")[1]: g[k] for k in g.files if k.startswith(n + "__")}) for n in names]


def sync():
    import torch
    torch.cuda.synchronize()


def assert_clean(tree):
    assert tree.error_flags() == 0, hex(tree.error_flags())


@pytest.mark.parametrize("nw", WIDTHS)
@pytest.mark.parametrize("kind", ["plain", "versioned"])
@pytest.mark.parametrize("name,c", golden_cases(), ids=[n for n, _ in golden_cases()])
def test_reference_btree_unittests(dev, kind, nw, name, c):
    """unittest_btree.cu: Validation, FindExist, FindNotExist, EraseAll,
    EraseNone, EraseAllInsertAll."""
    cls = mv.GpuBlinkTree if kind == "plain" else mv.GpuVersionedBTree
    tree = cls(pool_nodes=1 << 14, node_width=nw)
    keys, vals = to_dev(c["keys"], dev), to_dev(c["values"], dev)
    tree.insert(keys, vals)
    nk, h = tree.validate_tree_structure()
    assert nk == c["keys"].size
    assert np.array_equal(to_np(tree.find(keys)), c["find_exist"])
    assert np.array_equal(to_np(tree.find(to_dev(c["absent"], dev))), c["find_not_exist"])
    tree.erase(to_dev(c["absent"], dev))
    assert np.array_equal(to_np(tree.find(keys)), c["find_exist"])
    tree.erase(keys)
    assert np.array_equal(to_np(tree.find(keys)), c["find_not_exist"])
    assert tree.validate_tree_structure()[0] == 0
    tree.insert(keys, vals)
    assert np.array_equal(to_np(tree.find(keys)), c["find_exist"])
    assert tree.validate_tree_structure()[0] == c["keys"].size
    assert_clean(tree)
    tree.close()


@pytest.mark.parametrize("nw", WIDTHS)
@pytest.mark.parametrize("name,c", golden_cases(), ids=[n for n, _ in golden_cases()])
def test_reference_versioning_unittests(dev, nw, name, c):
    """unittest_versioning.cu: FindExist/FindNotExist/RangeQueryTest at the
    snapshot taken after the first half, then after the second half."""
    half, L = int(c["half"][0]), int(c["L"][0])
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 14, node_width=nw)
    tree.insert(to_dev(c["keys"][:half], dev), to_dev(c["values"][:half], dev))
    s0 = tree.take_snapshot()
    tree.insert(to_dev(c["keys"][half:], dev), to_dev(c["values"][half:], dev))
    s1 = tree.take_snapshot()
    assert (s0, s1) == (0, 1)
    assert np.array_equal(to_np(tree.find(to_dev(c["keys"][:half], dev), timestamp=s0)),
                          c["find_s0_first_half"])
    assert np.array_equal(to_np(tree.find(to_dev(c["keys"][half:], dev), timestamp=s0)),
                          c["find_s0_second_half"])
    assert np.array_equal(to_np(tree.find(to_dev(c["keys"], dev))), c["find_exist"])
    assert np.array_equal(to_np(tree.find(to_dev(c["absent"], dev))), c["find_not_exist"])
    p, n = tree.range_query(to_dev(c["lower"][:half], dev), to_dev(c["upper"][:half], dev), L,
                            timestamp=s0)
    assert np.array_equal(to_np(n), c["range_s0_counts"])
    assert np.array_equal(to_np(p, np.uint64), c["range_s0_pairs"])
    p, n = tree.range_query(to_dev(c["lower"], dev), to_dev(c["upper"], dev), L, timestamp=s1)
    assert np.array_equal(to_np(n), c["range_s1_counts"])
    assert np.array_equal(to_np(p, np.uint64), c["range_s1_pairs"])
    assert_clean(tree)
    tree.close()


def test_cpp_api_example():
    """The C++ host API (include/gpu_btree.h) over pinned host buffers."""
    exe = os.path.join(os.path.dirname(mv.__file__), "capi_example")
    for args in (["1024", "2"], ["5000", "32"]):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "OK" in r.stdout


@pytest.mark.parametrize("nw", WIDTHS)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_history_all_snapshots(dev, nw, seed):
    """Random rounds of inserts, value updates and erases (distinct keys within
    a round), a snapshot after each round; every snapshot is read back with
    finds and ranges and compared word-for-word with the oracle."""
    rng = np.random.default_rng(100 + seed)
    universe = wl.unique_keys(rng, 30000)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 16, node_width=nw)
    ref = oracle.MVMap()
    snaps = []
    present = np.zeros(universe.size, dtype=bool)
    for rnd in range(8):
        idx = rng.choice(universe.size, size=int(rng.integers(500, 6000)), replace=False)
        ins = idx[rng.random(idx.size) < 0.7]
        era = np.setdiff1d(idx, ins)
        k = universe[ins]
        v = rng.integers(0, 0xFFFFFFFF, size=k.size, dtype=np.uint64).astype(np.uint32)
        tree.insert(to_dev(k, dev), to_dev(v, dev))
        ref.insert(k, v)
        if era.size:
            tree.erase(to_dev(universe[era], dev))
            ref.erase(universe[era])
        present[ins] = True
        present[era] = False
        s = tree.take_snapshot()
        assert s == ref.take_snapshot()
        snaps.append(s)
    q = np.concatenate([universe, wl.absent_keys(rng, 3000, np.sort(universe))])
    lo = rng.integers(1, 0xFFFFFF00, size=4000, dtype=np.uint64).astype(np.uint32)
    hi = np.minimum(lo.astype(np.uint64) + wl.span_for(12, 20000), 0xFFFFFFFF).astype(np.uint32)
    dq, dlo, dhi = to_dev(q, dev), to_dev(lo, dev), to_dev(hi, dev)
    for ts in snaps + [mv.LATEST]:
        assert np.array_equal(to_np(tree.find(dq, timestamp=ts)), ref.find(q, ts)), f"find@{ts}"
        p, n = tree.range_query(dlo, dhi, 48, timestamp=ts)
        rp, rn = ref.range_query(lo, hi, 48, ts)
        assert np.array_equal(to_np(n), rn), f"counts@{ts}"
        assert np.array_equal(to_np(p, np.uint64), rp), f"pairs@{ts}"
    nk, _ = tree.validate_tree_structure()
    assert nk == ref.size()
    exp = tree.export()
    assert np.array_equal(exp, ref.export())
    assert_clean(tree)
    tree.close()


def test_config1_plain_full_size(dev):
    """Config 1 at full size: 1M unique uniform keys, 1M finds (50% hits)."""
    c = wl.config1(seed=1, n=1_000_000)
    tree = mv.GpuBlinkTree(pool_nodes=1 << 20)
    tree.insert(to_dev(c.keys, dev), to_dev(c.values, dev))
    got = to_np(tree.find(to_dev(c.queries, dev)))
    want = oracle.SortedState(c.keys, c.values).find(c.queries)
    assert np.array_equal(got, want)
    assert (got != MISS).sum() == 500_000
    nk, h = tree.validate_tree_structure()
    assert nk == 1_000_000
    assert_clean(tree)
    tree.close()


@pytest.mark.parametrize("kind", ["plain", "versioned"])
def test_large_single_launch_build(dev, kind):
    """8M random keys into a one-key tree in one insert launch: the tree
    grows from one leaf while the launch runs, so routing done early in a
    chunk goes stale by millions of leaves; writes must still land (no step
    bound, no error flag) and every key must be found with its value. (An
    insert into a fresh tree runs in growing stages instead; the first key
    goes in alone so that this one is a single launch.)"""
    rng = np.random.default_rng(21)
    keys = wl.unique_keys(rng, 8_000_000)
    vals = wl.value_of(keys)
    cls = mv.GpuBlinkTree if kind == "plain" else mv.GpuVersionedBTree
    tree = cls(pool_nodes=1 << 21)
    tree.insert(to_dev(keys[:1], dev), to_dev(vals[:1], dev))
    tree.insert(to_dev(keys[1:], dev), to_dev(vals[1:], dev))
    sync()
    assert_clean(tree)
    q = np.concatenate([keys[::8], wl.absent_keys(rng, 500_000, np.sort(keys))])
    got = to_np(tree.find(to_dev(q, dev)))
    want = oracle.SortedState(keys, vals).find(q)
    assert np.array_equal(got, want)
    nk, _ = tree.validate_tree_structure()
    assert nk == keys.size
    tree.close()


@pytest.mark.parametrize("nw", WIDTHS)
def test_concurrent_snapshot_readers_beside_writers(dev, nw):
    """Finds and ranges at snapshot s run on one stream while inserts and
    erases (stamped after s) run on another: results must equal the oracle
    at s exactly, whatever the interleaving."""
    import torch
    rng = np.random.default_rng(7)
    keys = wl.unique_keys(rng, 300000)
    base, extra = keys[:200000], keys[200000:]
    vals = wl.value_of(keys)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 18, node_width=nw)
    tree.insert(to_dev(base, dev), to_dev(vals[:200000], dev))
    s = tree.take_snapshot()
    ref = oracle.SortedState(base, vals[:200000])
    q = np.concatenate([keys, wl.absent_keys(rng, 50000, np.sort(keys))])
    rng.shuffle(q)
    lo = rng.integers(1, 0xFFFFFF00, size=100000, dtype=np.uint64).astype(np.uint32)
    hi = np.minimum(lo.astype(np.uint64) + wl.span_for(32, 200000), 0xFFFFFFFF).astype(np.uint32)
    dq, dlo, dhi = to_dev(q, dev), to_dev(lo, dev), to_dev(hi, dev)
    dins, dinsv = to_dev(extra, dev), to_dev(vals[200000:], dev)
    dera = to_dev(base[::3].copy(), dev)
    sync()
    sr, sw = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(sw):
        tree.insert(dins, dinsv, stream=sw)
        tree.erase(dera, stream=sw)
    with torch.cuda.stream(sr):
        got = tree.find(dq, timestamp=s, stream=sr, concurrent=True)
        p, n = tree.range_query(dlo, dhi, 96, timestamp=s, stream=sr, concurrent=True)
    sync()
    assert np.array_equal(to_np(got), ref.find(q))
    rp, rn = ref.range_query(lo, hi, 96)
    assert np.array_equal(to_np(n), rn)
    assert np.array_equal(to_np(p, np.uint64), rp)
    # the writes themselves landed
    ref.insert(extra, vals[200000:])
    ref.erase(base[::3])
    assert np.array_equal(to_np(tree.find(dq)), ref.find(q))
    assert tree.validate_tree_structure()[0] == len(ref)
    assert_clean(tree)
    tree.close()


def test_reclamation_frees_unreadable_versions(dev):
    """Config-3 shape (small): after each batch the pre-batch snapshot is
    released and reclaimed; memory stays bounded and reads stay exact."""
    w = wl.config3(seed=11, n_initial=200000, num_batches=6, ops_per_batch=40000)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 18)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    ref = oracle.SortedState(w.initial_keys, w.initial_values)
    base_live = tree.memory_stats().live_nodes
    for b in w.batches:
        s = tree.take_snapshot()
        got = tree.find(to_dev(b.find_keys, dev), timestamp=s)
        tree.erase(to_dev(b.erase_keys, dev))
        assert np.array_equal(to_np(got), ref.find(b.find_keys))
        ref.erase(b.erase_keys)
        # the erases are invisible at s and visible at latest
        assert np.all(to_np(tree.find(to_dev(b.erase_keys, dev), timestamp=s)) != MISS)
        assert np.all(to_np(tree.find(to_dev(b.erase_keys, dev))) == MISS)
        tree.release_snapshot(s)
        tree.reclaim()
    st = tree.memory_stats()
    assert st.reclaimed_nodes > 0
    assert st.retired_pending == 0
    assert st.live_nodes <= base_live + 2000  # only split siblings remain
    assert st.peak_live_nodes < base_live + 40000
    assert tree.validate_tree_structure()[0] == len(ref)
    assert_clean(tree)
    tree.close()


def test_live_snapshot_keeps_its_versions(dev):
    """Reclaim with an older snapshot still live must keep what it reads."""
    rng = np.random.default_rng(5)
    k = wl.unique_keys(rng, 50000)
    v = wl.value_of(k)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 16)
    tree.insert(to_dev(k, dev), to_dev(v, dev))
    s0 = tree.take_snapshot()
    tree.erase(to_dev(k[:25000], dev))
    s1 = tree.take_snapshot()
    tree.insert(to_dev(k[:10000], dev), to_dev(v[:10000] ^ np.uint32(1), dev))
    tree.release_snapshot(s1)
    tree.reclaim()
    got = to_np(tree.find(to_dev(k, dev), timestamp=s0))
    assert np.array_equal(got, v)
    tree.release_snapshot(s0)
    tree.reclaim()
    latest = to_np(tree.find(to_dev(k, dev)))
    want = v.copy()
    want[:10000] ^= 1
    want[10000:25000] = MISS
    assert np.array_equal(latest, want)
    assert tree.memory_stats().retired_pending == 0
    assert_clean(tree)
    tree.close()


def test_reserved_keys_and_pool_exhaustion_are_flagged(dev):
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 12)
    bad = np.array([0, 0xFFFFFFFF, 5], dtype=np.uint32)
    tree.insert(to_dev(bad, dev), to_dev(bad, dev))
    assert tree.error_flags() == mv.FLAG_INVALID_KEY
    assert np.array_equal(to_np(tree.find(to_dev(bad, dev))), np.array([MISS, MISS, 5], np.uint32))
    tree.clear_error_flags()
    tree.close()
    small = mv.GpuBlinkTree(pool_nodes=64)
    rng = np.random.default_rng(3)
    k = wl.unique_keys(rng, 20000)
    small.insert(to_dev(k, dev), to_dev(k, dev))
    assert small.error_flags() & mv.FLAG_POOL_EXHAUSTED
    nk, _ = small.validate_tree_structure()  # still a valid tree
    assert 0 < nk < k.size
    got = to_np(small.find(to_dev(k, dev)))
    hits = got != MISS
    assert np.array_equal(got[hits], k[hits]) and hits.sum() == nk
    small.close()


def test_device_api_user_kernels():
    """User kernels on include/mvbt_device.hpp: cooperative_insert, an
    in-kernel take_snapshot, then tiles inserting while other tiles read that
    snapshot with cooperative_find / concurrent_cooperative_range_query."""
    exe = os.path.join(os.path.dirname(mv.__file__), "device_api_example")
    for n in ("1000", "200000"):
        r = subprocess.run([exe, n], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "OK" in r.stdout


@pytest.mark.parametrize("kind", ["plain", "versioned"])
def test_export_dot(dev, kind):
    """Graphviz text (reference plot_dot): every latest key appears once in a
    leaf record, child and sibling edges are present, versioned leaves show
    their version chains, and max_nodes bounds the output."""
    rng = np.random.default_rng(8)
    keys = wl.unique_keys(rng, 400)
    cls = mv.GpuBlinkTree if kind == "plain" else mv.GpuVersionedBTree
    tree = cls(pool_nodes=1 << 12)
    tree.insert(to_dev(keys[:300], dev), to_dev(wl.value_of(keys[:300]), dev))
    if kind == "versioned":
        tree.take_snapshot()
    tree.insert(to_dev(keys[300:], dev), to_dev(wl.value_of(keys[300:]), dev))
    dot = tree.to_dot()
    assert dot.startswith("digraph") and dot.rstrip().endswith("}")
    leaf_lines = [l for l in dot.splitlines() if " leaf" in l and "label=" in l]
    for k in keys[::17]:
        field = f">{int(k)}:{int(wl.value_of(np.array([k]))[0])}|"  # record field "<pN>key:value|"
        assert sum(field in l for l in leaf_lines) == 1, k
    assert ":h;" in dot  # child edges
    if kind == "versioned":
        assert "older" in dot and "(copy)" in dot
    small = tree.to_dot(max_nodes=3)
    assert small.count("label=") <= 3
    tree.close()


@pytest.mark.parametrize("kind", ["plain", "versioned"])
def test_concurrent_writers_on_two_streams_route_through_stale_caches(dev, kind):
    """32-pair nodes, a tree large enough that the insert launch routes
    through its LDS copy of the top levels: two streams insert disjoint key
    sets at once, so each launch's images go stale under the other's splits
    (they must route left and repair by right moves). The result must equal
    the oracle and the structure must validate."""
    import torch
    rng = np.random.default_rng(21)
    keys = wl.unique_keys(rng, 700_000)
    base, a, b = keys[:300_000], keys[300_000:500_000], keys[500_000:]
    cls = mv.GpuVersionedBTree if kind == "versioned" else mv.GpuBlinkTree
    tree = cls(pool_nodes=1 << 17, node_width=32)
    tree.insert(to_dev(base, dev), to_dev(wl.value_of(base), dev))
    sync()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    da, dva = to_dev(a, dev), to_dev(wl.value_of(a), dev)
    db, dvb = to_dev(b, dev), to_dev(wl.value_of(b), dev)
    for part in range(4):  # batches of 50K: small against the tree, so the cache is used
        sl = slice(part * 50_000, (part + 1) * 50_000)
        tree.insert(da[sl], dva[sl], stream=s1)
        tree.insert(db[sl], dvb[sl], stream=s2)
    sync()
    ref = oracle.SortedState(keys, wl.value_of(keys))
    assert tree.validate_tree_structure()[0] == keys.size
    assert np.array_equal(tree.export(), (ref.keys.astype(np.uint64) | (ref.values.astype(np.uint64) << 32)))
    q = np.concatenate([keys[::7], wl.absent_keys(rng, 20_000, np.sort(keys))])
    assert np.array_equal(to_np(tree.find(to_dev(q, dev))), ref.find(q))
    assert_clean(tree)
    tree.close()


@pytest.mark.parametrize("nw", WIDTHS)
def test_plain_concurrent_ranges_beside_inserts(dev, nw):
    """Plain tree, MVBT_CONCURRENT ranges: with no writer running they equal
    the oracle; on one stream while inserts run on another, every range
    counts at least the keys present before the inserts and at most those
    present after them (any linearization, as the reference's tolerant
    validators allow, validation.hpp:266-369), and every pair it returns is a
    key of the final state inside the range, with its value, in key order."""
    import torch
    rng = np.random.default_rng(77 + nw)
    keys = wl.unique_keys(rng, 300000)
    base, extra = keys[:200000], keys[200000:]
    tree = mv.GpuBlinkTree(pool_nodes=1 << 18, node_width=nw)
    tree.insert(to_dev(base, dev), to_dev(wl.value_of(base), dev))
    before = oracle.SortedState(base, wl.value_of(base))
    after = oracle.SortedState(keys, wl.value_of(keys))
    lo = rng.integers(1, 0xFFFFFF00, size=60000, dtype=np.uint64).astype(np.uint32)
    hi = np.minimum(lo.astype(np.uint64) + wl.span_for(32, 200000), 0xFFFFFFFF).astype(np.uint32)
    dlo, dhi = to_dev(lo, dev), to_dev(hi, dev)
    L = 128
    p, n = tree.range_query(dlo, dhi, L, concurrent=True)
    rp, rn = before.range_query(lo, hi, L)
    assert np.array_equal(to_np(n), rn)
    assert np.array_equal(to_np(p, np.uint64), rp)
    sync()
    sr, sw = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(sw):
        tree.insert(to_dev(extra, dev), to_dev(wl.value_of(extra), dev), stream=sw)
    with torch.cuda.stream(sr):
        p, n = tree.range_query(dlo, dhi, L, stream=sr, concurrent=True)
    sync()
    got_n = to_np(n).astype(np.int64)
    _, n_after = after.range_query(lo, hi, L)
    assert np.all(got_n >= rn) and np.all(got_n <= n_after)
    assert got_n.max() <= L
    pairs = to_np(p, np.uint64).reshape(lo.size, L)
    used = np.arange(L)[None, :] < got_n[:, None]
    k = (pairs & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    v = (pairs >> np.uint64(32)).astype(np.uint32)
    ku = k[used]
    assert np.all(np.isin(ku, keys))
    assert np.all(v[used] == wl.value_of(ku))
    assert np.all((k >= lo[:, None]) | ~used) and np.all((k < hi[:, None]) | ~used)
    assert np.all((np.diff(k.astype(np.int64), axis=1) > 0) | ~used[:, 1:])
    assert np.all(pairs[~used] == np.uint64(0xFFFFFFFFFFFFFFFF))
    assert_clean(tree)
    tree.close()
