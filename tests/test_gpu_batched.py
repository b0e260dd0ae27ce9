"""GPU parity tests of the persistent batched form (mvbt_run_batches) and of
the plain per-batch path on the same mixed workloads (config-4 shape, scaled
down): both bit-exact against the oracle, which answers every read of a batch
from the state before the batch."""
from __future__ import annotations

import numpy as np
import pytest

import mvgpubtree_amd as mv
import oracle
from mvgpubtree_amd import drivers
from mvgpubtree_amd import workloads as wl

from conftest import to_dev, to_np

pytestmark = pytest.mark.gpu


def small_config4(seed, n_initial=200_000, num_batches=120, max_batch=8192):
    return wl.config4(seed=seed, n_initial=n_initial, num_batches=num_batches, min_batch=32,
                      max_batch=max_batch, range_capacity=256)


def build_tree(w, dev, nw=16):
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 19, node_width=nw)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    return tree


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("nw", [16, 32, 64])
def test_batched_form_matches_oracle(dev, nw, use_graph, mode):
    """mode 0: one operation per tile; mode 1: runs of operations per tile;
    mode 2: writes sorted by leaf, then every read at its snapshot."""
    w = small_config4(seed=40 + nw)
    tree = build_tree(w, dev, nw)
    tree.set_worker_mode(mode)
    r = drivers.BatchedRunner(w.batches, w.range_capacity, dev, seed=nw)
    s0 = tree.take_snapshot()
    tree.release_snapshot(s0)
    res = to_np(r.run(tree, use_graph=use_graph))
    want, wpairs, final = oracle.batch_expected(w.initial_keys, w.initial_values, r.ops_host,
                                                r.offsets_host, w.range_capacity, r.slots)
    assert int(to_np(r.base)[0]) == s0 + 1
    bad = np.nonzero(res != want)[0]
    assert bad.size == 0, (bad[:10], r.ops_host[bad[:10]], res[bad[:10]], want[bad[:10]])
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * w.range_capacity], wpairs)
    # clock advanced by one snapshot per batch
    assert tree.take_snapshot() == s0 + 1 + len(w.batches)
    exp = tree.export()
    assert np.array_equal(exp, final.keys.astype(np.uint64) | (final.values.astype(np.uint64) << np.uint64(32)))
    assert tree.error_flags() == 0
    tree.close()


@pytest.mark.parametrize("nw", [16, 32])
@pytest.mark.parametrize("waves", [1, 2])
def test_chunked_worker_multi_op_chunks(dev, waves, nw):
    """Few worker waves and large batches, so tiles claim runs of up to NW
    operations that span batch boundaries (batch size / tiles in the grid).
    At 32-pair nodes the worker also routes through its LDS copy of the top
    levels, which the job's own splits make stale."""
    w = wl.config4(seed=77 + waves, n_initial=300_000, num_batches=40, min_batch=32, max_batch=65536,
                   range_capacity=256)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 20, node_width=nw)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    tree.set_worker_mode(1)
    tree.set_worker_waves(waves)
    r = drivers.BatchedRunner(w.batches, w.range_capacity, dev, seed=waves)
    res = to_np(r.run(tree, use_graph=True))
    want, wpairs, final = oracle.batch_expected(w.initial_keys, w.initial_values, r.ops_host,
                                                r.offsets_host, w.range_capacity, r.slots)
    bad = np.nonzero(res != want)[0]
    assert bad.size == 0, (bad[:10], r.ops_host[bad[:10]], res[bad[:10]], want[bad[:10]])
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * w.range_capacity], wpairs)
    exp = tree.export()
    assert np.array_equal(exp, final.keys.astype(np.uint64) | (final.values.astype(np.uint64) << np.uint64(32)))
    assert tree.error_flags() == 0
    tree.close()


def test_per_batch_path_matches_oracle(dev):
    w = small_config4(seed=7, num_batches=60)
    tree = build_tree(w, dev)
    r = drivers.PerBatchRunner(w.batches, w.range_capacity, dev, keep_pairs=True)
    r.run(tree)
    got = [r.batch_results(i) for i in range(len(w.batches))]
    st = oracle.SortedState(w.initial_keys, w.initial_values)
    for (f, c, p), b in zip(got, w.batches):
        assert np.array_equal(f, st.find(b.find_keys))
        wp, wc = st.range_query(b.range_lo, b.range_hi, w.range_capacity)
        assert np.array_equal(c, wc)
        assert np.array_equal(p, wp)
        st.insert(b.insert_keys, b.insert_values)
        st.erase(b.erase_keys)
    assert tree.validate_tree_structure()[0] == len(st)
    assert tree.error_flags() == 0
    tree.close()


def test_graph_replays_with_new_inputs(dev):
    """Two launches through one captured graph (the second smaller, with new
    parameters), then a direct launch: state carries across all three."""
    w = small_config4(seed=9, num_batches=90)
    parts = [w.batches[:50], w.batches[50:70], w.batches[70:]]
    tree = build_tree(w, dev)
    st_keys, st_vals = w.initial_keys, w.initial_values
    for i, part in enumerate(parts):
        r = drivers.BatchedRunner(part, w.range_capacity, dev, seed=i)
        res = to_np(r.run(tree, use_graph=(i < 2)))
        want, wpairs, final = oracle.batch_expected(st_keys, st_vals, r.ops_host, r.offsets_host,
                                                    w.range_capacity, r.slots)
        assert np.array_equal(res, want), f"part {i}"
        assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * w.range_capacity], wpairs)
        st_keys, st_vals = final.keys, final.values
        tree.reclaim()
    assert tree.validate_tree_structure()[0] == st_keys.size
    assert tree.memory_stats().retired_pending == 0
    assert tree.error_flags() == 0
    tree.close()


def test_batched_snapshot_reads_older_than_the_run(dev):
    """A snapshot taken before a batched run still reads its own state."""
    w = small_config4(seed=11, num_batches=40)
    tree = build_tree(w, dev)
    s = tree.take_snapshot()
    r = drivers.BatchedRunner(w.batches, w.range_capacity, dev)
    r.run(tree)
    q = np.concatenate([w.initial_keys[:50000], np.concatenate([b.insert_keys for b in w.batches])])
    got = to_np(tree.find(to_dev(q, dev), timestamp=s))
    assert np.array_equal(got, oracle.SortedState(w.initial_keys, w.initial_values).find(q))
    assert tree.error_flags() == 0
    tree.close()


def test_autotuner_sweeps_and_caches(dev, tmp_path):
    """The autotuner times every (node width, worker mode, waves/CU) and picks
    the fastest; a second call with the same key reads the cache."""
    from mvgpubtree_amd import autotune
    w = wl.config4(seed=5, n_initial=100_000, num_batches=32, min_batch=32, max_batch=4096, range_capacity=256)
    waves, modes = (4, 8), (0, 1)
    groups = [w.batches[0::2], w.batches[1::2]]  # (warm, timed)
    cache = str(tmp_path / "tune.json")
    best, allr = autotune.tune_batched(w.initial_keys, w.initial_values, groups, w.range_capacity, dev,
                                       node_widths=(16, 32), waves=waves, modes=modes, cache=cache)
    assert len(allr) == 2 * len(waves) * len(modes)
    assert {(r.node_width, r.worker_mode, r.waves_per_cu) for r in allr} == \
        {(nw, m, wv) for nw in (16, 32) for m in modes for wv in waves}
    assert best.mops == max(r.mops for r in allr)
    again, _ = autotune.tune_batched(w.initial_keys, w.initial_values, groups, w.range_capacity, dev,
                                     node_widths=(16, 32), waves=waves, modes=modes, cache=cache)
    assert again == best


def test_autotuner_sweeps_read_tiles_of_the_sorted_form(dev):
    """Node width x read tile x waves/CU: the sorted form is timed with every
    read tile up to its node width (16- and 32-lane tiles at 32-pair nodes,
    16-, 32- and 64-lane tiles at 64-pair nodes)."""
    from mvgpubtree_amd import autotune
    w = wl.config4(seed=6, n_initial=100_000, num_batches=24, min_batch=32, max_batch=4096, range_capacity=256)
    groups = [w.batches[0::2], w.batches[1::2]]
    best, allr = autotune.tune_batched(w.initial_keys, w.initial_values, groups, w.range_capacity, dev,
                                       node_widths=(16, 32, 64), waves=(8,), modes=(2,))
    got = {(r.node_width, r.read_tile or r.node_width) for r in allr}
    assert got == {(16, 16), (32, 16), (32, 32), (64, 16), (64, 32), (64, 64)}
    assert best.mops == max(r.mops for r in allr)


def hot_leaf_batches(seed, n_batches, hot):
    """Batches whose writes all fall on a few hot leaves: every leaf's writes
    span hundreds of batches (the sorted form's leaf groups then run past a
    tile's window many times over and split repeatedly), values change on
    every re-insert, and erased keys come back."""
    rng = np.random.default_rng(seed)
    hot_keys = wl.unique_keys(rng, hot)
    hs = np.sort(hot_keys)
    span = int(wl.span_for(48, hot))
    batches = []
    for _ in range(n_batches):
        w = rng.choice(hot_keys, size=int(rng.integers(0, 48)), replace=False)
        k = int(rng.integers(0, w.size + 1))
        ik, ek = w[:k], w[k:]
        iv = rng.integers(1, 0xFFFFFFFE, size=ik.size, dtype=np.uint64).astype(np.uint32)
        fk = rng.choice(hot_keys, size=int(rng.integers(0, 12)))
        lo = hs[rng.integers(0, hs.size, size=int(rng.integers(0, 6)))]
        hi = np.minimum(lo.astype(np.uint64) + span, 0xFFFFFFFF).astype(np.uint32)
        batches.append(wl.Batch(insert_keys=ik, insert_values=iv, erase_keys=ek, find_keys=fk, range_lo=lo,
                                range_hi=hi))
    return hot_keys, batches


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("nw", [16, 32, 64])
def test_sorted_form_hot_leaves(dev, nw, use_graph):
    """Many batches writing the same few leaves, through the sorted form:
    per leaf, the writes of each batch are applied in batch order (each read
    sees exactly the earlier batches), against the oracle."""
    hot_keys, batches = hot_leaf_batches(91 + nw, 700, 3000)
    init = np.sort(hot_keys[:1500])
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 18, node_width=nw)
    tree.insert(to_dev(init, dev), to_dev(wl.value_of(init), dev))
    tree.set_worker_mode(2)
    L = 256
    r = drivers.BatchedRunner(batches, L, dev, seed=nw)
    res = to_np(r.run(tree, use_graph=use_graph))
    want, wpairs, final = oracle.batch_expected(init, wl.value_of(init), r.ops_host, r.offsets_host, L, r.slots)
    bad = np.nonzero(res != want)[0]
    assert bad.size == 0, (bad[:10], r.ops_host[bad[:10]], res[bad[:10]], want[bad[:10]])
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * L], wpairs)
    exp = tree.export()
    assert np.array_equal(exp, final.keys.astype(np.uint64) | (final.values.astype(np.uint64) << np.uint64(32)))
    assert tree.error_flags() == 0
    nk, _ = tree.validate_tree_structure()
    assert nk == final.keys.size
    tree.close()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batched_read_only_and_write_only_calls(dev, mode):
    """A call with no writes, then one with no reads, then one with reserved
    keys among its writes and reads (the writes fail with 0xFFFFFFFF and the
    invalid-key flag; reads of them miss)."""
    w = small_config4(seed=300 + mode, num_batches=30)
    tree = build_tree(w, dev, 32)
    tree.set_worker_mode(mode)
    L = w.range_capacity
    ks = np.sort(w.initial_keys)
    vs = wl.value_of(ks)
    ro = [wl.Batch(find_keys=b.find_keys, range_lo=b.range_lo, range_hi=b.range_hi) for b in w.batches]
    wo = [wl.Batch(insert_keys=b.insert_keys, insert_values=b.insert_values, erase_keys=b.erase_keys)
          for b in w.batches]
    for batches in (ro, wo):
        r = drivers.BatchedRunner(batches, L, dev, seed=mode)
        res = to_np(r.run(tree, use_graph=True))
        want, wpairs, final = oracle.batch_expected(ks, vs, r.ops_host, r.offsets_host, L, r.slots)
        assert np.array_equal(res, want)
        assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * L], wpairs)
        ks, vs = final.keys, final.values
    assert tree.error_flags() == 0
    bad = [wl.Batch(insert_keys=np.array([0, 5], np.uint32), insert_values=np.array([1, 2], np.uint32),
                    erase_keys=np.array([0xFFFFFFFF], np.uint32), find_keys=np.array([0, 5, 0xFFFFFFFF], np.uint32)),
           wl.Batch(find_keys=np.array([5], np.uint32), range_lo=np.array([0], np.uint32),
                    range_hi=np.array([0xFFFFFFFF], np.uint32))]
    r = drivers.BatchedRunner(bad, 1 << 20, dev, seed=0)
    res = to_np(r.run(tree, use_graph=False))
    want, _, final = oracle.batch_expected(ks, vs, r.ops_host, r.offsets_host, 1 << 20, r.slots)
    assert np.array_equal(res, want)
    assert tree.error_flags() == 4  # MVBT_FLAG_INVALID_KEY
    tree.close()


@pytest.mark.parametrize("live_snapshot", [False, True])
def test_sorted_form_in_windows_reclaims_between_them(dev, live_snapshot):
    """A call cut into many windows (small window size): results identical to
    the oracle. With no snapshot live the call frees the copies of each window
    once its reads are done (one timeline epoch per window, nothing left
    retired); a live snapshot keeps them all, and still reads its own state."""
    w = small_config4(seed=140, num_batches=200)
    tree = build_tree(w, dev, 32)
    tree.set_worker_mode(2)
    tree.set_batch_window(20_000)
    s = tree.take_snapshot() if live_snapshot else None
    epochs0 = len(tree.memory_series())
    r = drivers.BatchedRunner(w.batches, w.range_capacity, dev, seed=5)
    nwin = min(16, -(-r.nops // 20_000))
    res = to_np(r.run(tree, use_graph=True))
    want, wpairs, final = oracle.batch_expected(w.initial_keys, w.initial_values, r.ops_host, r.offsets_host,
                                                w.range_capacity, r.slots)
    bad = np.nonzero(res != want)[0]
    assert bad.size == 0, (bad[:10], r.ops_host[bad[:10]], res[bad[:10]], want[bad[:10]])
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * w.range_capacity], wpairs)
    st = tree.memory_stats()
    ser = tree.memory_series()
    if live_snapshot:
        assert len(ser) == epochs0
        assert st.retired_pending > 0
        q = w.initial_keys[:20000]
        assert np.array_equal(to_np(tree.find(to_dev(q, dev), timestamp=s)),
                              oracle.SortedState(w.initial_keys, w.initial_values).find(q))
        tree.release_snapshot(s)
    else:
        assert len(ser) == epochs0 + nwin
        assert st.retired_pending == 0
        assert int(ser.freed_nodes[-nwin:].sum()) > 0
    exp = tree.export()
    assert np.array_equal(exp, final.keys.astype(np.uint64) | (final.values.astype(np.uint64) << np.uint64(32)))
    assert tree.error_flags() == 0
    tree.close()


def test_sorted_form_fits_a_pool_that_holds_one_window_of_copies(dev):
    """A stream whose version copies together exceed the pool's spare slots
    runs to the end, bit-exact, when the call reclaims between windows."""
    w = small_config4(seed=141, n_initial=100_000, num_batches=300, max_batch=8192)
    n_ops = sum(b.num_ops for b in w.batches)
    writes = sum(b.insert_keys.size + b.erase_keys.size for b in w.batches)
    spare = writes // 3  # far fewer spare slots than the call makes copies
    tree = mv.GpuVersionedBTree(pool_nodes=100_000 // 10 + 2_000 + spare, node_width=32)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    tree.set_worker_mode(2)
    tree.set_batch_window(max(n_ops // 16, 1))
    assert tree.memory_stats().free_nodes < writes
    r = drivers.BatchedRunner(w.batches, w.range_capacity, dev, seed=6)
    res = to_np(r.run(tree, use_graph=True))
    assert tree.error_flags() == 0, hex(tree.error_flags())
    want, wpairs, final = oracle.batch_expected(w.initial_keys, w.initial_values, r.ops_host, r.offsets_host,
                                                w.range_capacity, r.slots)
    assert np.array_equal(res, want)
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * w.range_capacity], wpairs)
    assert tree.validate_tree_structure()[0] == final.keys.size
    tree.close()


def test_no_reclaim_flag_keeps_every_copy(dev):
    """MVBT_NO_RECLAIM: a call in many windows frees nothing between them
    (the timeline gains no epoch); a later mvbt_reclaim frees the copies."""
    w = small_config4(seed=142, num_batches=120)
    tree = build_tree(w, dev, 16)
    tree.set_worker_mode(2)
    tree.set_batch_window(10_000)
    e0 = len(tree.memory_series())
    r = drivers.BatchedRunner(w.batches, w.range_capacity, dev, seed=7)
    tree.run_batches(r.ops, r.offsets, results=r.results, pairs=r.pairs, pair_slots=max(r.slots, 1),
                     average_range_length=w.range_capacity, use_graph=True, no_reclaim=True)
    want, wpairs, _ = oracle.batch_expected(w.initial_keys, w.initial_values, r.ops_host, r.offsets_host,
                                            w.range_capacity, r.slots)
    assert np.array_equal(to_np(r.results)[: r.nops], want)
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * w.range_capacity], wpairs)
    assert len(tree.memory_series()) == e0
    pending = tree.memory_stats().retired_pending
    assert pending > 0
    tree.reclaim()
    ser = tree.memory_series()
    assert len(ser) == e0 + 1 and int(ser.freed_nodes[-1]) == pending
    assert tree.memory_stats().retired_pending == 0
    assert tree.error_flags() == 0
    tree.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_auto_mode_switches_by_call_size(dev, use_graph):
    """Mode 3 (default): calls below 16384 operations run on the persistent
    worker, larger ones in the sorted form; a run of small, large, small
    calls on one tree (graphs of both forms held at once) stays bit-exact."""
    w = small_config4(seed=143, num_batches=60, max_batch=2048)
    tree = build_tree(w, dev, 32)
    tree.reserve_batches(40_000, 64)
    keys, vals = w.initial_keys, w.initial_values
    cuts = [0, 4, 50, 60]  # batches of each call
    for i, (a, b) in enumerate(zip(cuts[:-1], cuts[1:])):
        r = drivers.BatchedRunner(w.batches[a:b], w.range_capacity, dev, seed=20 + i)
        assert (r.nops < 16384) == (i != 1), r.nops
        res = to_np(r.run(tree, use_graph=use_graph))
        want, wpairs, final = oracle.batch_expected(keys, vals, r.ops_host, r.offsets_host, w.range_capacity,
                                                    r.slots)
        bad = np.nonzero(res != want)[0]
        assert bad.size == 0, (i, bad[:10], r.ops_host[bad[:10]], res[bad[:10]], want[bad[:10]])
        assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * w.range_capacity], wpairs)
        keys, vals = final.keys, final.values
    exp = tree.export()
    assert np.array_equal(exp, keys.astype(np.uint64) | (vals.astype(np.uint64) << np.uint64(32)))
    assert tree.error_flags() == 0
    tree.close()
