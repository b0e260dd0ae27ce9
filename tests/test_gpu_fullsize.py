"""GPU parity at the configurations' stated sizes (the benchmark's own seeded
workloads), bit-exact against the oracle:

* config 2: 40M-key versioned tree (32-pair nodes read by 16-lane tiles, as
  measured), 10 batches of 500K range-32 queries at the pre-batch snapshot +
  500K inserts: every count and every pair, then the final key set;
* config 3: 40M-key tree, 10 batches of 500K erases + 500K finds at the
  pre-batch snapshot, reclaiming after each batch: every find, the final key
  set, and reclaimed memory;
* config 4: 40M-key tree (32-pair nodes, 16-lane read passes, as measured),
  4096 batches (log-uniform sizes in [32, 65536], 60/20/10/10
  find/insert/erase/range) in one batched launch: every result word and
  every range pair against the write-once replay;
* the config-2 and config-3 batches through the one-launch mixed calls
  (concurrent_insert_range, concurrent_find_erase), 32-pair nodes: every
  read of a call against the state before it;
* config 5: 500M-key tree (32-pair nodes, 16-lane read passes, as measured),
  requests of 1..16384 ops (80/10/10 find/range/insert) through the served
  loop: every result and pair against the write-once replay of the same
  request stream.
"""
from __future__ import annotations

import sys
import time

import numpy as np
import pytest

import mvgpubtree_amd as mv
import oracle
from mvgpubtree_amd import drivers
from mvgpubtree_amd import workloads as wl

from conftest import to_dev, to_np

pytestmark = pytest.mark.gpu
MISS = 0xFFFFFFFF


def note(*a):
    print(*a, file=sys.stderr, flush=True)


def sync():
    import torch
    torch.cuda.synchronize()


def packed(state):
    return state.keys.astype(np.uint64) | (state.values.astype(np.uint64) << np.uint64(32))


def test_config2_full_size(dev):
    t0 = time.time()
    w = wl.config2(seed=2)  # 40M keys, 10 x 1M ops: the benchmark's workload
    L = w.range_capacity
    tree = mv.GpuVersionedBTree(pool_nodes=int((40_000_000 + 10 * 1_000_000) / 4) + (1 << 20), node_width=32)
    tree.set_read_tile(16)  # as measured: 16-lane read tiles, two pairs per lane
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    ref = oracle.SortedState(w.initial_keys, w.initial_values)
    note(f"config2: built in {time.time() - t0:.1f}s")
    for i, b in enumerate(w.batches):
        s = tree.take_snapshot()
        p, n = tree.range_query(to_dev(b.range_lo, dev), to_dev(b.range_hi, dev), L, timestamp=s)
        tree.insert(to_dev(b.insert_keys, dev), to_dev(b.insert_values, dev))
        tree.release_snapshot(s)
        tree.reclaim()
        rp, rn = ref.range_query(b.range_lo, b.range_hi, L)
        assert np.array_equal(to_np(n), rn), f"batch {i}: counts"
        got = to_np(p, np.uint64)
        assert np.array_equal(got, rp), f"batch {i}: pairs"
        assert rn.max() <= L
        ref.insert(b.insert_keys, b.insert_values)
        note(f"config2: batch {i} ok ({time.time() - t0:.1f}s)")
        del p, n, got, rp
    assert np.array_equal(tree.export(), packed(ref))
    assert tree.error_flags() == 0
    tree.close()


def test_config3_full_size(dev):
    t0 = time.time()
    w = wl.config3(seed=3)  # 40M keys, 10 x (500K erases + 500K finds)
    tree = mv.GpuVersionedBTree(pool_nodes=int(40_000_000 / 4) + 10 * 1_000_000 + (1 << 20))
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    ref = oracle.SortedState(w.initial_keys, w.initial_values)
    live0 = tree.memory_stats().live_nodes
    nk, height = tree.validate_tree_structure()
    assert nk == w.initial_keys.size
    # the height the algorithmic byte counts assume (16-pair nodes, 40M random keys)
    assert height == 8, height
    note(f"config3: built in {time.time() - t0:.1f}s, height {height}")
    for i, b in enumerate(w.batches):
        s = tree.take_snapshot()
        f = tree.find(to_dev(b.find_keys, dev), timestamp=s)
        tree.erase(to_dev(b.erase_keys, dev))
        tree.release_snapshot(s)
        tree.reclaim()
        want = ref.find(b.find_keys)
        assert np.array_equal(to_np(f), want), f"batch {i}"
        assert (want != MISS).sum() == b.find_keys.size // 2  # half hits
        assert ref.erase(b.erase_keys) == b.erase_keys.size
        note(f"config3: batch {i} ok ({time.time() - t0:.1f}s)")
    st = tree.memory_stats()
    assert st.retired_pending == 0 and st.reclaimed_nodes > 0
    assert st.live_nodes <= live0 + 20_000  # version copies all freed; erases never merge leaves
    assert np.array_equal(tree.export(), packed(ref))
    assert tree.error_flags() == 0
    tree.close()


def test_mixed_calls_full_size(dev):
    """The config-2 and config-3 batches at 40M keys through the one-launch
    mixed calls (concurrent_insert_range, concurrent_find_erase) on 32-pair
    trees, as measured: every read of a call against the state before it,
    then the final key set. Four calls of each kind (the first ones of the
    benchmark's workloads), reclaiming after each."""
    t0 = time.time()
    w = wl.config2(seed=2, num_batches=4)
    L = w.range_capacity
    tree = mv.GpuVersionedBTree(pool_nodes=int((40_000_000 + 4 * 1_000_000) / 4) + (1 << 20), node_width=32)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    ref = oracle.SortedState(w.initial_keys, w.initial_values)
    for i, b in enumerate(w.batches):
        p, n = tree.concurrent_insert_range(to_dev(b.insert_keys, dev), to_dev(b.insert_values, dev),
                                            to_dev(b.range_lo, dev), to_dev(b.range_hi, dev), L)
        tree.reclaim()
        rp, rn = ref.range_query(b.range_lo, b.range_hi, L)
        assert np.array_equal(to_np(n), rn), f"insert_range call {i}: counts"
        assert np.array_equal(to_np(p, np.uint64), rp), f"insert_range call {i}: pairs"
        ref.insert(b.insert_keys, b.insert_values)
        note(f"mixed insert_range: call {i} ok ({time.time() - t0:.1f}s)")
        del p, n, rp
    assert np.array_equal(tree.export(), packed(ref))
    assert tree.error_flags() == 0, hex(tree.error_flags())
    tree.close()
    del ref

    w = wl.config3(seed=3, num_batches=4)
    tree = mv.GpuVersionedBTree(pool_nodes=int(40_000_000 / 8) + 4 * 1_000_000 + (1 << 20), node_width=32)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    ref = oracle.SortedState(w.initial_keys, w.initial_values)
    for i, b in enumerate(w.batches):
        f = tree.concurrent_find_erase(to_dev(b.find_keys, dev), to_dev(b.erase_keys, dev))
        tree.reclaim()
        want = ref.find(b.find_keys)
        assert np.array_equal(to_np(f)[: b.find_keys.size], want), f"find_erase call {i}"
        assert ref.erase(b.erase_keys) == b.erase_keys.size
        note(f"mixed find_erase: call {i} ok ({time.time() - t0:.1f}s)")
    assert np.array_equal(tree.export(), packed(ref))
    assert tree.error_flags() == 0, hex(tree.error_flags())
    tree.close()


def compare_batched(res_dev, pairs_dev, L, ops, off, replay, chunk=1 << 21):
    """Every result word, and every range's L pair slots, against `replay`."""
    import torch
    res = to_np(res_dev)[: ops.shape[0]]
    n = ops.shape[0]
    for a in range(0, n, chunk):
        idx = np.arange(a, min(n, a + chunk), dtype=np.uint64)
        want, _ = replay.answer(idx)
        bad = np.nonzero(res[a:a + idx.size] != want)[0]
        assert bad.size == 0, (a + bad[:8], ops[a + bad[:8]], res[a + bad[:8]], want[bad[:8]])
    rq = np.nonzero(ops[:, 0] == wl.OP_RANGE)[0]
    rows = pairs_dev.view(-1, L)
    for a in range(0, rq.size, 1 << 17):
        sel = rq[a:a + (1 << 17)]
        _, want = replay.answer(sel.astype(np.uint64), L)
        slots = torch.from_numpy(ops[sel, 3].astype(np.int64)).to(rows.device)
        got = rows.index_select(0, slots).cpu().numpy().view(np.uint64).reshape(-1)
        assert np.array_equal(got, want), f"range pairs, ranges {a}.."


def test_config4_full_size(dev):
    t0 = time.time()
    w = wl.config4(seed=4)  # 40M keys, 4096 batches: the benchmark's workload
    L = w.range_capacity
    n_ops = sum(b.num_ops for b in w.batches)
    note(f"config4: {n_ops} ops generated in {time.time() - t0:.1f}s")
    nw = 32
    # spare slots: a fifth of the operation count (the call reclaims between
    # its windows, so it never holds more than about one window of copies)
    tree = mv.GpuVersionedBTree(pool_nodes=int(40_000_000 / 4 * (16 / nw) + n_ops / 5), node_width=nw)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    tree.set_worker_mode(2)
    tree.set_read_tile(16)  # as measured: 16-lane read passes over 32-pair nodes
    spare = tree.memory_stats().free_nodes
    assert spare < n_ops / 2
    e0 = len(tree.memory_series())
    r = drivers.BatchedRunner(w.batches, L, dev)
    r.run(tree, use_graph=True)
    sync()
    note(f"config4: ran ({time.time() - t0:.1f}s), {spare / n_ops:.3f} spare slots per op")
    assert tree.error_flags() == 0, hex(tree.error_flags())
    ser = tree.memory_series()[e0:]
    assert len(ser) >= 8  # one reclamation per window
    assert tree.memory_stats().retired_pending == 0
    ks = np.sort(w.initial_keys)
    replay = oracle.WriteOnceReplay(ks, wl.value_of(ks), r.ops_host, r.offsets_host)
    compare_batched(r.results, r.pairs, L, r.ops_host, r.offsets_host, replay)
    note(f"config4: compared ({time.time() - t0:.1f}s)")
    nk, _ = tree.validate_tree_structure()
    ins = r.ops_host[:, 0] == wl.OP_INSERT
    er = r.ops_host[:, 0] == wl.OP_ERASE
    assert nk == ks.size + ins.sum() - er.sum()
    tree.close()


def test_config5_full_size(dev):
    """500M-key tree: the served loop's results for one QPS level against the
    write-once replay of the request stream."""
    import torch
    from mvgpubtree_amd import serving
    t0 = time.time()
    n = 500_000_000
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    want = n + 2_000_000
    u = torch.empty(0, dtype=torch.int64, device=dev)
    while u.numel() < want:
        d = torch.randint(1, 0xFFFFFFFF, (int((want - u.numel()) * 1.03) + 4096,), generator=g, device=dev,
                          dtype=torch.int64)
        u = torch.unique(torch.cat([u, d]))
        u = u[u != wl.SENTINEL_COLLISION]
    perm = torch.randperm(u.numel(), generator=g, device=dev)[:want]
    to32 = lambda x: torch.where(x >= 2**31, x - 2**32, x).to(torch.int32)  # noqa: E731
    init_sorted = torch.sort(u[perm[:n]]).values  # the tree's keys, sorted
    fresh = to32(u[perm[n:]])
    del u, perm
    ks_dev = to32(init_sorted)
    del init_sorted
    ks = ks_dev.cpu().numpy().view(np.uint32)
    note(f"config5: {n} keys generated ({time.time() - t0:.1f}s)")
    tree = mv.GpuVersionedBTree(pool_nodes=int(n / 12) + (8 << 20), node_width=32)
    tree.set_worker_mode(3)  # by step size: small steps on the persistent worker, large ones sorted
    tree.set_read_tile(16)  # as measured: 32-pair nodes, 16-lane read passes
    order = torch.randperm(n, generator=g, device=dev)
    k_ins = ks_dev[order]
    del order
    vx = torch.tensor(wl.VALUE_XOR - 2**32, dtype=torch.int32, device=dev)
    tree.insert(k_ins, k_ins ^ vx)
    sync()
    del k_ins, ks_dev
    note(f"config5: built ({time.time() - t0:.1f}s)")
    rng = np.random.default_rng(55)
    sample = ks[rng.integers(0, n, 4_000_000)]
    sample.sort()
    fresh_h = fresh.cpu().numpy().view(np.uint32)
    reqs = serving.make_requests(rng, 2000, 16000.0, sample, fresh_h, range_capacity=128, n_live=n)
    srv = serving.ContinuousBatchingServer(tree, reqs, dev)
    summary, _ = srv.run()
    sync()
    note(f"config5: served {summary['ops']} ops in {summary['steps']} steps ({time.time() - t0:.1f}s)")
    assert tree.error_flags() == 0, hex(tree.error_flags())
    vs = wl.value_of(ks)
    replay = oracle.WriteOnceReplay(ks, vs, srv.runner.ops_host, srv.runner.offsets_host)
    compare_batched(srv.runner.results, srv.runner.pairs, 128, srv.runner.ops_host, srv.runner.offsets_host, replay)
    note(f"config5: compared ({time.time() - t0:.1f}s)")
    nk, _ = tree.validate_tree_structure()
    assert nk == n + int((srv.runner.ops_host[:, 0] == wl.OP_INSERT).sum())
    tree.close()
