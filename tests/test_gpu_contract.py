"""GPU tests of the API contract around concurrency and the batched form.

* latest-version reads with MVBT_CONCURRENT on a versioned tree beside writers
  on another stream: keys no writer touches read exactly, touched keys read
  as their old or their new state, range results stay sorted and in bounds;
* mvbt_run_batches rejects a call in which a batch writes one key twice (or
  names an unknown op type, or whose offsets disagree with num_ops): nothing
  is applied and MVBT_FLAG_BAD_BATCH is raised;
* a batched call is ordered against writes on other streams in host call
  order (it waits for earlier ones, later ones wait for it);
* the retire ring is raised to the pool size, so copies kept alive by a long
  snapshot never overflow it;
* a launch over a slice of a larger op array with default results.
"""
from __future__ import annotations

import numpy as np
import pytest

import mvgpubtree_amd as mv
import oracle
from mvgpubtree_amd import drivers
from mvgpubtree_amd import workloads as wl

from conftest import to_dev, to_np

pytestmark = pytest.mark.gpu
MISS = 0xFFFFFFFF


def sync():
    import torch
    torch.cuda.synchronize()


@pytest.mark.parametrize("nw", [16, 32])
def test_versioned_latest_concurrent_reads_beside_writers(dev, nw):
    """Latest reads under MVBT_CONCURRENT read each leaf under its lock:
    in-epoch rewrites shift pairs in place, and a torn read would miss an
    untouched key or repeat one in a range."""
    import torch
    rng = np.random.default_rng(31 + nw)
    keys = wl.unique_keys(rng, 500_000)
    base, fresh = keys[:300_000], keys[300_000:]
    vals = wl.value_of(keys)
    bv = vals[:300_000]
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 19, node_width=nw)
    tree.insert(to_dev(base, dev), to_dev(bv, dev))
    s = tree.take_snapshot()  # later writes copy each leaf once, then rewrite it in place
    tree.release_snapshot(s)
    erased, updated, untouched = base[0::3], base[1::3], base[2::3]
    q = np.concatenate([base, fresh, wl.absent_keys(rng, 50_000, np.sort(keys))])
    rng.shuffle(q)
    span = wl.span_for(24, 300_000)
    lo = rng.integers(1, 0xFFFFFF00, size=8_000, dtype=np.uint64).astype(np.uint32)
    hi = np.minimum(lo.astype(np.uint64) + span, 0xFFFFFFFF).astype(np.uint32)
    L = 256
    dq, dlo, dhi = to_dev(q, dev), to_dev(lo, dev), to_dev(hi, dev)
    parts = 8
    dfresh = [to_dev(x, dev) for x in np.array_split(fresh, parts)]
    dfreshv = [to_dev(wl.value_of(x), dev) for x in np.array_split(fresh, parts)]
    der = [to_dev(x, dev) for x in np.array_split(erased, parts)]
    dup = [to_dev(x, dev) for x in np.array_split(updated, parts)]
    dupv = [to_dev(wl.value_of(x) ^ np.uint32(1), dev) for x in np.array_split(updated, parts)]
    sync()
    sr, sw = torch.cuda.Stream(), torch.cuda.Stream()
    finds, ranges = [], []
    for i in range(parts):  # writes and reads interleave on two streams
        with torch.cuda.stream(sw):
            tree.insert(dfresh[i], dfreshv[i], stream=sw)
            tree.erase(der[i], stream=sw)
            tree.insert(dup[i], dupv[i], stream=sw)
        with torch.cuda.stream(sr):
            finds.append(tree.find(dq, stream=sr, concurrent=True))
            ranges.append(tree.range_query(dlo, dhi, L, stream=sr, concurrent=True))
    sync()
    v_of = dict(zip(keys.tolist(), vals.tolist()))
    qs = q
    is_untouched = np.isin(qs, untouched)
    is_erased = np.isin(qs, erased)
    is_updated = np.isin(qs, updated)
    is_fresh = np.isin(qs, fresh)
    want = wl.value_of(qs)
    for f in finds:
        got = to_np(f)
        assert np.array_equal(got[is_untouched], want[is_untouched])
        e = got[is_erased]
        assert np.all((e == want[is_erased]) | (e == MISS))
        u = got[is_updated]
        assert np.all((u == want[is_updated]) | (u == (want[is_updated] ^ np.uint32(1))))
        fr = got[is_fresh]
        assert np.all((fr == want[is_fresh]) | (fr == MISS))
        other = ~(is_untouched | is_erased | is_updated | is_fresh)
        assert np.all(got[other] == MISS)
    su = np.sort(untouched)
    a0 = np.searchsorted(su, lo)
    a1 = np.searchsorted(su, hi)
    for p, n in ranges[::3]:
        pairs = to_np(p, np.uint64).reshape(-1, L)
        cnt = to_np(n)
        assert cnt.max() <= L
        for i in range(lo.size):
            row = pairs[i, : cnt[i]]
            k = (row & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            v = (row >> np.uint64(32)).astype(np.uint32)
            assert np.all(np.diff(k.astype(np.int64)) > 0), i
            assert np.all((k >= lo[i]) & (k < hi[i])), i
            # every untouched key of the range is there, with its value
            need = su[a0[i]:a1[i]]
            assert np.isin(need, k).all(), i
            for kk, vv in zip(k.tolist(), v.tolist()):
                w0 = v_of[kk]
                assert vv == w0 or vv == (w0 ^ 1), (i, kk)
    assert tree.error_flags() == 0, hex(tree.error_flags())
    ref = oracle.SortedState(base, bv)
    ref.insert(fresh, wl.value_of(fresh))
    ref.erase(erased)
    ref.insert(updated, wl.value_of(updated) ^ np.uint32(1))
    assert np.array_equal(to_np(tree.find(dq)), ref.find(q))
    assert tree.validate_tree_structure()[0] == len(ref)
    tree.close()


def _stream_with(w, extra_batch_ops, at=3):
    ops, off, slots = wl.to_ops(w.batches, seed=1)
    ops = ops.copy()
    b0, b1 = int(off[at]), int(off[at + 1])
    blk = np.concatenate([ops[b0:b1], np.asarray(extra_batch_ops, dtype=np.uint32).reshape(-1, 4)])
    ops = np.concatenate([ops[:b0], blk, ops[b1:]])
    off = off.copy()
    off[at + 1:] += len(extra_batch_ops)
    return ops, off, slots


@pytest.mark.parametrize("case", ["insert_insert", "erase_insert", "erase_erase", "unknown_type", "offsets"])
def test_run_batches_rejects_bad_calls(dev, case):
    w = wl.config4(seed=61, n_initial=50_000, num_batches=12, min_batch=32, max_batch=2048, range_capacity=256)
    base_ops, base_off, slots = wl.to_ops(w.batches, seed=1)
    b3 = base_ops[int(base_off[3]):int(base_off[4])]
    ek = int(b3[b3[:, 0] == wl.OP_ERASE][0, 1])
    ik = int(b3[b3[:, 0] == wl.OP_INSERT][0, 1])
    extra = {"insert_insert": [(wl.OP_INSERT, ik, 5, 0)],
             "erase_insert": [(wl.OP_INSERT, ek, 5, 0)],
             "erase_erase": [(wl.OP_ERASE, ek, 0, 0)],
             "unknown_type": [(9, 12345, 0, 0)],
             "offsets": []}[case]
    ops, off, slots = _stream_with(w, extra) if extra else (base_ops, base_off, slots)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 17)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    s0 = tree.take_snapshot()
    tree.release_snapshot(s0)
    before = tree.export()
    import torch
    dops = to_dev(ops, dev).reshape(-1, 4)
    doff = to_dev(off, dev)
    res = torch.full((ops.shape[0],), 0x5A5A5A5A, dtype=torch.int32, device=dev)
    pairs = torch.full((max(slots, 1) * 256,), -1, dtype=torch.int64, device=dev)
    n = ops.shape[0] if case != "offsets" else ops.shape[0] - 1
    tree.run_batches(dops, doff, results=res, pairs=pairs, pair_slots=max(slots, 1),
                     average_range_length=256, num_ops=n, use_graph=(case == "erase_erase"))
    sync()
    assert tree.error_flags() == mv.FLAG_BAD_BATCH
    assert np.all(to_np(res) == 0x5A5A5A5A)  # nothing ran
    assert np.array_equal(tree.export(), before)
    nb = off.size - 1
    assert tree.take_snapshot() == s0 + 1 + nb  # the ids were reserved all the same
    tree.clear_error_flags()
    # the same handle then runs a valid call correctly
    r = drivers.BatchedRunner(w.batches, 256, dev, seed=2)
    got = to_np(r.run(tree, use_graph=True))[: r.nops]
    want, wpairs, _ = oracle.replay_batches(w.initial_keys, w.initial_values, r.ops_host, r.offsets_host, 256,
                                            r.slots)
    assert np.array_equal(got, want)
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * 256], wpairs)
    assert tree.error_flags() == 0
    tree.close()


def test_batched_call_is_ordered_against_other_streams(dev):
    """insert X on stream B, run_batches on stream A, insert Y on stream B:
    the batches see X and not Y, whatever the device interleaving."""
    import torch
    # ranges sized for the initial keys read a tree that X doubles: capacity 512
    w = wl.config4(seed=62, n_initial=200_000, num_batches=64, min_batch=32, max_batch=8192, range_capacity=512)
    rng = np.random.default_rng(5)
    live = np.sort(np.concatenate([w.initial_keys] + [b.insert_keys for b in w.batches]))
    xy = wl.unique_keys(rng, 400_000, exclude_sorted=live)
    x, y = xy[:200_000], xy[200_000:]
    # make the batches read X and Y: finds of both in every batch
    for b in w.batches:
        extra = np.concatenate([x[rng.integers(0, x.size, 16)], y[rng.integers(0, y.size, 16)]])
        b.find_keys = np.concatenate([b.find_keys, extra])
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 19)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    r = drivers.BatchedRunner(w.batches, 512, dev, seed=3)
    dx, dxv = to_dev(x, dev), to_dev(wl.value_of(x), dev)
    dy, dyv = to_dev(y, dev), to_dev(wl.value_of(y), dev)
    sync()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    tree.insert(dx, dxv, stream=sb)
    r.run(tree, stream=sa, use_graph=True)
    tree.insert(dy, dyv, stream=sb)
    sync()
    k0 = np.concatenate([w.initial_keys, x])
    want, wpairs, m = oracle.replay_batches(k0, wl.value_of(k0), r.ops_host, r.offsets_host, 512, r.slots)
    assert np.array_equal(to_np(r.results)[: r.nops], want)
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * 512], wpairs)
    m.insert(y, wl.value_of(y))
    assert np.array_equal(tree.export(), m.export())
    assert tree.error_flags() == 0
    tree.close()


def test_retire_ring_is_raised_to_the_pool(dev):
    """A ring asked for 256 entries holds at least the pool: thousands of
    copies kept by a live snapshot are all retired and later all freed."""
    rng = np.random.default_rng(12)
    k = wl.unique_keys(rng, 120_000)
    v = wl.value_of(k)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 16, retire_capacity=256)
    tree.insert(to_dev(k[:40_000], dev), to_dev(v[:40_000], dev))
    s0 = tree.take_snapshot()
    snaps = []
    for i in range(8):
        snaps.append(tree.take_snapshot())
        part = k[40_000 + i * 10_000: 40_000 + (i + 1) * 10_000]
        tree.insert(to_dev(part, dev), to_dev(wl.value_of(part), dev))
        tree.erase(to_dev(k[i * 2000:(i + 1) * 2000], dev))
        tree.reclaim()  # s0 is live: nothing it reads may go
    st = tree.memory_stats()
    assert st.retired_pending > 4 * 256
    got = to_np(tree.find(to_dev(k, dev), timestamp=s0))
    want = np.where(np.arange(k.size) < 40_000, v, MISS).astype(np.uint32)
    assert np.array_equal(got, want)
    for s in [s0] + snaps:
        tree.release_snapshot(s)
    tree.reclaim()
    st = tree.memory_stats()
    assert st.retired_pending == 0
    assert st.live_nodes + st.free_nodes == st.capacity_nodes
    assert tree.error_flags() == 0, hex(tree.error_flags())
    # the freed slots are reused without harm
    tree.insert(to_dev(k[:16_000], dev), to_dev(v[:16_000], dev))
    ref = oracle.SortedState(k[16_000:], v[16_000:])
    ref.insert(k[:16_000], v[:16_000])
    assert np.array_equal(to_np(tree.find(to_dev(k, dev))), ref.find(k))
    assert tree.validate_tree_structure()[0] == len(ref)
    assert tree.error_flags() == 0
    tree.close()


def test_slice_of_a_larger_op_array_with_default_results(dev):
    """Run batches [10, 30) of a pre-staged stream: offsets are absolute op
    indices, the default results tensor covers every row of the op array."""
    import torch
    w = wl.config4(seed=63, n_initial=100_000, num_batches=40, min_batch=32, max_batch=4096, range_capacity=256)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 18)
    tree.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
    ops, off, slots = wl.to_ops(w.batches, seed=4)
    dops = to_dev(ops, dev).reshape(-1, 4)
    L = 256
    pairs = torch.full((slots * L,), -1, dtype=torch.int64, device=dev)
    first, last = 10, 30
    doff = to_dev(off[first:last + 1].copy(), dev)
    n = int(off[last] - off[first])
    res = tree.run_batches(dops, doff, pairs=pairs, pair_slots=slots, average_range_length=L, num_ops=n)
    sync()
    assert res.numel() == ops.shape[0]
    # expected: batches [0, 10) are not applied; the slice starts from the initial keys
    sub = ops[int(off[first]):int(off[last])]
    sub_off = off[first:last + 1] - off[first]
    want, wpairs, _ = oracle.replay_batches(w.initial_keys, w.initial_values, sub, sub_off, L, slots)
    assert np.array_equal(to_np(res)[int(off[first]):int(off[last])], want)
    rq = sub[:, 0] == wl.OP_RANGE
    sl = sub[rq, 3].astype(np.int64)
    got_pairs = to_np(pairs, np.uint64).reshape(-1, L)[sl]
    assert np.array_equal(got_pairs, wpairs.reshape(-1, L)[sl])
    assert tree.error_flags() == 0
    tree.close()
