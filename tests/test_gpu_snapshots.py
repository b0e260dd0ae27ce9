"""GPU tests of stream-ordered snapshots (mvbt_take_snapshot returns the id
from the host's copy of the clock and does not wait for the GPU).

Checked against the oracle: the ids follow the device clock across every
call that advances it (snapshots, mixed calls, batched launches; the device
flags MVBT_FLAG_SNAPSHOT_ID otherwise), writes issued on another stream right
after a snapshot stay invisible at it, reclamation on another stream keeps
the versions a live snapshot needs, and after a device view is handed out
the ids are read back from the device and stay in sequence.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import mvgpubtree_amd as mv
import oracle
from mvgpubtree_amd import workloads as wl

from conftest import to_dev, to_np

pytestmark = pytest.mark.gpu


def test_ids_follow_clock_across_clock_advancing_calls(dev):
    import torch
    rng = np.random.default_rng(3)
    keys = wl.unique_keys(rng, 30_000)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 16)
    ref = oracle.MVMap()
    tree.insert(to_dev(keys[:10_000], dev), to_dev(wl.value_of(keys[:10_000]), dev))
    ref.insert(keys[:10_000], wl.value_of(keys[:10_000]))
    ids = [tree.take_snapshot() for _ in range(3)]
    assert ids == [ref.take_snapshot() for _ in range(3)]
    # a mixed call reserves one id
    lo = np.sort(rng.integers(1, 1 << 31, 100, dtype=np.uint64)).astype(np.uint32)
    hi = (lo.astype(np.uint64) + (1 << 20)).astype(np.uint32)
    new = keys[10_000:12_000]
    tree.concurrent_insert_range(to_dev(new, dev), to_dev(wl.value_of(new), dev), to_dev(lo, dev),
                                 to_dev(hi, dev), 64)
    ref.take_snapshot()
    ref.insert(new, wl.value_of(new))
    s = tree.take_snapshot()
    assert s == ref.take_snapshot()
    # a batched launch of three batches reserves three ids
    bk = keys[12_000:12_009]
    ops = np.zeros((9, 4), dtype=np.uint32)  # {type, key, arg, slot}
    ops[:, 0] = mv.OP_INSERT
    ops[:, 1] = bk
    ops[:, 2] = wl.value_of(bk)
    offsets = np.array([0, 3, 6, 9], dtype=np.uint64)
    base = torch.zeros(1, dtype=torch.int32, device=dev)
    tree.run_batches(to_dev(ops, dev), to_dev(offsets, dev), snapshot_base=base)
    for b in range(3):
        ref.take_snapshot()
        ref.insert(bk[3 * b:3 * b + 3], wl.value_of(bk[3 * b:3 * b + 3]))
    s2 = tree.take_snapshot()
    assert s2 == ref.take_snapshot()
    assert int(to_np(base)[0]) == s + 1
    torch.cuda.synchronize()
    for ts in ids + [s, s2]:
        q = keys[:13_000]
        assert np.array_equal(to_np(tree.find(to_dev(q, dev), timestamp=ts)), ref.find(q, ts)), ts
    assert tree.error_flags() == 0
    tree.close()


def test_write_on_other_stream_after_snapshot_is_invisible(dev):
    import torch
    rng = np.random.default_rng(4)
    keys = wl.unique_keys(rng, 200_000)
    base, late = keys[:100_000], keys[100_000:]
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 18)
    tree.insert(to_dev(base, dev), to_dev(wl.value_of(base), dev))
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    dl, dv = to_dev(late, dev), to_dev(wl.value_of(late), dev)
    q = to_dev(np.concatenate([base[:20_000], late[:20_000]]), dev)
    torch.cuda.synchronize()
    for rnd in range(3):
        # a long write on stream a, then a snapshot on a and a write on b issued at once
        s = tree.take_snapshot(stream=a)
        part = slice(rnd * 30_000, (rnd + 1) * 30_000)
        tree.insert(dl[part], dv[part], stream=b)
        got = torch.empty_like(q)
        tree.find(q, got, timestamp=s, stream=b)
        torch.cuda.synchronize()
        r = to_np(got)
        want_base = wl.value_of(base[:20_000])
        assert np.array_equal(r[:20_000], want_base)
        # keys written in this pass are absent at s; those of earlier passes are present
        lk = late[:20_000]
        present = np.isin(lk, late[:rnd * 30_000])
        want = np.where(present, wl.value_of(lk), np.uint32(mv.NOT_FOUND))
        assert np.array_equal(r[20_000:], want), rnd
        # reclaim on b while s is live keeps what s reads
        tree.reclaim(stream=b)
        tree.find(q, got, timestamp=s, stream=b)
        torch.cuda.synchronize()
        assert np.array_equal(to_np(got)[20_000:], want), rnd
        tree.release_snapshot(s)
    tree.reclaim()
    assert tree.error_flags() == 0
    tree.close()


def test_ids_stay_in_sequence_after_device_view(dev):
    rng = np.random.default_rng(5)
    keys = wl.unique_keys(rng, 5000)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 14)
    ref = oracle.MVMap()
    tree.insert(to_dev(keys, dev), to_dev(wl.value_of(keys), dev))
    ref.insert(keys, wl.value_of(keys))
    before = [tree.take_snapshot() for _ in range(2)]
    assert before == [ref.take_snapshot() for _ in range(2)]
    handle = ctypes.create_string_buffer(512)
    assert mv.load_library().mvbt_device_view(tree._h, ctypes.cast(handle, ctypes.c_void_p)) == 0
    after = [tree.take_snapshot() for _ in range(2)]  # read back from the device now
    assert after == [ref.take_snapshot() for _ in range(2)]
    tree.erase(to_dev(keys[:1000], dev))
    ref.erase(keys[:1000])
    for ts in before + after + [mv.LATEST]:
        assert np.array_equal(to_np(tree.find(to_dev(keys, dev), timestamp=ts)), ref.find(keys, ts))
    assert tree.error_flags() == 0
    tree.close()


def test_memory_series_records_every_reclamation(dev):
    """Each reclamation appends one epoch to the device timeline: consecutive
    epoch numbers, the clock at the time, live slots equal to memory_stats,
    allocated = live + reclaimed, and the freed counts summing to the total."""
    import torch
    rng = np.random.default_rng(17)
    keys = wl.unique_keys(rng, 60_000)
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 16)
    tree.insert(to_dev(keys[:40_000], dev), to_dev(wl.value_of(keys[:40_000]), dev))
    rows = []
    for i in range(5):
        s = tree.take_snapshot()
        lo, hi = 40_000 + 4000 * i, 40_000 + 4000 * (i + 1)
        tree.insert(to_dev(keys[lo:hi], dev), to_dev(wl.value_of(keys[lo:hi]), dev))
        tree.erase(to_dev(keys[i * 3000:(i + 1) * 3000], dev))
        tree.release_snapshot(s)
        tree.reclaim()
        torch.cuda.synchronize()
        st = tree.memory_stats()
        s2 = tree.take_snapshot()  # the clock after the reclamation
        tree.release_snapshot(s2)
        rows.append((st.live_nodes, st.reclaimed_nodes, s2))
    ser = tree.memory_series()
    assert len(ser) == 5
    assert list(ser.epoch) == list(range(5))
    for r, (live, reclaimed, clock_after) in zip(ser, rows):
        assert r.live_nodes == live
        assert r.reclaimed_total == reclaimed
        assert r.allocated_total == r.live_nodes + r.reclaimed_total
        assert r.retired_pending == 0
        assert r.clock == clock_after  # no snapshot between the reclamation and s2
    assert int(ser.freed_nodes.sum()) == rows[-1][1]
    assert int(ser.freed_nodes.min()) > 0  # every batch made version copies
    assert len(tree.memory_series(capacity=2)) == 2
    assert tree.error_flags() == 0
    tree.close()
