"""The allocation-free batched entry point (mvbt_run_batches_noalloc) against
mvbt_run_batches: two calls in a row on two trees built alike, the second
call smaller than the first, every result word, range pair and snapshot base
equal (and equal to the oracle); a call beyond the limits is refused before
anything runs."""
from __future__ import annotations

import numpy as np
import pytest

import mvgpubtree_amd as mv
import oracle
from mvgpubtree_amd import drivers
from mvgpubtree_amd import workloads as wl

from conftest import to_dev, to_np

pytestmark = pytest.mark.gpu


def test_noalloc_matches_run_batches(dev):
    w = wl.config4(seed=500, n_initial=200_000, num_batches=160, min_batch=32, max_batch=8192, range_capacity=256)
    L = w.range_capacity
    calls = [w.batches[:120], w.batches[120:]]
    trees = []
    for _ in range(2):
        t = mv.GpuVersionedBTree(pool_nodes=1 << 19, node_width=32)
        t.insert(to_dev(w.initial_keys, dev), to_dev(w.initial_values, dev))
        trees.append(t)
    ref, na = trees
    ks, vs = np.sort(w.initial_keys), wl.value_of(np.sort(w.initial_keys))
    sizes = []
    for i, batches in enumerate(calls):
        r1 = drivers.BatchedRunner(batches, L, dev, seed=i)
        r2 = drivers.BatchedRunner(batches, L, dev, seed=i)
        sizes.append(r1.nops)
        assert r1.nops <= mv.NOALLOC_MAX_OPS and len(batches) <= mv.NOALLOC_MAX_BATCHES
        r1.run(ref, use_graph=True)
        na.run_batches(r2.ops, r2.offsets, results=r2.results, pairs=r2.pairs, pair_slots=max(r2.slots, 1),
                       average_range_length=L, snapshot_base=r2.base, no_alloc=True)
        res1, res2 = to_np(r1.results)[: r1.nops], to_np(r2.results)[: r2.nops]
        assert np.array_equal(res1, res2), f"call {i}: results"
        assert np.array_equal(to_np(r1.pairs, np.uint64), to_np(r2.pairs, np.uint64)), f"call {i}: pairs"
        assert int(to_np(r1.base)[0]) == int(to_np(r2.base)[0])
        want, wpairs, final = oracle.batch_expected(ks, vs, r1.ops_host, r1.offsets_host, L, r1.slots)
        assert np.array_equal(res2, want), f"call {i}: oracle"
        assert np.array_equal(to_np(r2.pairs, np.uint64)[: r2.slots * L], wpairs)
        ks, vs = final.keys, final.values
    assert sizes[1] < sizes[0]
    assert np.array_equal(ref.export(), na.export())
    assert ref.error_flags() == 0 and na.error_flags() == 0
    for t in trees:
        t.close()


def test_noalloc_refuses_calls_beyond_its_limits(dev):
    import torch
    tree = mv.GpuVersionedBTree(pool_nodes=1 << 16, node_width=16)
    n = mv.NOALLOC_MAX_OPS + 1
    ops = torch.zeros((n, 4), dtype=torch.int32, device=dev)
    ops[:, 1] = 7  # finds of key 7
    offsets = torch.tensor([0, n], dtype=torch.int64, device=dev)
    s0 = tree.take_snapshot()
    with pytest.raises(mv.MvbtError, match="at most"):
        tree.run_batches(ops, offsets, no_alloc=True)
    many = torch.arange(0, mv.NOALLOC_MAX_BATCHES + 2, dtype=torch.int64, device=dev)
    with pytest.raises(mv.MvbtError, match="at most"):
        tree.run_batches(ops[: mv.NOALLOC_MAX_BATCHES + 1], many, no_alloc=True)
    assert tree.take_snapshot() == s0 + 1  # nothing ran: the clock did not move
    assert tree.error_flags() == 0
    tree.close()
