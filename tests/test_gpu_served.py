"""GPU tests of the served form (continuous batching): every request's
results equal those of the requests run one at a time, in arrival order,
through the plain per-request path (checked word-for-word against the
oracle), whatever the step boundaries and overlap; admission by free pool
memory keeps the pool from running out."""
from __future__ import annotations

import numpy as np
import pytest

import mvgpubtree_amd as mv
import oracle
from mvgpubtree_amd import serving
from mvgpubtree_amd import workloads as wl

from conftest import to_dev, to_np

pytestmark = pytest.mark.gpu


def setup(dev, seed, n_initial=100_000, n_requests=300, qps=1e6, max_ops=2048, pool=1 << 18, nw=16, read_tile=0):
    rng = np.random.default_rng(seed)
    keys = wl.unique_keys(rng, n_initial + 400_000)
    init, fresh = keys[:n_initial], keys[n_initial:]
    reqs = serving.make_requests(rng, n_requests, qps, np.sort(init), fresh, max_ops=max_ops,
                                 range_capacity=128)
    tree = mv.GpuVersionedBTree(pool_nodes=pool, node_width=nw)
    tree.set_read_tile(read_tile)
    tree.insert(to_dev(init, dev), to_dev(wl.value_of(init), dev))
    return init, reqs, tree


@pytest.mark.parametrize("max_inflight,use_graph", [(1, False), (2, True), (3, True)])
def test_served_results_equal_sequential_requests(dev, max_inflight, use_graph):
    init, reqs, tree = setup(dev, seed=max_inflight)
    srv = serving.ContinuousBatchingServer(tree, reqs, dev, max_step_ops=4096,
                                           max_inflight=max_inflight, use_graph=use_graph)
    summary, lat = srv.run()
    r = srv.runner
    want, wpairs, final = oracle.batch_expected(init, wl.value_of(init), r.ops_host, r.offsets_host,
                                                r.L, r.slots)
    assert np.array_equal(to_np(r.results)[: r.nops], want)
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * r.L], wpairs)
    assert summary["steps"] > 1 and np.all(lat >= 0)
    assert tree.validate_tree_structure()[0] == len(final)
    assert tree.error_flags() == 0
    tree.close()


def test_served_steps_of_both_forms_on_a_wide_tree(dev):
    """The served configuration: 32-pair nodes read by 16-lane tiles, worker
    mode by step size. Bursts of requests make steps both below and above
    the size at which the sorted form takes over; every result equals the
    requests run one at a time."""
    init, reqs, tree = setup(dev, seed=9, n_requests=400, qps=2e7, max_ops=8192, pool=1 << 19, nw=32,
                             read_tile=16)
    srv = serving.ContinuousBatchingServer(tree, reqs, dev, max_step_ops=1 << 16, max_inflight=2)
    summary, lat = srv.run()
    r = srv.runner
    assert min(srv.step_ops) < 16384 <= max(srv.step_ops)  # both forms ran
    want, wpairs, final = oracle.batch_expected(init, wl.value_of(init), r.ops_host, r.offsets_host,
                                                r.L, r.slots)
    assert np.array_equal(to_np(r.results)[: r.nops], want)
    assert np.array_equal(to_np(r.pairs, np.uint64)[: r.slots * r.L], wpairs)
    assert summary["steps"] > 1 and np.all(lat >= 0)
    assert tree.validate_tree_structure()[0] == len(final)
    assert tree.error_flags() == 0
    tree.close()


def test_per_request_path_matches_oracle(dev):
    init, reqs, tree = setup(dev, seed=5, n_requests=80)
    srv = serving.PerRequestServer(tree, reqs, dev, keep_pairs=True)
    summary, lat = srv.run()
    assert summary["requests"] == len(reqs.batches) and np.all(lat >= 0)
    got = [srv.runner.batch_results(i) for i in range(len(reqs.batches))]
    st = oracle.SortedState(init, wl.value_of(init))
    for (f, c, p), b in zip(got, reqs.batches):
        assert np.array_equal(f, st.find(b.find_keys))
        wp, wc = st.range_query(b.range_lo, b.range_hi, reqs.range_capacity)
        assert np.array_equal(c, wc) and np.array_equal(p, wp)
        st.insert(b.insert_keys, b.insert_values)
    tree.close()


def test_admission_waits_for_free_memory(dev):
    """A pool with little headroom: the server must split the stream into many
    steps and reclaim between them; no write may fail."""
    rng = np.random.default_rng(9)
    init = wl.unique_keys(rng, 50_000)
    fresh = wl.unique_keys(rng, 100_000, exclude_sorted=np.sort(init))
    reqs = serving.make_requests(rng, 200, 1e6, np.sort(init), fresh, max_ops=1024,
                                 range_capacity=64, mix=(0.5, 0.1, 0.4))
    tree = mv.GpuVersionedBTree(pool_nodes=24_000)
    tree.insert(to_dev(init, dev), to_dev(wl.value_of(init), dev))
    srv = serving.ContinuousBatchingServer(tree, reqs, dev, max_step_ops=1 << 20, max_inflight=2)
    summary, _ = srv.run()
    r = srv.runner
    want, _, _ = oracle.batch_expected(init, wl.value_of(init), r.ops_host, r.offsets_host, r.L, r.slots)
    assert np.array_equal(to_np(r.results)[: r.nops], want)
    assert tree.error_flags() == 0
    assert summary["steps"] >= 3
    tree.close()
