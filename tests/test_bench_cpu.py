"""CPU checks of bench.py's contract pieces that need no GPU: the refusal of
more than one GPU, the algorithmic byte counts per operation, the measured
traffic lookup, and the sharded CPU baseline, whose key-range clipping must
return the range counts of a single map (including ranges that straddle
shard splitters and ranges that run to the top of the key space)."""
from __future__ import annotations

import importlib.util
import os
import subprocess
import sys

import numpy as np

import oracle
from mvgpubtree_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_follow_the_reference_layout():
    b = load_bench()
    # h = 8 levels of 128-byte nodes at 40M keys, plus the operation's own bytes
    assert b.BYTES == {"find": 1032, "insert": 1288, "range32": 1932, "erase": 1284}


def test_traffic_is_reported_only_for_the_node_width_it_was_measured_at():
    b = load_bench()
    import json
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        data = json.load(f)
    nw = data["node_width"]
    assert b.load_traffic("insert", nw) == data["insert"]
    assert b.load_traffic("insert", 64 if nw != 64 else 16) is None


def test_sharded_cpu_baseline_matches_one_map():
    b = load_bench()
    w = wl.config2(seed=12, n_initial=300_000, num_batches=1, ops_per_batch=40_000)
    batch = w.batches[0]
    # ranges at both ends of the key space and one across the median key (a shard
    # boundary at 2 and 8 threads); the random ranges cross the others
    ks = np.sort(w.initial_keys)
    lo, hi = batch.range_lo, batch.range_hi  # edited in place: the baseline reads the batch
    lo[:3] = [0xFFFFF000, 1, ks[ks.size // 2] - 5]
    hi[:3] = [0xFFFFFFFF, 100, ks[ks.size // 2] + 5]
    m = oracle.MVMap()
    m.build_sorted(ks, wl.value_of(ks))
    s = m.take_snapshot()
    want = int(m.range_query(lo, hi, w.range_capacity, s)[1].sum())
    m.close()
    for threads in (1, 3, 8):
        _, _, got = b.cpu_baseline_config2_sharded(w, batch, threads)
        assert got == want, threads
