"""Served form: continuous batching of requests over the batched launch.

A request is a small batch of operations (find / range / insert) that reads
one snapshot: the state after every earlier request. Requests arrive
open-loop (Poisson). Two servers answer the same request stream:

* ``ContinuousBatchingServer``: each step is one batched launch
  (``run_batches``) over every request admitted since the previous step,
  one batch per request, in arrival order. Up to ``max_inflight`` steps are
  queued on the stream, so requests that arrive while a step runs are
  admitted into the next one; a request's work overlaps with its neighbours'
  on the device. Admission is bounded by step slots (ops per step) and by
  free node-pool memory: each admitted request reserves pool slots for its
  writes until its step completes (a write uses at most a leaf copy and a
  split sibling, and a tile's magazine holds at most about twice what it
  used, hence 7 slots per write), and the free count is refreshed by
  stream-ordered statistics after each step's reclamation.
* ``PerRequestServer``: the plain per-request launch path - for each
  request, take a snapshot, launch find and range_query at it, then insert,
  release, one request at a time (the reference host API's pattern), from a
  C++ loop over the C ABI (drivers.PerBatchRunner.run_requests).

Both produce identical results for the same request order (a request's
reads see exactly the requests before it).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import drivers
from . import workloads as wl


@dataclass
class RequestStream:
    batches: list            # one workloads.Batch per request
    arrivals: np.ndarray     # seconds from the start, non-decreasing
    range_capacity: int

    @property
    def num_ops(self) -> int:
        return sum(b.num_ops for b in self.batches)


def make_requests(rng, n_requests: int, qps: float, live_sorted: np.ndarray, fresh_keys: np.ndarray,
                  min_ops: int = 1, max_ops: int = 16384, expected_range: int = 32,
                  range_capacity: int = 128, mix=(0.8, 0.1, 0.1),
                  n_live: int | None = None) -> RequestStream:
    """Log-uniform request sizes, mix (find, range, insert), Poisson arrivals.
    Finds are half hits among `live_sorted` (live keys, or a sample of them),
    half uniform draws; inserts take keys from `fresh_keys` (absent from the
    tree) in order; range spans hold ~expected_range keys of n_live."""
    sizes = wl.log_uniform_sizes(rng, n_requests, min_ops, max_ops)
    gaps = rng.exponential(1.0 / qps, size=n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    span = wl.span_for(expected_range, n_live or live_sorted.size)
    fi = 0
    batches = []
    for s in sizes:
        s = int(s)
        kinds = rng.choice(3, size=s, p=mix)
        n_f, n_r, n_i = int((kinds == 0).sum()), int((kinds == 1).sum()), int((kinds == 2).sum())
        fk = np.concatenate([live_sorted[rng.integers(0, live_sorted.size, size=n_f - n_f // 2)],
                             wl._draw(rng, n_f // 2)])
        rng.shuffle(fk)
        lo = wl._draw(rng, n_r)
        ik = fresh_keys[fi:fi + n_i]
        fi += n_i
        batches.append(wl.Batch(insert_keys=ik, insert_values=wl.value_of(ik), find_keys=fk,
                                range_lo=lo, range_hi=wl._hi(lo, span)))
    return RequestStream(batches, arrivals, range_capacity)


def _percentiles(lat_s: np.ndarray):
    return {"p50_ms": round(float(np.percentile(lat_s, 50)) * 1e3, 4),
            "p99_ms": round(float(np.percentile(lat_s, 99)) * 1e3, 4),
            "max_ms": round(float(lat_s.max()) * 1e3, 4)}


class ContinuousBatchingServer:
    def __init__(self, tree, reqs: RequestStream, dev, max_step_ops: int = 1 << 20,
                 max_inflight: int = 2, slots_per_write: int = 7, use_graph: bool = True,
                 stream=None):
        import torch
        self.tree, self.reqs, self.dev = tree, reqs, dev
        self.max_step_ops = max_step_ops
        self.max_inflight = max_inflight
        self.slots_per_write = slots_per_write
        self.use_graph = use_graph
        self.runner = drivers.BatchedRunner(reqs.batches, reqs.range_capacity, dev)
        self.offsets = self.runner.offsets_host.astype(np.int64)
        ty = self.runner.ops_host[:, 0]
        w = ((ty == wl.OP_INSERT) | (ty == wl.OP_ERASE)).astype(np.int64)
        cw = np.concatenate([[0], np.cumsum(w)])
        self.writes = cw[self.offsets[1:]] - cw[self.offsets[:-1]]
        # A long-lived server keeps one stream: the first launches on a new
        # stream pay one-time runtime set-up, which belongs outside the loop.
        self.stream = stream if stream is not None else torch.cuda.Stream()
        self.stats = [torch.zeros(8, dtype=torch.int64).pin_memory() for _ in range(max_inflight + 1)]

    def run(self):
        import torch
        tree, R = self.tree, len(self.reqs.batches)
        arrivals = self.reqs.arrivals
        done_at = np.zeros(R)
        st = tree.memory_stats()
        free_nodes, reserved = st.free_nodes, 0
        inflight = deque()
        nxt = k = 0
        steps = 0
        admitted_at = np.zeros(R)
        step_gpu_ms = []
        step_ops = []
        enqueue_ms = []
        # scratch and launch graph for the largest step, before the clock starts
        sizes = np.diff(self.offsets)
        tree.reserve_batches(int(min(self.offsets[-1], self.max_step_ops + (sizes.max() if R else 0))), max(R, 1))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while k < R:
            now = time.perf_counter() - t0
            while inflight and inflight[0][0].query():
                ev, a, b, res, stat, ev0 = inflight.popleft()
                t = time.perf_counter() - t0
                step_gpu_ms.append(ev0.elapsed_time(ev))
                done_at[a:b] = t
                k = b
                reserved -= res
                free_nodes = int(stat[4])
                now = t
            if nxt < R and arrivals[nxt] <= now and len(inflight) < self.max_inflight:
                # admit every arrived request that fits the step and the pool
                b = nxt
                ops = 0
                res = 0
                while b < R and arrivals[b] <= now:
                    n_b = int(self.offsets[b + 1] - self.offsets[b])
                    r_b = int(self.writes[b]) * self.slots_per_write + 64
                    if b > nxt and (ops + n_b > self.max_step_ops):
                        break
                    if reserved + res + r_b > free_nodes:
                        break
                    ops += n_b
                    res += r_b
                    b += 1
                if b == nxt and not inflight:
                    raise RuntimeError("a request needs more node-pool slots than the pool has free")
                if b > nxt:
                    stat = self.stats[steps % len(self.stats)]
                    admitted_at[nxt:b] = now
                    step_ops.append(ops)
                    te = time.perf_counter()
                    with torch.cuda.stream(self.stream):
                        ev0 = torch.cuda.Event(enable_timing=True)
                        ev0.record(self.stream)
                        tree.run_batches(self.runner.ops, self.runner.offsets[nxt:b + 1],
                                         results=self.runner.results, pairs=self.runner.pairs,
                                         pair_slots=max(self.runner.slots, 1),
                                         average_range_length=self.runner.L, stream=self.stream,
                                         use_graph=self.use_graph, num_ops=ops)
                        tree.reclaim(self.stream)
                        tree.memory_stats_async(stat, self.stream)
                        ev = torch.cuda.Event(enable_timing=True)
                        ev.record(self.stream)
                    enqueue_ms.append((time.perf_counter() - te) * 1e3)
                    inflight.append((ev, nxt, b, res, stat, ev0))
                    reserved += res
                    nxt = b
                    steps += 1
                    continue
            if not inflight and nxt < R and arrivals[nxt] > now:
                time.sleep(min(arrivals[nxt] - now, 1e-3) if arrivals[nxt] - now > 2e-4 else 0)
        total = time.perf_counter() - t0
        self.step_ops = step_ops  # operations of each step, in order
        lat = done_at - arrivals
        gpu = np.asarray(step_gpu_ms) if step_gpu_ms else np.zeros(1)
        return {"requests": R, "ops": self.reqs.num_ops, "steps": steps,
                "Mops": round(self.reqs.num_ops / total / 1e6, 3), "wall_s": round(total, 4),
                **_percentiles(lat),
                # where the latency goes: waiting for admission, then the step
                "queue_p99_ms": round(float(np.percentile(admitted_at - arrivals, 99)) * 1e3, 4),
                "step_gpu_p50_ms": round(float(np.percentile(gpu, 50)), 4),
                "step_gpu_p99_ms": round(float(np.percentile(gpu, 99)), 4),
                "step_ops_mean": round(float(np.mean(step_ops)) if step_ops else 0.0, 1),
                "enqueue_p50_ms": round(float(np.percentile(enqueue_ms, 50)), 4) if enqueue_ms else 0.0,
                "enqueue_max_ms": round(float(np.max(enqueue_ms)), 4) if enqueue_ms else 0.0,
                "worst": [(int(i), round(float(lat[i]) * 1e3, 3),
                           round(float(admitted_at[i] - arrivals[i]) * 1e3, 3))
                          for i in np.argsort(lat)[-4:]]}, lat


class PerRequestServer:
    """Plain per-request launch path: each request is issued at its arrival
    (or when the previous one completes) and waited for, from C++."""

    def __init__(self, tree, reqs: RequestStream, dev, reclaim_every: int = 64, keep_pairs: bool = False):
        self.tree, self.reqs = tree, reqs
        self.runner = drivers.PerBatchRunner(reqs.batches, reqs.range_capacity, dev,
                                             reclaim_every=reclaim_every, keep_pairs=keep_pairs)

    def run(self):
        import torch
        R = len(self.reqs.batches)
        arrivals = self.reqs.arrivals
        stream = torch.cuda.current_stream()
        torch.cuda.synchronize()
        done_at = self.runner.run_requests(self.tree, arrivals, stream)
        total = float(done_at[-1]) if R else 0.0
        lat = done_at - arrivals
        return {"requests": R, "ops": self.reqs.num_ops, "steps": R,
                "Mops": round(self.reqs.num_ops / total / 1e6, 3) if total else 0.0, "wall_s": round(total, 4),
                **_percentiles(lat)}, lat
