"""Batch drivers over the tree API (used by the benchmark and the tests).

* ``PerBatchRunner`` - the plain per-batch launch path: for every batch take a
  snapshot, launch find and range_query at it, then insert and erase, then
  release the snapshot (one launch per operation type, as with the
  reference's host API). The loop runs in C++ over the C ABI
  (``libmvbt_driver.so``, csrc/per_batch_driver.cpp), so the comparison with
  the batched form measures device paths, not an interpreter's loop.
* ``BatchedRunner`` - the persistent work-queue form: the whole stream of
  batches in one ``run_batches`` launch, optionally through the captured
  hipGraph.

Both produce the same observable results for the same batches.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import workloads as wl

DRIVER_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmvbt_driver.so")
_drv = None


def driver():
    """libmvbt_driver.so (the plain caller loops); raises if it is not built."""
    global _drv
    if _drv is None:
        from . import load_library
        load_library()  # libmvbt.so first (libmvbt_driver.so links against it)
        if not os.path.exists(DRIVER_PATH):
            raise ImportError(f"{DRIVER_PATH} is not built; run __graft_entry__.build()")
        D = ctypes.CDLL(DRIVER_PATH)
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        D.mvbt_driver_per_batch.argtypes = [vp, u32, vp, vp, vp, vp, vp, vp, vp, vp, u32, vp, vp, vp, vp, vp,
                                            u32, u32, vp]
        D.mvbt_driver_per_batch.restype = ctypes.c_int
        D.mvbt_driver_per_request.argtypes = [vp, u32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, u32, vp, vp, vp,
                                              u32, u32, vp]
        D.mvbt_driver_per_request.restype = ctypes.c_int
        _drv = D
    return _drv


def _dev(a, dev):
    import torch
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view({4: np.int32, 8: np.int64}[a.dtype.itemsize])).to(dev)


def _cat(arrs):
    return np.ascontiguousarray(np.concatenate(arrs).astype(np.uint32)) if arrs else np.zeros(0, np.uint32)


def _offsets(arrs):
    return np.concatenate([[0], np.cumsum([a.size for a in arrs])]).astype(np.uint64)


class PerBatchRunner:
    """Device-resident inputs for the per-batch path, grouped by operation
    type with per-batch host offsets; run() issues the launches from C++.
    Find results and range counts are kept for every batch; range pairs for
    every batch with keep_pairs=True, else for the last batch only."""

    def __init__(self, batches, L: int, dev, reclaim_every: int = 64, keep_pairs: bool = False):
        import torch
        self.L = L
        self.nb = len(batches)
        self.reclaim_every = reclaim_every
        self.keep_pairs = keep_pairs
        self.f_off = _offsets([b.find_keys for b in batches])
        self.r_off = _offsets([b.range_lo for b in batches])
        self.i_off = _offsets([b.insert_keys for b in batches])
        self.e_off = _offsets([b.erase_keys for b in batches])
        z = lambda a: a if a.size else np.zeros(1, np.uint32)  # noqa: E731
        self.f_keys = _dev(z(_cat([b.find_keys for b in batches])), dev)
        self.r_lo = _dev(z(_cat([b.range_lo for b in batches])), dev)
        self.r_hi = _dev(z(_cat([b.range_hi for b in batches])), dev)
        self.i_keys = _dev(z(_cat([b.insert_keys for b in batches])), dev)
        self.i_vals = _dev(z(_cat([b.insert_values for b in batches])), dev)
        self.e_keys = _dev(z(_cat([b.erase_keys for b in batches])), dev)
        nf, nr = int(self.f_off[-1]), int(self.r_off[-1])
        self.f_out = torch.empty(max(nf, 1), dtype=torch.int32, device=dev)
        self.r_counts = torch.empty(max(nr, 1), dtype=torch.int32, device=dev)
        w = nr if keep_pairs else int(np.max(np.diff(self.r_off))) if self.nb else 0
        self.pairs = torch.full((max(w, 1) * L,), -1, dtype=torch.int64, device=dev)

    @property
    def num_ops(self) -> int:
        return int(self.f_off[-1] + self.r_off[-1] + self.i_off[-1] + self.e_off[-1])

    def _args(self, first: int):
        p = lambda a: ctypes.c_void_p(a.ctypes.data + 8 * first)  # noqa: E731
        d = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        return (p(self.f_off), d(self.f_keys), d(self.f_out), p(self.r_off), d(self.r_lo), d(self.r_hi),
                d(self.pairs), d(self.r_counts), self.L, p(self.i_off), d(self.i_keys), d(self.i_vals))

    def run(self, tree, stream=None, first: int = 0, count: int | None = None):
        """Runs batches [first, first + count) (default: all), asynchronously."""
        from . import _check, _stream
        count = self.nb - first if count is None else count
        if count <= 0:
            return
        a = self._args(first)
        _check(driver().mvbt_driver_per_batch(tree.handle, count, *a, ctypes.c_void_p(self.e_off.ctypes.data + 8 * first),
                                              ctypes.c_void_p(self.e_keys.data_ptr()), self.reclaim_every,
                                              0 if self.keep_pairs else 1, _stream(stream)), "driver_per_batch")

    def run_one(self, tree, i: int, stream=None):
        self.run(tree, stream, first=i, count=1)

    def batch_results(self, i: int):
        """Host copies of batch i's (find values, range counts, range pairs);
        pairs need keep_pairs=True (or i the last batch run)."""
        f = self.f_out[int(self.f_off[i]):int(self.f_off[i + 1])].cpu().numpy().view(np.uint32)
        a, b = int(self.r_off[i]), int(self.r_off[i + 1])
        c = self.r_counts[a:b].cpu().numpy().view(np.uint32)
        base = a if self.keep_pairs else 0
        p = self.pairs[base * self.L:(base + b - a) * self.L].cpu().numpy().view(np.uint64)
        return f, c, p

    def run_requests(self, tree, arrivals, stream=None):
        """Plain per-request server: batch b is issued at host time
        arrivals[b] (seconds from the call) or when b - 1 has completed,
        whichever is later, and waited for. Returns completion times."""
        from . import _check, _stream
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        done = np.zeros(self.nb, dtype=np.float64)
        a = self._args(0)
        _check(driver().mvbt_driver_per_request(tree.handle, self.nb, ctypes.c_void_p(arr.ctypes.data),
                                                ctypes.c_void_p(done.ctypes.data), *a, self.reclaim_every,
                                                0 if self.keep_pairs else 1, _stream(stream)), "driver_per_request")
        return done


class BatchedRunner:
    """Device-resident op stream for the persistent form."""

    def __init__(self, batches, L: int, dev, seed: int = 0):
        import torch
        ops, offsets, slots = wl.to_ops(batches, seed=seed)
        self.ops_host, self.offsets_host, self.slots = ops, offsets, slots
        self.L = L
        self.nops = ops.shape[0]
        self.ops = _dev(ops, dev).reshape(-1, 4)
        self.offsets = _dev(offsets, dev)
        self.results = torch.empty(max(self.nops, 1), dtype=torch.int32, device=dev)
        self.pairs = torch.full((max(slots, 1) * L,), -1, dtype=torch.int64, device=dev)
        self.base = torch.zeros(1, dtype=torch.int32, device=dev)

    def run(self, tree, stream=None, use_graph: bool = True):
        tree.run_batches(self.ops, self.offsets, results=self.results, pairs=self.pairs,
                         pair_slots=max(self.slots, 1), average_range_length=self.L,
                         snapshot_base=self.base, stream=stream, use_graph=use_graph)
        return self.results
