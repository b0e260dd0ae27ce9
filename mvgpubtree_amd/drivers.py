"""Batch drivers over the tree API (used by the benchmark and the tests).

* ``PerBatchRunner`` - the plain per-batch launch path: for every batch take a
  snapshot, launch find and range_query at it, then insert and erase, then
  release the snapshot (one launch per operation type, as with the
  reference's host API).
* ``BatchedRunner`` - the persistent work-queue form: the whole stream of
  batches in one ``run_batches`` launch, optionally through the captured
  hipGraph.

Both produce the same observable results for the same batches.
"""
from __future__ import annotations

import numpy as np

from . import workloads as wl


def _dev(a, dev):
    import torch
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view({4: np.int32, 8: np.int64}[a.dtype.itemsize])).to(dev)


class PerBatchRunner:
    """Device-resident inputs for the per-batch path; run() issues the launches."""

    def __init__(self, batches, L: int, dev, reclaim_every: int = 64):
        import torch
        self.L = L
        self.reclaim_every = reclaim_every
        self.batches = []
        nf = max((b.find_keys.size for b in batches), default=0)
        nr = max((b.range_lo.size for b in batches), default=0)
        for b in batches:
            self.batches.append(dict(
                find=_dev(b.find_keys, dev) if b.find_keys.size else None,
                lo=_dev(b.range_lo, dev) if b.range_lo.size else None,
                hi=_dev(b.range_hi, dev) if b.range_lo.size else None,
                ik=_dev(b.insert_keys, dev) if b.insert_keys.size else None,
                iv=_dev(b.insert_values, dev) if b.insert_keys.size else None,
                ek=_dev(b.erase_keys, dev) if b.erase_keys.size else None))
        self.find_out = torch.empty(max(nf, 1), dtype=torch.int32, device=dev)
        self.pairs = torch.empty(max(nr, 1) * L, dtype=torch.int64, device=dev)
        self.counts = torch.empty(max(nr, 1), dtype=torch.int32, device=dev)

    def run(self, tree, stream=None, collect: bool = False):
        """Runs every batch; with collect=True returns per-batch host copies of
        (find values, range counts, range pairs) for parity checks."""
        out = []
        for i in range(len(self.batches)):
            r = self.run_one(tree, i, stream, collect)
            if collect:
                out.append(r)
        return out

    def run_one(self, tree, i: int, stream=None, collect: bool = False):
        """Snapshot, reads at it, writes, release (and a periodic reclaim)."""
        b = self.batches[i]
        res = None
        s = tree.take_snapshot(stream)
        if True:
            if b["find"] is not None:
                n = b["find"].numel()
                tree.find(b["find"], self.find_out[:n], timestamp=s, stream=stream)
            if b["lo"] is not None:
                n = b["lo"].numel()
                if collect:
                    self.pairs.fill_(-1)  # unused slots read back as all-ones
                tree.range_query(b["lo"], b["hi"], self.L, result=self.pairs[:n * self.L],
                                 counts=self.counts[:n], timestamp=s, stream=stream)
            if collect:
                f = self.find_out[:b["find"].numel()].cpu().numpy().view(np.uint32).copy() \
                    if b["find"] is not None else np.empty(0, np.uint32)
                c = self.counts[:b["lo"].numel()].cpu().numpy().view(np.uint32).copy() \
                    if b["lo"] is not None else np.empty(0, np.uint32)
                p = self.pairs[:b["lo"].numel() * self.L].cpu().numpy().view(np.uint64).copy() \
                    if b["lo"] is not None else np.empty(0, np.uint64)
                res = (f, c, p)
            if b["ik"] is not None:
                tree.insert(b["ik"], b["iv"], stream=stream)
            if b["ek"] is not None:
                tree.erase(b["ek"], stream=stream)
            tree.release_snapshot(s)
            if self.reclaim_every and (i + 1) % self.reclaim_every == 0:
                tree.reclaim(stream)
        return res


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
