"""ORACLE - test infrastructure only.

Batch-state restatement of the reference's validation semantics
(owensgroup/MVGpuBTree include/validation.hpp:75-104 for finds,
:184-264 for ranges packed at i*L) for full-size workloads: the latest key set
is kept as sorted numpy arrays; every read of a batch is answered against the
state before the batch (the batch's pre-batch snapshot), then the batch's
writes are applied. Equivalent to ``MVMap`` reading at the snapshot taken
before each batch (cross-checked in tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np

EMPTY = np.uint32(0xFFFFFFFF)
KEY_MAX = 0xFFFFFFFF


def range_bounds(lower: np.ndarray, span: int) -> np.ndarray:
    """Upper bounds lower + span, saturated at 0xFFFFFFFF (exclusive bound)."""
    hi = lower.astype(np.uint64) + np.uint64(span)
    return np.minimum(hi, np.uint64(KEY_MAX)).astype(np.uint32)


class SortedState:
    def __init__(self, keys=None, values=None):
        if keys is None:
            self.keys = np.empty(0, dtype=np.uint32)
            self.values = np.empty(0, dtype=np.uint32)
        else:
            k = np.asarray(keys, dtype=np.uint32)
            v = np.asarray(values, dtype=np.uint32)
            order = np.argsort(k, kind="stable")
            k, v = k[order], v[order]
            # last write wins for duplicate keys
            last = np.ones(k.size, dtype=bool)
            if k.size > 1:
                last[:-1] = k[1:] != k[:-1]
            self.keys, self.values = k[last].copy(), v[last].copy()

    def __len__(self):
        return int(self.keys.size)

    def _locate(self, q: np.ndarray):
        q = np.asarray(q, dtype=np.uint32)
        idx = np.searchsorted(self.keys, q, side="left")
        safe = np.minimum(idx, max(self.keys.size - 1, 0))
        hit = (idx < self.keys.size) & (self.keys[safe] == q) if self.keys.size else np.zeros(q.size, bool)
        return idx, hit

    def find(self, q) -> np.ndarray:
        q = np.asarray(q, dtype=np.uint32)
        idx, hit = self._locate(q)
        out = np.full(q.size, EMPTY, dtype=np.uint32)
        if self.keys.size:
            out[hit] = self.values[idx[hit]]
        out[(q == 0) | (q == EMPTY)] = EMPTY
        return out

    def range_index(self, lower, upper):
        lo = np.maximum(np.asarray(lower, dtype=np.uint32), np.uint32(1))
        hi = np.asarray(upper, dtype=np.uint32)
        a = np.searchsorted(self.keys, lo, side="left")
        b = np.searchsorted(self.keys, hi, side="left")
        return a, np.maximum(b - a, 0).astype(np.int64)

    def range_counts(self, lower, upper) -> np.ndarray:
        return self.range_index(lower, upper)[1].astype(np.uint32)

    def range_query(self, lower, upper, L: int):
        """Returns (pairs u64[n*L], counts u32[n]); unused slots are all-ones."""
        a, c = self.range_index(lower, upper)
        n = a.size
        j = np.arange(L, dtype=np.int64)[None, :]
        take = j < np.minimum(c, L)[:, None]
        src = np.minimum(a[:, None] + j, max(self.keys.size - 1, 0))
        packed = self.keys.astype(np.uint64) | (self.values.astype(np.uint64) << np.uint64(32))
        out = np.full((n, L), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
        if self.keys.size:
            out[take] = packed[src[take]]
        return out.reshape(-1), c.astype(np.uint32)

    def insert(self, keys, values):
        k = np.asarray(keys, dtype=np.uint32)
        v = np.asarray(values, dtype=np.uint32)
        ok = (k != 0) & (k != EMPTY)
        k, v = k[ok], v[ok]
        batch = SortedState(k, v)  # dedupe inside the batch, last wins
        k, v = batch.keys, batch.values
        idx, hit = self._locate(k)
        self.values[idx[hit]] = v[hit]
        new = ~hit
        self.keys = np.insert(self.keys, idx[new], k[new])
        self.values = np.insert(self.values, idx[new], v[new])

    def erase(self, keys) -> int:
        k = np.unique(np.asarray(keys, dtype=np.uint32))
        idx, hit = self._locate(k)
        self.keys = np.delete(self.keys, idx[hit])
        self.values = np.delete(self.values, idx[hit])
        return int(hit.sum())


def batch_expected(initial_keys, initial_values, ops: np.ndarray, offsets: np.ndarray, L: int,
                   num_slots: int):
    """Expected results of the batched form: per batch, every read sees the
    state before the batch; then the batch's writes apply (a write of a
    reserved key fails: 0xFFFFFFFF, no effect). A batch whose writes name one
    key twice raises ValueError: the order of the two writes would decide the
    result. Returns (results u32[n], pairs u64[num_slots*L] with empty slots
    all-ones, final SortedState)."""
    st = SortedState(initial_keys, initial_values)
    n = ops.shape[0]
    res = np.zeros(n, dtype=np.uint32)
    pairs = np.full(num_slots * L, np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    for b in range(offsets.size - 1):
        lo, hi = int(offsets[b]), int(offsets[b + 1])
        blk = ops[lo:hi]
        ty = blk[:, 0]
        idx = np.arange(lo, hi)
        f = ty == 0
        res[idx[f]] = st.find(blk[f, 1])
        r = ty == 3
        if r.any():
            p, c = st.range_query(blk[r, 1], blk[r, 2], L)
            res[idx[r]] = c
            slots = blk[r, 3].astype(np.int64)
            pairs.reshape(-1, L)[slots] = p.reshape(-1, L)
        ins = ty == 1
        er = ty == 2
        key = blk[:, 1]
        bad = (key == 0) | (key == EMPTY)
        wk = key[(ins | er) & ~bad]
        if np.unique(wk).size != wk.size:
            raise ValueError(f"batch {b} writes one key twice")
        res[idx[(ins | er) & bad]] = EMPTY
        ins &= ~bad
        er &= ~bad
        if ins.any():
            res[idx[ins]] = (st.find(blk[ins, 1]) == EMPTY).astype(np.uint32)
        if er.any():
            res[idx[er]] = (st.find(blk[er, 1]) != EMPTY).astype(np.uint32)
        if ins.any():
            st.insert(blk[ins, 1], blk[ins, 2])
        if er.any():
            st.erase(blk[er, 1])
    return res, pairs, st
