"""Seeded synthetic workloads for the five benchmark configurations.

Keys are uniform in [1, 2^32 - 2], unique, drawn with numpy's PCG64 from a
fixed seed; value = key ^ 0x9E3779B9. The key 0x61C88646 is excluded because
its value would be 0xFFFFFFFF, the miss sentinel, making a hit on it
indistinguishable from a miss. Configurations:

1. plain tree: insert n unique keys; n finds, half present, half absent;
2. versioned tree: build n_initial keys, then batches of ops: half inserts of
   new unique keys, half range queries [lo, lo + span) with span chosen so a
   range holds ~32 keys, all reading the snapshot taken before the batch;
3. versioned tree: build, then batches of half erases of live keys and half
   finds (half hits, half misses) at the pre-batch snapshot;
4. batched form: many batches with log-uniform sizes and a 60/20/10/10
   find/insert/erase/range mix (range sizes ~8-128 keys);
5. served form: log-uniform request sizes, 80/10/10 find/range/insert.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

KEY_LO = 1
KEY_HI = 0xFFFFFFFE  # inclusive
VALUE_XOR = 0x9E3779B9
SENTINEL_COLLISION = 0x61C88646  # key ^ VALUE_XOR == 0xFFFFFFFF
KEY_SPACE = KEY_HI - KEY_LO + 1

OP_FIND, OP_INSERT, OP_ERASE, OP_RANGE = 0, 1, 2, 3


def value_of(keys: np.ndarray) -> np.ndarray:
    return (np.asarray(keys, dtype=np.uint32) ^ np.uint32(VALUE_XOR)).astype(np.uint32)


def span_for(expected_keys: float, n_live: int) -> int:
    """Range width whose expected key count is `expected_keys` at n_live uniform keys."""
    return int(np.ceil(expected_keys * KEY_SPACE / max(n_live, 1)))


def _draw(rng: np.random.Generator, n: int) -> np.ndarray:
    return rng.integers(KEY_LO, KEY_HI + 1, size=n, dtype=np.uint64).astype(np.uint32)


def unique_keys(rng: np.random.Generator, n: int, exclude_sorted: np.ndarray | None = None) -> np.ndarray:
    """n distinct keys in random order, none in `exclude_sorted` (sorted uint32)."""
    have = np.empty(0, dtype=np.uint32)
    while have.size < n:
        need = n - have.size
        cand = np.unique(np.concatenate([have, _draw(rng, int(need * 1.05) + 1024)]))
        cand = cand[cand != SENTINEL_COLLISION]
        if exclude_sorted is not None and exclude_sorted.size:
            pos = np.searchsorted(exclude_sorted, cand)
            pos = np.minimum(pos, exclude_sorted.size - 1)
            cand = cand[exclude_sorted[pos] != cand]
        have = cand
    # `have` is sorted; choose n of them uniformly and shuffle
    pick = rng.choice(have.size, size=n, replace=False) if have.size > n else rng.permutation(have.size)
    return have[pick]


def absent_keys(rng: np.random.Generator, n: int, present_sorted: np.ndarray) -> np.ndarray:
    """n keys (not necessarily distinct) uniform over the domain, none present."""
    out = np.empty(0, dtype=np.uint32)
    while out.size < n:
        c = _draw(rng, (n - out.size) + 64)
        c = c[c != SENTINEL_COLLISION]
        if present_sorted.size:
            pos = np.minimum(np.searchsorted(present_sorted, c), present_sorted.size - 1)
            c = c[present_sorted[pos] != c]
        out = np.concatenate([out, c])
    return out[:n]


@dataclass
class Config1:
    keys: np.ndarray
    values: np.ndarray
    queries: np.ndarray


def config1(seed: int = 1, n: int = 1_000_000) -> Config1:
    rng = np.random.default_rng(seed)
    keys = unique_keys(rng, n)
    sk = np.sort(keys)
    hits = rng.choice(keys, size=n // 2, replace=False)
    miss = absent_keys(rng, n - n // 2, sk)
    q = np.concatenate([hits, miss])
    rng.shuffle(q)
    return Config1(keys, value_of(keys), q)


@dataclass
class Batch:
    insert_keys: np.ndarray = field(default_factory=lambda: np.empty(0, np.uint32))
    insert_values: np.ndarray = field(default_factory=lambda: np.empty(0, np.uint32))
    erase_keys: np.ndarray = field(default_factory=lambda: np.empty(0, np.uint32))
    find_keys: np.ndarray = field(default_factory=lambda: np.empty(0, np.uint32))
    range_lo: np.ndarray = field(default_factory=lambda: np.empty(0, np.uint32))
    range_hi: np.ndarray = field(default_factory=lambda: np.empty(0, np.uint32))

    @property
    def num_ops(self) -> int:
        return int(self.insert_keys.size + self.erase_keys.size + self.find_keys.size + self.range_lo.size)


@dataclass
class MixedWorkload:
    initial_keys: np.ndarray
    initial_values: np.ndarray
    batches: list
    range_capacity: int
    span: int = 0


def _hi(lo: np.ndarray, span: int) -> np.ndarray:
    return np.minimum(lo.astype(np.uint64) + np.uint64(span), np.uint64(0xFFFFFFFF)).astype(np.uint32)


def config2(seed: int = 2, n_initial: int = 40_000_000, num_batches: int = 10,
            ops_per_batch: int = 1_000_000, expected_range: int = 32,
            range_capacity: int = 128) -> MixedWorkload:
    rng = np.random.default_rng(seed)
    n_ins = ops_per_batch // 2
    n_rq = ops_per_batch - n_ins
    keys = unique_keys(rng, n_initial + num_batches * n_ins)
    initial = keys[:n_initial]
    span = span_for(expected_range, n_initial)
    batches = []
    for b in range(num_batches):
        ik = keys[n_initial + b * n_ins: n_initial + (b + 1) * n_ins]
        lo = _draw(rng, n_rq)
        batches.append(Batch(insert_keys=ik, insert_values=value_of(ik), range_lo=lo, range_hi=_hi(lo, span)))
    return MixedWorkload(initial, value_of(initial), batches, range_capacity, span)


def config3(seed: int = 3, n_initial: int = 40_000_000, num_batches: int = 10,
            ops_per_batch: int = 1_000_000) -> MixedWorkload:
    rng = np.random.default_rng(seed)
    initial = unique_keys(rng, n_initial)
    n_er = ops_per_batch // 2
    n_find = ops_per_batch - n_er
    order = rng.permutation(n_initial)  # erase order: distinct live keys
    live = np.sort(initial)
    batches = []
    for b in range(num_batches):
        er = initial[order[b * n_er:(b + 1) * n_er]]
        hits = live[rng.integers(0, live.size, size=n_find // 2)]
        miss = absent_keys(rng, n_find - n_find // 2, live)
        q = np.concatenate([hits, miss])
        rng.shuffle(q)
        batches.append(Batch(erase_keys=er, find_keys=q))
        live = np.setdiff1d(live, er, assume_unique=True)
    return MixedWorkload(initial, value_of(initial), batches, 0)


def log_uniform_sizes(rng: np.random.Generator, count: int, lo: int, hi: int) -> np.ndarray:
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size=count))).astype(np.int64)


def config4(seed: int = 4, n_initial: int = 40_000_000, num_batches: int = 4096,
            min_batch: int = 32, max_batch: int = 65536, range_capacity: int = 256) -> MixedWorkload:
    """Mix 60% find / 20% insert / 10% erase / 10% range (8-128 expected keys)."""
    rng = np.random.default_rng(seed)
    sizes = log_uniform_sizes(rng, num_batches, min_batch, max_batch)
    n_ins_total = int(np.ceil(0.2 * sizes.sum())) + num_batches
    keys = unique_keys(rng, n_initial + n_ins_total)
    initial = keys[:n_initial]
    fresh = keys[n_initial:]
    live = np.sort(initial)
    erase_pool = rng.permutation(initial)
    fi = ei = 0
    batches = []
    for s in sizes:
        s = int(s)
        n_f, n_i, n_e = int(0.6 * s), int(0.2 * s), int(0.1 * s)
        n_r = s - n_f - n_i - n_e
        ik = fresh[fi:fi + n_i]; fi += n_i
        ek = erase_pool[ei:ei + n_e]; ei += n_e
        fk = np.concatenate([live[rng.integers(0, live.size, size=n_f - n_f // 2)],
                             _draw(rng, n_f // 2)])
        rng.shuffle(fk)
        lo = _draw(rng, n_r)
        want = rng.integers(8, 129, size=n_r)
        span = np.ceil(want * (KEY_SPACE / live.size)).astype(np.uint64)
        hi = np.minimum(lo.astype(np.uint64) + span, np.uint64(0xFFFFFFFF)).astype(np.uint32)
        batches.append(Batch(insert_keys=ik, insert_values=value_of(ik), erase_keys=ek,
                             find_keys=fk, range_lo=lo, range_hi=hi))
    return MixedWorkload(initial, value_of(initial), batches, range_capacity)


def to_ops(batches, seed: int = 0, shuffle: bool = True):
    """Flattens batches into the batched-form op stream.

    Returns (ops uint32[n, 4] rows {type, key, arg, slot}, offsets uint64[nb+1],
    number of range slots). Operations of different types are interleaved
    inside each batch (a batch's semantics do not depend on order: reads see
    the pre-batch state and the batch's writes touch distinct keys)."""
    rng = np.random.default_rng(seed)
    rows, offsets, slot = [], [0], 0
    for b in batches:
        parts = []
        if b.find_keys.size:
            z = np.zeros(b.find_keys.size, np.uint32)
            parts.append(np.stack([z + OP_FIND, b.find_keys, z, z], 1))
        if b.insert_keys.size:
            z = np.zeros(b.insert_keys.size, np.uint32)
            parts.append(np.stack([z + OP_INSERT, b.insert_keys, b.insert_values, z], 1))
        if b.erase_keys.size:
            z = np.zeros(b.erase_keys.size, np.uint32)
            parts.append(np.stack([z + OP_ERASE, b.erase_keys, z, z], 1))
        if b.range_lo.size:
            z = np.zeros(b.range_lo.size, np.uint32)
            slots = np.arange(slot, slot + b.range_lo.size, dtype=np.uint32)
            slot += b.range_lo.size
            parts.append(np.stack([z + OP_RANGE, b.range_lo, b.range_hi, slots], 1))
        blk = np.concatenate(parts).astype(np.uint32) if parts else np.zeros((0, 4), np.uint32)
        if shuffle:
            blk = blk[rng.permutation(blk.shape[0])]
        rows.append(blk)
        offsets.append(offsets[-1] + blk.shape[0])
    ops = np.ascontiguousarray(np.concatenate(rows)) if rows else np.zeros((0, 4), np.uint32)
    return ops, np.array(offsets, dtype=np.uint64), slot
