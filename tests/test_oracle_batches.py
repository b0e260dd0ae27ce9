"""CPU tests pinning the batch-stream oracles.

Three restatements answer a stream of mixed-operation batches (reads of batch
b see exactly the writes of batches < b; reference per-batch pattern
take_snapshot -> find/range_query at it -> insert/erase,
gpu_versioned_blink_tree.hpp:106-302):

* ``oracle.replay_batches`` - C++ replay on the multiversion map;
* ``oracle.batch_expected`` - numpy sorted-array state;
* ``oracle.WriteOnceReplay`` - C++ per-op answers for streams that write no
  key twice (used at full size).

They are checked against a hand-computed stream, against a brute-force dict
replay on tiny key domains, against each other on config-4 and config-5
shaped streams (with empty, write-free and reserved-key batches), and for the
rejection of a batch that writes one key twice.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from mvgpubtree_amd import workloads as wl

MISS = 0xFFFFFFFF
EMPTY_PAIR = 0xFFFFFFFFFFFFFFFF
F, I, E, R = wl.OP_FIND, wl.OP_INSERT, wl.OP_ERASE, wl.OP_RANGE


def stream(batches):
    """batches: list of lists of (type, key, arg) -> (ops, offsets, slots)."""
    rows, offs, slot = [], [0], 0
    for b in batches:
        for t, k, a in b:
            s = 0
            if t == R:
                s, slot = slot, slot + 1
            rows.append((t, k, a, s))
        offs.append(len(rows))
    ops = np.array(rows, dtype=np.uint32).reshape(-1, 4)
    return ops, np.array(offs, dtype=np.uint64), slot


def write_once_all(initial_keys, initial_values, ops, offsets, L, slots):
    ks = np.asarray(initial_keys, dtype=np.uint32)
    vs = np.asarray(initial_values, dtype=np.uint32)
    o = np.argsort(ks)
    wo = oracle.WriteOnceReplay(ks[o], vs[o], ops, offsets)
    res, p = wo.answer(np.arange(ops.shape[0]), L)
    pairs = np.full((max(slots, 0), L), np.uint64(EMPTY_PAIR), dtype=np.uint64)
    rq = ops[:, 0] == R
    pairs[ops[rq, 3]] = p.reshape(-1, L)[rq]
    return res, pairs.reshape(-1)


def all_three(initial_keys, initial_values, ops, offsets, L, slots, write_once=True):
    a = oracle.replay_batches(initial_keys, initial_values, ops, offsets, L, slots)[:2]
    b = oracle.batch_expected(initial_keys, initial_values, ops, offsets, L, slots)[:2]
    out = [a, b]
    if write_once:
        out.append(write_once_all(initial_keys, initial_values, ops, offsets, L, slots))
    return out


def test_hand_computed_stream():
    """Expected values written out by hand: reads see the state before their
    batch; insert -> 1 new / 0 replaced; erase -> 1 erased / 0 absent; a
    reserved write key fails with 0xFFFFFFFF."""
    init_k, init_v = [10, 20, 30], [100, 200, 300]
    ops, off, slots = stream([
        [(F, 5, 0), (I, 5, 50), (F, 10, 0), (R, 1, 100), (E, 20, 0)],   # batch 0
        [(F, 5, 0), (F, 20, 0), (E, 5, 0), (R, 0, 31), (I, 10, 101)],   # batch 1
        [(F, 5, 0), (F, 10, 0), (I, 0, 7), (E, 99, 0), (R, 30, 30)],    # batch 2
        [],                                                            # batch 3 (empty)
        [(R, 1, 0xFFFFFFFF), (E, 0xFFFFFFFF, 0), (I, 40, 400)],         # batch 4
    ])
    L = 2
    want_res = [MISS, 1, 100, 3, 1,
                50, MISS, 1, 3, 0,
                MISS, 101, MISS, 0, 0,
                2, MISS, 1]
    p = lambda k, v: k | (v << 32)  # noqa: E731
    want_pairs = [p(10, 100), p(20, 200),          # slot 0: batch 0 [1, 100): 10 20 30 -> 3
                  p(5, 50), p(10, 100),            # slot 1: batch 1 [1, 31): 5 10 30 -> 3
                  EMPTY_PAIR, EMPTY_PAIR,          # slot 2: [30, 30) empty
                  p(10, 101), p(30, 300)]          # slot 3: batch 4: 10 30 -> 2
    for res, pairs in all_three(init_k, init_v, ops, off, L, slots, write_once=False):
        assert res.tolist() == want_res
        assert pairs.tolist() == want_pairs
    # this stream writes key 5 and key 10 twice overall, so only the exact
    # replays apply; the write-once replay refuses it
    with pytest.raises(oracle.BatchConflict):
        oracle.WriteOnceReplay(np.array(init_k, np.uint32), np.array(init_v, np.uint32), ops, off)


def brute(init, batches, L):
    """Dict replay: per batch, reads against a frozen copy, then writes in order."""
    st = dict(init)
    res, pairs = [], []
    for b in batches:
        snap = dict(st)
        for t, k, a in b:
            if t == F:
                res.append(snap.get(k, MISS) if 0 < k < MISS else MISS)
            elif t == R:
                lo = max(k, 1)
                ks = sorted(x for x in snap if lo <= x < a)
                res.append(len(ks))
                pairs.append([x | (snap[x] << 32) for x in ks[:L]] + [EMPTY_PAIR] * (L - min(L, len(ks))))
            else:
                res.append(None)
        j = len(res) - len(b)
        for t, k, a in b:
            if t in (I, E):
                if not (0 < k < MISS):
                    res[j] = MISS
                elif t == I:
                    res[j] = 0 if k in st else 1
                    st[k] = a
                else:
                    res[j] = 1 if k in st else 0
                    st.pop(k, None)
            j += 1
    return res, [x for row in pairs for x in row]


@pytest.mark.parametrize("seed", range(10))
def test_replays_match_brute_force_on_tiny_domains(seed):
    rng = np.random.default_rng(100 + seed)
    dom = int(rng.integers(3, 40))
    init = {int(k): int(rng.integers(0, 2**32 - 1)) for k in rng.choice(np.arange(1, dom + 1), dom // 2, replace=False)}
    batches = []
    for _ in range(int(rng.integers(1, 12))):
        n = int(rng.integers(0, 12))
        used, b = set(), []
        for _ in range(n):
            t = int(rng.integers(0, 4))
            k = int(rng.integers(0, dom + 2))
            if t in (I, E):
                if k in used:
                    continue
                used.add(k)
            a = int(rng.integers(0, 2**32 - 1)) if t == I else (int(rng.integers(0, dom + 3)) if t == R else 0)
            b.append((t, k, a))
        batches.append(b)
    L = int(rng.integers(1, 5))
    ops, off, slots = stream(batches)
    want_res, want_pairs = brute(init, batches, L)
    ik = np.array(list(init), dtype=np.uint32)
    iv = np.array([init[k] for k in init], dtype=np.uint32)
    for res, pairs in all_three(ik, iv, ops, off, L, slots, write_once=False):
        assert res.tolist() == want_res
        assert pairs.tolist() == want_pairs


def with_edge_cases(batches, rng, initial_sorted):
    """Adds empty and write-free batches, reserved keys, updates of present
    keys, erases of absent keys and unbounded ranges, keeping every key
    written at most once over the stream."""
    out = []
    used = set()
    for b in batches:
        for k in np.concatenate([b.insert_keys, b.erase_keys]):
            used.add(int(k))
    for i, b in enumerate(batches):
        out.append(b)
        if i % 7 == 0:
            out.append(wl.Batch())
        if i % 5 == 0:
            out.append(wl.Batch(find_keys=np.array([0, MISS, 1], np.uint32),
                                range_lo=np.array([0, 5, 0xFFFFFFFE], np.uint32),
                                range_hi=np.array([0x100000, 5, MISS], np.uint32)))
    upd = [int(k) for k in initial_sorted[::997] if int(k) not in used][:20]
    absent = [int(k) for k in wl.absent_keys(rng, 40, initial_sorted) if int(k) not in used][:20]
    extra = wl.Batch(insert_keys=np.array(upd + [0], np.uint32),
                     insert_values=np.array([7] * len(upd) + [1], np.uint32),
                     erase_keys=np.array(absent + [MISS], np.uint32),
                     find_keys=np.array(upd[:5], np.uint32))
    out.insert(len(out) // 2, extra)
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_replays_agree_on_config4_shape(seed):
    w = wl.config4(seed=seed, n_initial=30000, num_batches=96, min_batch=1, max_batch=3000, range_capacity=64)
    rng = np.random.default_rng(seed)
    batches = with_edge_cases(w.batches, rng, np.sort(w.initial_keys))
    ops, off, slots = wl.to_ops(batches, seed=seed)
    outs = all_three(w.initial_keys, w.initial_values, ops, off, w.range_capacity, slots)
    for res, pairs in outs[1:]:
        assert np.array_equal(res, outs[0][0])
        assert np.array_equal(pairs, outs[0][1])
    ty = ops[:, 0]
    res = outs[0][0]
    assert (res[ty == I] == MISS).sum() == 1 and (res[ty == E] == MISS).sum() == 1  # reserved write keys
    assert ((res[ty == I] == 0).sum() >= 1) and ((res[ty == E] == 0).sum() >= 1)     # updates, absent erases


def test_replays_agree_on_config5_shape():
    from mvgpubtree_amd import serving
    rng = np.random.default_rng(8)
    init = wl.unique_keys(rng, 50000)
    ks = np.sort(init)
    fresh = wl.unique_keys(rng, 200000, exclude_sorted=ks)
    reqs = serving.make_requests(rng, 150, 1000.0, ks, fresh, max_ops=2000, range_capacity=64)
    ops, off, slots = wl.to_ops(reqs.batches)
    outs = all_three(init, wl.value_of(init), ops, off, 64, slots)
    for res, pairs in outs[1:]:
        assert np.array_equal(res, outs[0][0]) and np.array_equal(pairs, outs[0][1])


def test_range_counts_exceed_capacity_exactly():
    """The true count is returned; only the first L pairs are written."""
    k = np.arange(1, 101, dtype=np.uint32)
    ops, off, slots = stream([[(R, 1, 101), (R, 50, 60)]])
    for res, pairs in all_three(k, k * 2, ops, off, 4, slots):
        assert res.tolist() == [100, 10]
        assert pairs.tolist()[:4] == [int(x) | (2 * int(x) << 32) for x in (1, 2, 3, 4)]
        assert pairs.tolist()[4:] == [int(x) | (2 * int(x) << 32) for x in (50, 51, 52, 53)]


@pytest.mark.parametrize("dup", [(I, I), (I, E), (E, I), (E, E)])
def test_batch_writing_a_key_twice_is_rejected(dup):
    k = np.array([10, 20], np.uint32)
    ops, off, slots = stream([[(F, 10, 0)], [(dup[0], 10, 1), (F, 20, 0), (dup[1], 10, 2)]])
    with pytest.raises(oracle.BatchConflict) as e:
        oracle.replay_batches(k, k, ops, off, 2, slots)
    assert e.value.op_index == 3
    with pytest.raises(ValueError):
        oracle.batch_expected(k, k, ops, off, 2, slots)
    with pytest.raises(oracle.BatchConflict):
        oracle.WriteOnceReplay(k, k, ops, off)
    # the same key in two different batches, or read and written in one, is fine
    ops, off, slots = stream([[(I, 10, 1), (F, 10, 0)], [(E, 10, 0), (R, 1, 50)]])
    res, _, _ = oracle.replay_batches(k, k, ops, off, 2, slots)
    assert res.tolist() == [0, 10, 1, 2]


def test_unknown_op_type_is_rejected():
    k = np.array([10], np.uint32)
    ops = np.array([[0, 10, 0, 0], [7, 10, 0, 0]], np.uint32)
    off = np.array([0, 2], np.uint64)
    with pytest.raises(oracle.BatchConflict):
        oracle.replay_batches(k, k, ops, off, 1, 0)
    with pytest.raises(oracle.BatchConflict):
        oracle.WriteOnceReplay(k, k, ops, off)


def test_write_once_answers_subsets_independently():
    w = wl.config4(seed=11, n_initial=20000, num_batches=40, min_batch=1, max_batch=2000)
    ops, off, slots = wl.to_ops(w.batches)
    ks = np.sort(w.initial_keys)
    wo = oracle.WriteOnceReplay(ks, wl.value_of(ks), ops, off)
    full, fp = wo.answer(np.arange(ops.shape[0]), 64, threads=4)
    sub = np.random.default_rng(1).choice(ops.shape[0], 5000, replace=False)
    r, p = wo.answer(sub, 64, threads=1)
    assert np.array_equal(r, full[sub])
    assert np.array_equal(p.reshape(-1, 64), fp.reshape(-1, 64)[sub])
