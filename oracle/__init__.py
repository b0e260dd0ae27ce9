"""ORACLE - test infrastructure only (never imported by the library).

Checker for the MI355X B-link tree. Two restatements of the reference's
observable semantics (owensgroup/MVGpuBTree include/validation.hpp and the
snapshot rules of include/gpu_versioned_blink_tree.hpp:297-353,1578-1603):

* ``MVMap``: ctypes wrapper over ``mvmap.cpp`` (std::map of per-key version
  histories). Exact for any interleaving of writes and snapshots; used on
  small and medium inputs and as the CPU baseline.
* ``SortedState`` (``sorted_ref.py``): numpy sorted-array state for batch
  workloads at full size (tens of millions of keys), where every read in a
  batch sees the state before the batch.
* ``replay_batches``: exact replay of a stream of mixed-operation batches on
  the multiversion map (one snapshot before each batch), the semantics the
  batched and served forms claim.
* ``WriteOnceReplay``: the same results for streams in which no key is
  written twice (the benchmark streams), answered op by op on several
  threads from the sorted initial pairs; used at full size (40M-500M keys).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from .sorted_ref import SortedState, batch_expected, range_bounds  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
LATEST = 0xFFFFFFFF
EMPTY = 0xFFFFFFFF

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build(verbose: bool = False) -> str:
    """Compile mvmap.cpp into liboracle.so (g++ -O2)."""
    src = os.path.join(HERE, "mvmap.cpp")
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-pthread", "-o", LIB, src]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_build_sorted.argtypes = [ctypes.c_void_p, _u32p, _u32p, ctypes.c_uint64]
        L.oracle_insert.argtypes = [ctypes.c_void_p, _u32p, _u32p, ctypes.c_uint64]
        L.oracle_erase.argtypes = [ctypes.c_void_p, _u32p, ctypes.c_uint64]
        L.oracle_erase.restype = ctypes.c_uint64
        L.oracle_take_snapshot.argtypes = [ctypes.c_void_p]
        L.oracle_take_snapshot.restype = ctypes.c_uint32
        L.oracle_find.argtypes = [ctypes.c_void_p, _u32p, _u32p, ctypes.c_uint64, ctypes.c_uint32]
        L.oracle_range.argtypes = [ctypes.c_void_p, _u32p, _u32p, _u64p, _u32p, ctypes.c_uint32,
                                   ctypes.c_uint64, ctypes.c_uint32]
        L.oracle_size.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.oracle_size.restype = ctypes.c_uint64
        L.oracle_export.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint64, ctypes.c_uint32]
        L.oracle_export.restype = ctypes.c_uint64
        L.oracle_replay_batches.argtypes = [ctypes.c_void_p, _u32p, ctypes.c_uint64, _u64p, ctypes.c_uint32,
                                            _u32p, _u64p, ctypes.c_uint64, ctypes.c_uint32]
        L.oracle_replay_batches.restype = ctypes.c_int64
        L.oracle_wo_create.argtypes = [_u32p, _u32p, ctypes.c_uint64, _u32p, ctypes.c_uint64, _u64p,
                                       ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32), _u64p]
        L.oracle_wo_create.restype = ctypes.c_void_p
        L.oracle_wo_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_wo_answer.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint64, _u32p, _u64p, ctypes.c_uint32,
                                       ctypes.c_uint32]
        _lib = L
    return _lib


def _p32(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def _p64(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


class MVMap:
    """Multiversion ordered map with the reference's snapshot semantics."""

    def __init__(self):
        self._h = lib().oracle_create()

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_sorted(self, keys, values):
        k, v = _u32(keys), _u32(values)
        lib().oracle_build_sorted(self._h, _p32(k), _p32(v), k.size)

    def insert(self, keys, values):
        k, v = _u32(keys), _u32(values)
        lib().oracle_insert(self._h, _p32(k), _p32(v), k.size)

    def erase(self, keys) -> int:
        k = _u32(keys)
        return int(lib().oracle_erase(self._h, _p32(k), k.size))

    def take_snapshot(self) -> int:
        return int(lib().oracle_take_snapshot(self._h))

    def find(self, keys, ts: int = LATEST) -> np.ndarray:
        k = _u32(keys)
        out = np.empty(k.size, dtype=np.uint32)
        lib().oracle_find(self._h, _p32(k), _p32(out), k.size, ts)
        return out

    def range_query(self, lower, upper, L: int, ts: int = LATEST):
        """Returns (pairs u64[n*L] with empty slots = all-ones, counts u32[n])."""
        lo, hi = _u32(lower), _u32(upper)
        out = np.full(lo.size * L, np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
        counts = np.empty(lo.size, dtype=np.uint32)
        lib().oracle_range(self._h, _p32(lo), _p32(hi), _p64(out), _p32(counts), L, lo.size, ts)
        return out, counts

    def size(self, ts: int = LATEST) -> int:
        return int(lib().oracle_size(self._h, ts))

    def export(self, ts: int = LATEST) -> np.ndarray:
        n = self.size(ts)
        out = np.empty(n, dtype=np.uint64)
        lib().oracle_export(self._h, _p64(out), n, ts)
        return out


class BatchConflict(ValueError):
    """A batch writes one key twice, or names an unknown operation type."""

    def __init__(self, op_index: int):
        super().__init__(f"batch stream rejected at op {op_index}")
        self.op_index = op_index


def replay_batches(initial_keys, initial_values, ops, offsets, L: int, num_slots: int):
    """Exact expected results of a batch stream (rows {type, key, arg, slot}):
    batch by batch, a snapshot, every read at it, then the writes in op order.
    Returns (results u32[n], pairs u64[num_slots * L] with unused slots
    all-ones, MVMap holding the final state). Raises BatchConflict when a
    batch writes one key twice."""
    k = _u32(initial_keys)
    order = np.argsort(k, kind="stable")
    k, v = k[order], _u32(initial_values)[order]
    last = np.ones(k.size, dtype=bool)
    if k.size > 1:
        last[:-1] = k[1:] != k[:-1]
    m = MVMap()
    m.build_sorted(np.ascontiguousarray(k[last]), np.ascontiguousarray(v[last]))
    o = np.ascontiguousarray(ops, dtype=np.uint32).reshape(-1, 4)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    res = np.zeros(o.shape[0], dtype=np.uint32)
    pairs = np.full(max(num_slots, 0) * L, np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    rc = lib().oracle_replay_batches(m._h, _p32(o), o.shape[0], _p64(off), off.size - 1, _p32(res),
                                     _p64(pairs) if pairs.size else None, num_slots, L)
    if rc < 0:
        raise BatchConflict(-rc - 1)
    return res, pairs, m


class WriteOnceReplay:
    """Expected results of a batch stream in which no key is written twice,
    for any subset of its ops (each answer depends only on the initial pairs
    and the stream's write events). `initial_keys` must be sorted, distinct
    and free of reserved keys."""

    def __init__(self, initial_keys, initial_values, ops, offsets):
        self._k = _u32(initial_keys)
        self._v = _u32(initial_values)
        self._ops = np.ascontiguousarray(ops, dtype=np.uint32).reshape(-1, 4)
        self._off = np.ascontiguousarray(offsets, dtype=np.uint64)
        st, bad = ctypes.c_int32(), ctypes.c_uint64()
        self._h = lib().oracle_wo_create(_p32(self._k), _p32(self._v), self._k.size, _p32(self._ops),
                                         self._ops.shape[0], _p64(self._off), self._off.size - 1,
                                         ctypes.byref(st), ctypes.byref(bad))
        if st.value == 1:
            raise ValueError(f"initial keys not strictly ascending or reserved at {bad.value}")
        if st.value in (2, 3):
            raise BatchConflict(int(bad.value))

    def close(self):
        if self._h:
            lib().oracle_wo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def answer(self, op_idx, L: int = 0, threads: int = 16):
        """Returns (results u32[m], pairs u64[m * L] or None) for ops op_idx."""
        idx = np.ascontiguousarray(op_idx, dtype=np.uint64)
        res = np.empty(idx.size, dtype=np.uint32)
        pairs = None
        if L:
            pairs = np.full(idx.size * L, np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
        lib().oracle_wo_answer(self._h, _p64(idx), idx.size, _p32(res), _p64(pairs) if L else None, L,
                               threads)
        return res, pairs
