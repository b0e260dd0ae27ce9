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
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", LIB, src]
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
