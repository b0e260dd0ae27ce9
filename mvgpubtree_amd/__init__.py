"""MI355X-native multiversion B-link tree (host mirror over the C ABI).

The product is ``libmvbt.so`` (HIP kernels for gfx950 + the C ABI declared in
``include/mvbt.h``). This module binds that ABI with ctypes and mirrors the
reference's host API (owensgroup/MVGpuBTree ``gpu_versioned_btree`` and
``gpu_blink_tree``, include/gpu_versioned_blink_tree.hpp:106-302 and
include/gpu_blink_tree.hpp:218-265): ``insert``, ``find``, ``erase``,
``range_query`` and ``take_snapshot`` with the same argument meaning. PyTorch
only provides device memory and streams: every array argument is a device
tensor with 4-byte elements (``torch.int32`` / ``torch.uint32``), or 8-byte
elements for range-query pairs.

Loading fails loudly when ``libmvbt.so`` is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

__all__ = [
    "GpuBlinkTree", "GpuVersionedBTree", "MvbtError", "LATEST", "NOT_FOUND",
    "FLAG_POOL_EXHAUSTED", "FLAG_RETIRE_OVERFLOW", "FLAG_INVALID_KEY", "FLAG_VERSION_GAP",
    "FLAG_RANGE_OVERFLOW", "FLAG_CORRUPT", "FLAG_SNAPSHOT_ID", "FLAG_STAMP_ORDER", "FLAG_BAD_BATCH",
    "CONCURRENT", "library_path", "load_library", "EXPORTED_SYMBOLS",
]

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmvbt.so")

LATEST = 0xFFFFFFFF
NOT_FOUND = 0xFFFFFFFF
CONCURRENT = 0x1
USE_GRAPH = 0x2
NO_RECLAIM = 0x4
OP_FIND, OP_INSERT, OP_ERASE, OP_RANGE = 0, 1, 2, 3
FLAG_POOL_EXHAUSTED = 0x1
FLAG_RETIRE_OVERFLOW = 0x2
FLAG_INVALID_KEY = 0x4
FLAG_VERSION_GAP = 0x8
FLAG_RANGE_OVERFLOW = 0x10
FLAG_CORRUPT = 0x20
FLAG_SNAPSHOT_ID = 0x40
FLAG_STAMP_ORDER = 0x80
FLAG_BAD_BATCH = 0x100
PLAIN, VERSIONED = 0, 1
NOALLOC_MAX_OPS, NOALLOC_MAX_BATCHES = 1 << 19, 1 << 14  # mvbt_run_batches_noalloc limits

_STATUS = {0: "OK", 1: "INVALID_ARG", 2: "OUT_OF_MEMORY", 3: "HIP", 4: "UNSUPPORTED", 5: "DEVICE_FLAGS"}

# Every symbol declared in include/mvbt.h.
EXPORTED_SYMBOLS = (
    "mvbt_last_error", "mvbt_create", "mvbt_destroy", "mvbt_insert", "mvbt_find", "mvbt_erase",
    "mvbt_range_query", "mvbt_take_snapshot", "mvbt_release_snapshot", "mvbt_reclaim",
    "mvbt_memory_stats", "mvbt_error_flags", "mvbt_clear_error_flags", "mvbt_validate",
    "mvbt_export", "mvbt_node_width", "mvbt_kind_of", "mvbt_run_batches", "mvbt_set_worker_waves",
    "mvbt_memory_stats_async", "mvbt_device_view", "mvbt_bulk_build", "mvbt_concurrent_insert_range",
    "mvbt_concurrent_find_erase", "mvbt_export_dot", "mvbt_set_worker_mode", "mvbt_set_read_tile", "mvbt_run_batches_noalloc",
    "mvbt_reserve_batches", "mvbt_memory_series", "mvbt_set_batch_window",
)


class MvbtError(RuntimeError):
    def __init__(self, status: int, op: str, msg: str):
        super().__init__(f"{op}: {_STATUS.get(status, status)}: {msg}")
        self.status = status


class _Config(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("node_width", ctypes.c_uint32),
                ("pool_nodes", ctypes.c_uint64), ("retire_capacity", ctypes.c_uint64),
                ("device", ctypes.c_int32), ("reserved", ctypes.c_uint32 * 3)]


# Fields of one memory-timeline record (mvbt_epoch_stat).
EPOCH_STAT_FIELDS = ("epoch", "clock", "live_nodes", "allocated_total", "freed_nodes", "retired_pending",
                     "reclaimed_total", "capacity_nodes")


class _MemStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "node_bytes", "capacity_nodes", "live_nodes", "peak_live_nodes", "free_nodes",
        "retired_pending", "reclaimed_nodes", "retired_total")]


@dataclass
class MemoryStats:
    node_bytes: int
    capacity_nodes: int
    live_nodes: int
    peak_live_nodes: int
    free_nodes: int
    retired_pending: int
    reclaimed_nodes: int
    retired_total: int

    @property
    def live_bytes(self) -> int:
        return self.live_nodes * self.node_bytes

    @property
    def peak_bytes(self) -> int:
        return self.peak_live_nodes * self.node_bytes

    @property
    def free_bytes(self) -> int:
        return self.free_nodes * self.node_bytes


_lib = None


def library_path() -> str:
    return LIB_PATH


def load_library():
    """Loads libmvbt.so and declares its C ABI. Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                          "(python -m mvgpubtree_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    u32p, u64p = ctypes.POINTER(u32), ctypes.POINTER(u64)
    sig = {
        "mvbt_last_error": ([], ctypes.c_char_p),
        "mvbt_create": ([ctypes.POINTER(_Config), ctypes.POINTER(vp)], i32),
        "mvbt_destroy": ([vp], i32),
        "mvbt_insert": ([vp, vp, vp, u64, vp, u32], i32),
        "mvbt_find": ([vp, vp, vp, u64, u32, vp, u32], i32),
        "mvbt_erase": ([vp, vp, u64, vp, u32], i32),
        "mvbt_range_query": ([vp, vp, vp, vp, vp, u32, u64, u32, vp, u32], i32),
        "mvbt_take_snapshot": ([vp, vp, u32p], i32),
        "mvbt_release_snapshot": ([vp, u32], i32),
        "mvbt_reclaim": ([vp, vp], i32),
        "mvbt_memory_stats": ([vp, ctypes.POINTER(_MemStats)], i32),
        "mvbt_error_flags": ([vp, u32p], i32),
        "mvbt_clear_error_flags": ([vp], i32),
        "mvbt_validate": ([vp, u64p, u32p], i32),
        "mvbt_export": ([vp, vp, u64, u64p], i32),
        "mvbt_node_width": ([vp, u32p], i32),
        "mvbt_kind_of": ([vp, u32p], i32),
        "mvbt_run_batches": ([vp, vp, u64, vp, u32, vp, vp, u64, u32, vp, vp, u32], i32),
        "mvbt_run_batches_noalloc": ([vp, vp, u64, vp, u32, vp, vp, u64, u32, vp, vp], i32),
        "mvbt_reserve_batches": ([vp, u64, u32], i32),
        "mvbt_memory_series": ([vp, vp, u32, u32p], i32),
        "mvbt_set_batch_window": ([vp, u64], i32),
        "mvbt_set_worker_waves": ([vp, u32], i32),
        "mvbt_set_worker_mode": ([vp, u32], i32),
        "mvbt_set_read_tile": ([vp, u32], i32),
        "mvbt_memory_stats_async": ([vp, vp, vp], i32),
        "mvbt_device_view": ([vp, vp], i32),
        "mvbt_bulk_build": ([vp, vp, vp, u64, u32, vp], i32),
        "mvbt_concurrent_insert_range": ([vp, vp, vp, u64, vp, vp, u64, vp, vp, u32, vp], i32),
        "mvbt_concurrent_find_erase": ([vp, vp, vp, u64, vp, u64, vp], i32),
        "mvbt_export_dot": ([vp, ctypes.c_char_p, u64, u64p, u32], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def sort_unique_pairs(keys, values):
    """Device sort of uint32 keys (int32 tensors holding the bit patterns) with
    their values; of repeated keys the last occurrence is kept."""
    torch = _torch()
    k64 = keys.to(torch.int64) & 0xFFFFFFFF
    order = torch.sort(k64, stable=True).indices
    ks, vs = k64[order], values[order]
    keep = torch.ones_like(ks, dtype=torch.bool)
    keep[:-1] = ks[1:] != ks[:-1]
    ks, vs = ks[keep], vs[keep]
    return torch.where(ks >= 2**31, ks - 2**32, ks).to(torch.int32).contiguous(), vs.contiguous()


def _check(status: int, op: str):
    if status != 0:
        msg = load_library().mvbt_last_error()
        raise MvbtError(status, op, msg.decode() if msg else "")


def _torch():
    import torch  # device memory and streams only
    return torch


def _ptr(t, elem_bytes: int, name: str):
    if t is None:
        return None
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on the GPU")
    if not t.is_cuda:
        raise ValueError(f"{name} must live in device memory")
    if t.element_size() != elem_bytes or not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous with {elem_bytes}-byte elements")
    return ctypes.c_void_p(t.data_ptr())


def _ptr_any(t, elem_bytes: int, name: str):
    """Like _ptr, but also accepts pinned host tensors."""
    torch = _torch()
    if not isinstance(t, torch.Tensor) or t.element_size() != elem_bytes or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous tensor with {elem_bytes}-byte elements")
    if not (t.is_cuda or t.is_pinned()):
        raise ValueError(f"{name} must be device or pinned host memory")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        torch = _torch()
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


class _TreeBase:
    _kind = PLAIN

    def __init__(self, pool_nodes: int = 1 << 22, node_width: int = 16,
                 retire_capacity: int = 1 << 22, device: int = -1):
        L = load_library()
        cfg = _Config(kind=self._kind, node_width=node_width, pool_nodes=pool_nodes,
                      retire_capacity=retire_capacity, device=device)
        h = ctypes.c_void_p()
        _check(L.mvbt_create(ctypes.byref(cfg), ctypes.byref(h)), "mvbt_create")
        self._h = h
        self.node_width = node_width
        self.branching_factor = node_width

    def close(self):
        if getattr(self, "_h", None):
            load_library().mvbt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # --- batch API (reference host methods) -------------------------------
    def bulk_build(self, keys, values, fill: int = 0, sorted_input: bool = True, stream=None):
        """Loads (key, value) pairs into this new tree, `fill` entries per node
        (reference sorted-input constructor, gpu_blink_tree.hpp:85-136). With
        sorted_input=False the pairs are first sorted by key on the device and
        repeated keys keep their last value, as a sequence of inserts would."""
        n = keys.numel()
        if values.numel() != n:
            raise ValueError("keys and values differ in length")
        if not sorted_input and n:
            keys, values = sort_unique_pairs(keys, values)
            n = keys.numel()
        _check(load_library().mvbt_bulk_build(self._h, _ptr(keys, 4, "keys"), _ptr(values, 4, "values"),
                                              n, fill, _stream(stream)), "bulk_build")

    def concurrent_insert_range(self, keys, values, lower, upper, average_range_length: int,
                                result=None, counts=None, stream=None):
        """Inserts and range queries in one launch (reference concurrent_insert_range).
        Returns (pairs int64[num_ranges * L], counts int32[num_ranges])."""
        torch = _torch()
        n_i, n_r, L = keys.numel(), lower.numel(), average_range_length
        if values.numel() != n_i or upper.numel() != n_r:
            raise ValueError("keys/values or lower/upper differ in length")
        dev = (keys if n_i else lower).device
        if result is None:
            result = torch.full((max(n_r, 1) * L,), -1, dtype=torch.int64, device=dev)
        if counts is None:
            counts = torch.empty(max(n_r, 1), dtype=torch.int32, device=dev)
        if result.numel() < n_r * L or counts.numel() < n_r:
            raise ValueError("result or counts too short")
        _check(load_library().mvbt_concurrent_insert_range(
            self._h, _ptr(keys, 4, "keys") if n_i else None, _ptr(values, 4, "values") if n_i else None, n_i,
            _ptr(lower, 4, "lower") if n_r else None, _ptr(upper, 4, "upper") if n_r else None, n_r,
            _ptr(result, 8, "result"), _ptr(counts, 4, "counts"), L, _stream(stream)), "concurrent_insert_range")
        return result, counts

    def concurrent_find_erase(self, find_keys, erase_keys, find_results=None, stream=None):
        """Finds and erases in one launch (reference concurrent_find_erase). Returns find results."""
        torch = _torch()
        n_f, n_e = find_keys.numel(), erase_keys.numel()
        if find_results is None:
            find_results = torch.empty(max(n_f, 1), dtype=torch.int32,
                                       device=(find_keys if n_f else erase_keys).device)
        _check(load_library().mvbt_concurrent_find_erase(
            self._h, _ptr(find_keys, 4, "find_keys") if n_f else None, _ptr(find_results, 4, "find_results"), n_f,
            _ptr(erase_keys, 4, "erase_keys") if n_e else None, n_e, _stream(stream)), "concurrent_find_erase")
        return find_results

    def insert(self, keys, values, stream=None):
        n = keys.numel()
        if values.numel() != n:
            raise ValueError("keys and values differ in length")
        _check(load_library().mvbt_insert(self._h, _ptr(keys, 4, "keys"), _ptr(values, 4, "values"),
                                          n, _stream(stream), 0), "insert")

    def erase(self, keys, stream=None, concurrent: bool = False):
        _check(load_library().mvbt_erase(self._h, _ptr(keys, 4, "keys"), keys.numel(),
                                         _stream(stream), CONCURRENT if concurrent else 0), "erase")

    def find(self, keys, values=None, timestamp: int = LATEST, stream=None, concurrent: bool = False):
        torch = _torch()
        if values is None:
            values = torch.empty(keys.numel(), dtype=torch.int32, device=keys.device)
        if values.numel() < keys.numel():
            raise ValueError("values shorter than keys")
        _check(load_library().mvbt_find(self._h, _ptr(keys, 4, "keys"), _ptr(values, 4, "values"),
                                        keys.numel(), timestamp, _stream(stream),
                                        CONCURRENT if concurrent else 0), "find")
        return values

    def range_query(self, lower, upper, average_range_length: int, result=None, counts=None,
                    timestamp: int = LATEST, stream=None, concurrent: bool = False,
                    want_pairs: bool = True):
        """Returns (pairs int64[n*L] with key in the low word, counts int32[n])."""
        torch = _torch()
        n, L = lower.numel(), average_range_length
        if upper.numel() != n:
            raise ValueError("lower and upper differ in length")
        if result is None and want_pairs:
            result = torch.full((n * L,), -1, dtype=torch.int64, device=lower.device)
        if counts is None:
            counts = torch.empty(n, dtype=torch.int32, device=lower.device)
        if result is not None and result.numel() < n * L:
            raise ValueError("result shorter than n * average_range_length")
        _check(load_library().mvbt_range_query(
            self._h, _ptr(lower, 4, "lower"), _ptr(upper, 4, "upper"), _ptr(result, 8, "result"),
            _ptr(counts, 4, "counts"), L, n, timestamp, _stream(stream),
            CONCURRENT if concurrent else 0), "range_query")
        return result, counts

    # --- maintenance -----------------------------------------------------------
    def memory_stats(self) -> MemoryStats:
        s = _MemStats()
        _check(load_library().mvbt_memory_stats(self._h, ctypes.byref(s)), "memory_stats")
        return MemoryStats(**{f: getattr(s, f) for f, _ in _MemStats._fields_})

    def memory_series(self, capacity: int = 4096):
        """The memory timeline: one record per reclamation epoch (the latest
        `capacity`, oldest first) with fields EPOCH_STAT_FIELDS."""
        import numpy as np
        buf = np.zeros((capacity, len(EPOCH_STAT_FIELDS)), dtype=np.uint64)
        n = ctypes.c_uint32()
        _check(load_library().mvbt_memory_series(self._h, ctypes.c_void_p(buf.ctypes.data), capacity,
                                                 ctypes.byref(n)), "memory_series")
        return np.rec.fromarrays(buf[: n.value].T, names=list(EPOCH_STAT_FIELDS))

    def memory_stats_async(self, out, stream=None):
        """Stream-ordered stats into `out` (8 uint64 words in device or pinned
        host memory: a torch int64 tensor of 8 elements)."""
        _check(load_library().mvbt_memory_stats_async(self._h, _ptr_any(out, 8, "out"), _stream(stream)),
               "memory_stats_async")

    def set_read_tile(self, lanes: int):
        """Lanes per tile of the launches in which no writer runs (finds and
        ranges without `concurrent`, the sorted form's read passes): 0 or the
        node width (one pair per lane), or 16 / 32 below it (node_width /
        lanes pairs per lane)."""
        _check(load_library().mvbt_set_read_tile(self._h, lanes), "set_read_tile")

    def error_flags(self) -> int:
        f = ctypes.c_uint32()
        _check(load_library().mvbt_error_flags(self._h, ctypes.byref(f)), "error_flags")
        return f.value

    def clear_error_flags(self):
        _check(load_library().mvbt_clear_error_flags(self._h), "clear_error_flags")

    def validate_tree_structure(self):
        """Returns (num_keys, height); raises MvbtError when an invariant fails."""
        nk, h = ctypes.c_uint64(), ctypes.c_uint32()
        _check(load_library().mvbt_validate(self._h, ctypes.byref(nk), ctypes.byref(h)), "validate")
        return nk.value, h.value

    def to_dot(self, max_nodes: int = 0) -> str:
        """Graphviz text of the latest tree (reference plot_dot); render with `dot -Tpdf`."""
        n = ctypes.c_uint64()
        _check(load_library().mvbt_export_dot(self._h, None, 0, ctypes.byref(n), max_nodes), "export_dot")
        buf = ctypes.create_string_buffer(n.value + 1)
        _check(load_library().mvbt_export_dot(self._h, buf, n.value + 1, ctypes.byref(n), max_nodes), "export_dot")
        return buf.value.decode()

    def export(self):
        """Latest pairs in key order as a numpy uint64 array (key in the low word)."""
        import numpy as np
        cnt = ctypes.c_uint64()
        _check(load_library().mvbt_export(self._h, None, 0, ctypes.byref(cnt)), "export")
        out = np.empty(cnt.value, dtype=np.uint64)
        _check(load_library().mvbt_export(self._h, ctypes.c_void_p(out.ctypes.data), out.size,
                                          ctypes.byref(cnt)), "export")
        return out[: cnt.value]


class GpuBlinkTree(_TreeBase):
    """Plain concurrent B-link tree (reference ``gpu_blink_tree``)."""
    _kind = PLAIN


class GpuVersionedBTree(_TreeBase):
    """Multiversion B-link tree (reference ``gpu_versioned_btree``)."""
    _kind = VERSIONED

    def take_snapshot(self, stream=None) -> int:
        sid = ctypes.c_uint32()
        _check(load_library().mvbt_take_snapshot(self._h, _stream(stream), ctypes.byref(sid)),
               "take_snapshot")
        return sid.value

    def release_snapshot(self, snapshot_id: int):
        _check(load_library().mvbt_release_snapshot(self._h, snapshot_id), "release_snapshot")

    def reclaim(self, stream=None):
        _check(load_library().mvbt_reclaim(self._h, _stream(stream)), "reclaim")

    def run_batches(self, ops, batch_offsets, results=None, pairs=None, pair_slots: int = 0,
                    average_range_length: int = 0, snapshot_base=None, stream=None,
                    use_graph: bool = False, num_ops: int | None = None, no_alloc: bool = False,
                    no_reclaim: bool = False):
        """Batched form: `ops` is an int32 device tensor of shape (n, 4) holding
        {type, key, arg, slot} (OP_* types); `batch_offsets` an int64 device
        tensor of num_batches + 1 absolute op indices into `ops`; results go to
        the same indices (the default results tensor has one entry per row of
        `ops`). `num_ops` = offsets[-1] - offsets[0] (default: all of `ops`).
        A batch must not write one key twice (the call is then rejected:
        FLAG_BAD_BATCH). no_alloc=True goes through mvbt_run_batches_noalloc
        (static scratch, at most NOALLOC_MAX_OPS operations and
        NOALLOC_MAX_BATCHES batches; pass `results` to allocate nothing here
        either). Returns the results tensor."""
        torch = _torch()
        n = ops.shape[0] if num_ops is None else num_ops
        if results is None:
            results = torch.empty(ops.shape[0], dtype=torch.int32, device=ops.device)
        nb = batch_offsets.numel() - 1
        if no_alloc:
            _check(load_library().mvbt_run_batches_noalloc(
                self._h, _ptr(ops, 4, "ops"), n, _ptr(batch_offsets, 8, "batch_offsets"), nb,
                _ptr(results, 4, "results"), _ptr(pairs, 8, "pairs"), pair_slots,
                average_range_length, _ptr(snapshot_base, 4, "snapshot_base"), _stream(stream)), "run_batches_noalloc")
            return results
        _check(load_library().mvbt_run_batches(
            self._h, _ptr(ops, 4, "ops"), n, _ptr(batch_offsets, 8, "batch_offsets"), nb,
            _ptr(results, 4, "results"), _ptr(pairs, 8, "pairs"), pair_slots,
            average_range_length, _ptr(snapshot_base, 4, "snapshot_base"), _stream(stream),
            (USE_GRAPH if use_graph else 0) | (NO_RECLAIM if no_reclaim else 0)), "run_batches")
        return results

    def reserve_batches(self, num_ops: int, num_batches: int):
        """Sizes the batched form's scratch (and builds its launch graph) for
        calls of up to num_ops operations and num_batches batches."""
        _check(load_library().mvbt_reserve_batches(self._h, num_ops, num_batches), "reserve_batches")

    def set_batch_window(self, ops_per_window: int):
        """Operations per window of the sorted form (0: one window)."""
        _check(load_library().mvbt_set_batch_window(self._h, ops_per_window), "set_batch_window")

    def set_worker_waves(self, waves_per_cu: int):
        _check(load_library().mvbt_set_worker_waves(self._h, waves_per_cu), "set_worker_waves")

    def set_worker_mode(self, mode: int):
        """0: one operation per tile; 1: runs of operations per tile; 2: the
        sorted form: writes grouped by leaf in batch order, then every read at
        its batch's snapshot; 3 (default): mode 0 for calls of fewer than
        16384 operations, mode 2 from there up."""
        _check(load_library().mvbt_set_worker_mode(self._h, mode), "set_worker_mode")
