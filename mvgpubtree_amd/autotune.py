"""Autotuner for the batched form.

Sweeps node width (16, 32, 64 pairs), read tile width (16, 32, 64 lanes, at
most the node width: the lanes of the sorted form's read passes, each reading
node width / lanes pairs; writers and the persistent worker keep one pair per
lane), worker mode (one operation per tile, runs of operations per tile, or
the sorted form: writes grouped by leaf, then every read at its snapshot) and
worker occupancy (waves per CU: of the persistent worker, or of the sorted
form's write pass), times the batched
form on a sample of the target workload against a tree of the target size,
and returns the fastest configuration. Results can be cached in a JSON file
keyed by device name and workload shape.
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import asdict, dataclass

import numpy as np

from . import drivers


@dataclass
class TuneResult:
    node_width: int
    waves_per_cu: int
    ms: float
    mops: float
    worker_mode: int = 1
    read_tile: int = 0  # lanes per read tile (0: the node width)


def _device_name() -> str:
    import torch
    return torch.cuda.get_device_name(0)


def tune_batched(initial_keys, initial_values, groups, L: int, dev,
                 node_widths=(16, 32, 64), waves=(8, 16, 24, 32), cache: str | None = None,
                 log=None, modes=(1, 2), tiles=(16, 32, 64)):
    """Returns (best TuneResult, all results).

    `groups` is (warm, timed): two disjoint lists of batches cut from one
    workload. Read tiles apply to the sorted form (mode 2) only; a tile
    wider than the node or reading more than 4 pairs per lane is skipped.
    Every configuration (node width, read tile, worker mode, waves per CU)
    gets a tree freshly built from the full initial key set, runs `warm`
    (graph build for that grid, code load) and then `timed`, which is timed:
    all configurations do exactly the same work from the same start. (Timing
    each configuration on its own slice of a workload with log-uniform batch
    sizes let the slices' different mixes decide the winner.)"""
    import torch

    import mvgpubtree_amd as mv

    if len(groups) != 2:
        raise ValueError("groups must be (warm batches, timed batches)")
    key = f"{_device_name()}|{initial_keys.size}|{[len(g) for g in groups]}|{L}|{list(modes)}|{list(tiles)}|same"
    if cache and os.path.exists(cache):
        try:
            data = json.load(open(cache))
            if key in data:
                best = TuneResult(**data[key]["best"])
                return best, [TuneResult(**r) for r in data[key]["all"]]
        except (OSError, ValueError, TypeError):
            pass
    warm, timed = (drivers.BatchedRunner(g, L, dev, seed=i) for i, g in enumerate(groups))
    total_ops = warm.nops + timed.nops
    dk = drivers._dev(initial_keys, dev)
    dv = drivers._dev(initial_values, dev)
    results = []
    for nw in node_widths:
        pool = int(initial_keys.size / 4 * (16 / nw) + 2 * total_ops) + (1 << 20)
        wide = [t for t in tiles if t < nw and nw // t <= 4]
        combos = [(m, t, wv) for m in modes for t in ([0] + wide if m == 2 else [0]) for wv in waves]
        for mode, rt, wv in combos:
            tree = mv.GpuVersionedBTree(pool_nodes=pool, node_width=nw, retire_capacity=1 << 24)
            tree.insert(dk, dv)
            tree.set_worker_mode(mode)
            tree.set_worker_waves(wv)
            tree.set_read_tile(rt)
            tree.reserve_batches(max(warm.nops, timed.nops), max(len(g) for g in groups))
            warm.run(tree, use_graph=True)
            tree.reclaim()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            timed.run(tree, use_graph=True)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            if tree.error_flags():
                raise RuntimeError(f"autotune: device error flags 0x{tree.error_flags():x}")
            tree.close()
            r = TuneResult(nw, wv, ms, timed.nops / ms / 1e3, mode, rt)
            results.append(r)
            if log:
                log(f"autotune nw={nw} tile={rt or nw} mode={mode} waves/CU={wv}: {ms:.2f} ms "
                    f"({r.mops:.1f} M ops/s)")
    best = max(results, key=lambda r: r.mops)
    if cache:
        data = {}
        if os.path.exists(cache):
            try:
                data = json.load(open(cache))
            except (OSError, ValueError):
                data = {}
        data[key] = {"best": asdict(best), "all": [asdict(r) for r in results]}
        with open(cache, "w") as f:
            json.dump(data, f, indent=1)
    return best, results
