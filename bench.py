"""Benchmark: M ops/s of the multiversion B-link tree on one MI355X.

Default workload (the metric's configuration): versioned tree with 32-pair
nodes, uint32 keys and values, bulk insert of 40M unique uniform keys, then
batches of 1M ops each:
500K inserts of new keys and 500K range queries [lo, lo + span) holding ~32
keys (result slots L = 128), every range reading the snapshot taken before its
batch. One step = one batch: take_snapshot, range_query at it, insert,
release the snapshot, reclaim. Inputs are resident in HBM before timing.

Prints one JSON line with the contract fields plus:
  roofline      dominant kernel's algorithmic bytes / average launch time
                against 8 TB/s (HIP events on the launch stream);
  cpu_baseline  the oracle (std::map restatement of the reference's serial
                validation path) on one batch of the same workload on this
                host: one thread, and key-range shards over the host's cores;
  breakdown     per-operation rates, plus 1M finds (50% hits) on the measured
                tree (--find-node-width picks another width: a fresh tree of the
                same 40M keys).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per operation at 40M keys: reference layout, 128-B nodes,
# h = 8 levels (insert adds leaf write + version copy; range-32 adds 5 leaves
# and 32 result pairs).
H_REF = 8
BYTES = {
    "find": H_REF * 128 + 8,
    "insert": H_REF * 128 + 2 * 128 + 8,
    "range32": H_REF * 128 + 5 * 128 + 32 * 8 + 12,
    "erase": H_REF * 128 + 2 * 128 + 4,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def dev_tensor(a, dev):
    import torch
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view(np.int32 if a.dtype.itemsize == 4 else np.int64)).to(dev)


def cpu_baseline_config2(w, batch):
    """Times the oracle (std::map version histories) on one batch of the
    workload: range queries at the pre-batch snapshot, then the inserts."""
    import oracle
    m = oracle.MVMap()
    ks = np.sort(w.initial_keys)
    m.build_sorted(ks, (ks ^ np.uint32(0x9E3779B9)).astype(np.uint32))
    t0 = time.perf_counter()
    s = m.take_snapshot()
    total1 = int(m.range_query(batch.range_lo, batch.range_hi, w.range_capacity, s)[1].sum())
    m.insert(batch.insert_keys, batch.insert_values)
    dt = time.perf_counter() - t0
    ops = batch.num_ops
    m.close()
    # the same batch over key-range shards, one host thread each (the box's CPU
    # share for one GPU is 16 cores)
    P = max(1, min(16, os.cpu_count() or 1))
    vP, dtP, totalP = cpu_baseline_config2_sharded(w, batch, P)
    if totalP != total1:
        raise RuntimeError(f"sharded CPU baseline disagrees with one map: {totalP} vs {total1} range pairs")
    return {"value": round(vP, 3), "unit": "M ops/s", "cores": P, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"1 batch ({ops} ops: {batch.range_lo.size} range-32 + {batch.insert_keys.size} "
                      f"inserts) on the {w.initial_keys.size}-key map: {P} threads over key-range shards "
                      f"{dtP:.2f} s; one thread on one map {ops / dt / 1e6:.3f} M ops/s ({dt:.2f} s)"}


def cpu_baseline_config2_sharded(w, batch, threads: int):
    """The same batch on `threads` host threads: the key space is split into
    key-range shards (splitters from the initial keys), each shard is its own
    oracle map served by one thread, and a range query is clipped to every
    shard it overlaps (its counts add up). The oracle's C calls run outside
    the interpreter lock, so the threads run in parallel."""
    import threading

    import oracle
    ks = np.sort(w.initial_keys)
    vs = (ks ^ np.uint32(0x9E3779B9)).astype(np.uint32)
    P = threads
    cut = np.concatenate([[0], ks[(np.arange(1, P) * ks.size) // P].astype(np.uint64), [1 << 32]])
    maps = [oracle.MVMap() for _ in range(P)]
    lo64, hi64 = batch.range_lo.astype(np.uint64), batch.range_hi.astype(np.uint64)
    work = []
    for p in range(P):
        a, b = cut[p], cut[p + 1]
        sel = (ks >= a) & (ks < b)
        ins = (batch.insert_keys >= a) & (batch.insert_keys < b)
        hit = (lo64 < b) & (hi64 > a)
        work.append((ks[sel], vs[sel], batch.insert_keys[ins], batch.insert_values[ins],
                     np.maximum(lo64[hit], a).astype(np.uint32),
                     np.minimum(hi64[hit], b).astype(np.uint32) if b < (1 << 32)
                     else hi64[hit].astype(np.uint32)))
    builders = [threading.Thread(target=maps[p].build_sorted, args=(work[p][0], work[p][1])) for p in range(P)]
    for th in builders:
        th.start()
    for th in builders:
        th.join()
    counts = [None] * P

    def serve(p):
        _, _, ik, iv, lo, hi = work[p]
        s = maps[p].take_snapshot()
        counts[p] = maps[p].range_query(lo, hi, w.range_capacity, s)[1]
        maps[p].insert(ik, iv)

    runners = [threading.Thread(target=serve, args=(p,)) for p in range(P)]
    t0 = time.perf_counter()
    for th in runners:
        th.start()
    for th in runners:
        th.join()
    dt = time.perf_counter() - t0
    for m in maps:
        m.close()
    return batch.num_ops / dt / 1e6, dt, int(sum(int(c.sum()) for c in counts))


def run_config2(args):
    import torch

    import mvgpubtree_amd as mv
    from mvgpubtree_amd import workloads as wl

    dev = torch.device("cuda:0")
    n_initial = args.keys
    t0 = time.perf_counter()
    w = wl.config2(seed=2, n_initial=n_initial, num_batches=args.steps, ops_per_batch=args.ops)
    # Warmup runs the same step on a separate 1M-key tree, so the timed batches
    # start from exactly the configured build.
    ww = wl.config2(seed=99, n_initial=min(n_initial, 1_000_000), num_batches=max(args.warmup, 1),
                    ops_per_batch=args.ops)
    log(f"workload generated in {time.perf_counter() - t0:.1f}s (span {w.span})")
    L = w.range_capacity
    batches = w.batches
    dbat = [(dev_tensor(b.insert_keys, dev), dev_tensor(b.insert_values, dev),
             dev_tensor(b.range_lo, dev), dev_tensor(b.range_hi, dev)) for b in batches]
    n_rq = batches[0].range_lo.size
    out_pairs = torch.empty(n_rq * L, dtype=torch.int64, device=dev)
    out_counts = torch.empty(n_rq, dtype=torch.int32, device=dev)

    def make_tree(n):
        pool = int((n + args.steps * args.ops) / 4) + (1 << 20)
        t = mv.GpuVersionedBTree(pool_nodes=pool, node_width=args.node_width, retire_capacity=1 << 23)
        t.set_read_tile(args.read_tile if args.read_tile < args.node_width else 0)
        return t

    scratch = make_tree(ww.initial_keys.size)
    scratch.insert(dev_tensor(ww.initial_keys, dev), dev_tensor(ww.initial_values, dev))
    wbat = [(dev_tensor(b.insert_keys, dev), dev_tensor(b.insert_values, dev),
             dev_tensor(b.range_lo, dev), dev_tensor(b.range_hi, dev)) for b in ww.batches]

    tree = make_tree(n_initial)
    dk, dv = dev_tensor(w.initial_keys, dev), dev_tensor(w.initial_values, dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree.insert(dk, dv)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    log(f"built {n_initial} keys in {build_s:.3f}s ({n_initial / build_s / 1e6:.1f} M inserts/s)")
    del dk, dv

    stream = torch.cuda.current_stream()
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    kt = {"range32": [], "insert": []}

    def step(tree, batch, timed):
        ik, iv, lo, hi = batch
        s = tree.take_snapshot(stream)
        e0, e1, e2 = ev(), ev(), ev()
        e0.record(stream)
        tree.range_query(lo, hi, L, result=out_pairs, counts=out_counts, timestamp=s, stream=stream)
        e1.record(stream)
        tree.insert(ik, iv, stream=stream)
        e2.record(stream)
        tree.release_snapshot(s)
        tree.reclaim(stream)
        if timed:
            kt["range32"].append((e0, e1))
            kt["insert"].append((e1, e2))

    for b in range(args.warmup):
        step(scratch, wbat[b], False)
    torch.cuda.synchronize()
    scratch.close()
    t0 = time.perf_counter()
    for b in range(args.steps):
        step(tree, dbat[b], True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ops = sum(b.num_ops for b in batches)

    kernel_ms = {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in kt.items()}
    n_ins = batches[0].insert_keys.size
    rates = {"range32_Mops": n_rq / kernel_ms["range32"] / 1e3,
             "insert_Mops": n_ins / kernel_ms["insert"] / 1e3}
    flags = tree.error_flags()
    if flags:
        raise RuntimeError(f"device error flags 0x{flags:x}")

    # 1M finds (50% hits) on the same tree at the latest snapshot.
    rng = np.random.default_rng(77)
    present = w.initial_keys[rng.integers(0, n_initial, size=args.ops // 2)]
    absent = wl.absent_keys(rng, args.ops - args.ops // 2, np.sort(np.concatenate(
        [w.initial_keys] + [b.insert_keys for b in batches])))
    q = np.concatenate([present, absent])
    rng.shuffle(q)
    dq = dev_tensor(q, dev)
    fout = torch.empty(q.size, dtype=torch.int32, device=dev)
    mem = tree.memory_stats()
    # Finds run on a tree of the node width that suits them (a tile walks one
    # key at a time, so 16-pair tiles keep four walks per wave in flight):
    # the same 40M initial keys, freshly built when the widths differ.
    ftree = tree
    if args.find_node_width != args.node_width:
        ftree = mv.GpuVersionedBTree(pool_nodes=int(n_initial / 4) + (1 << 20), node_width=args.find_node_width,
                                     retire_capacity=1 << 20)
        ftree.insert(dev_tensor(w.initial_keys, dev), dev_tensor(w.initial_values, dev), stream=stream)
        ftree.set_read_tile(args.read_tile if args.read_tile < args.find_node_width else 0)
    sf = ftree.take_snapshot(stream)
    for _ in range(2):
        ftree.find(dq, fout, timestamp=sf, stream=stream)
    fe = []
    for _ in range(5):
        a, b = ev(), ev()
        a.record(stream)
        ftree.find(dq, fout, timestamp=sf, stream=stream)
        b.record(stream)
        fe.append((a, b))
    torch.cuda.synchronize()
    find_ms = float(np.mean([a.elapsed_time(b) for a, b in fe]))
    hits = int((fout.cpu().numpy().view(np.uint32) != 0xFFFFFFFF).sum())
    rates["find_Mops"] = q.size / find_ms / 1e3
    ftree.release_snapshot(sf)
    if ftree is not tree:
        if ftree.error_flags():
            raise RuntimeError(f"device error flags 0x{ftree.error_flags():x} (find tree)")
        ftree.close()

    # The same batches through the reference's one-launch mixed call
    # (concurrent_insert_range: every range of the batch at the call's
    # snapshot beside its inserts) on a fresh tree of the same keys: the
    # first batch warms up, the others are timed with their reclamations.
    mtree = make_tree(n_initial)
    mtree.insert(dev_tensor(w.initial_keys, dev), dev_tensor(w.initial_values, dev))
    for b in range(args.steps):
        if b == 1:
            torch.cuda.synchronize()
            tm = time.perf_counter()
        ik, iv, lo, hi = dbat[b]
        mtree.concurrent_insert_range(ik, iv, lo, hi, L, result=out_pairs, counts=out_counts, stream=stream)
        mtree.reclaim(stream)
    torch.cuda.synchronize()
    mixed_mops = sum(b.num_ops for b in batches[1:]) / (time.perf_counter() - tm) / 1e6 if args.steps > 1 else None
    if mtree.error_flags():
        raise RuntimeError(f"device error flags 0x{mtree.error_flags():x} (mixed call)")
    mtree.close()

    # Dominant kernel of the step and its roofline.
    share = {"range32": kernel_ms["range32"], "insert": kernel_ms["insert"]}
    dom = max(share, key=share.get)
    per_launch_bytes = (n_rq if dom == "range32" else n_ins) * BYTES[dom]
    achieved = per_launch_bytes / (kernel_ms[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": (f"range_wide_kernel<{args.node_width},versioned,{args.read_tile}>"
                           if 0 < args.read_tile < args.node_width else f"range_kernel<{args.node_width},versioned>")
                + f" ({n_rq} range-32 queries/launch)"
                if dom == "range32" else f"write_kernel<{args.node_width},versioned,insert> ({n_ins} inserts/launch)",
                "avg_launch_ms": round(kernel_ms[dom], 4)}
    traffic = load_traffic(dom, args.node_width)
    if traffic is not None:
        roofline["traffic"] = traffic

    value = ops / elapsed / 1e6
    result = {
        "metric": "M ops/s (find, insert, range-32) at 40M keys; p99 latency served vs plain",
        "value": round(value, 2),
        "unit": "M ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "uint32",
        "data": "synthetic (seeded uniform unique uint32 keys, value = key ^ 0x9E3779B9)",
        "config": {"workload": "versioned tree, 40M-key build, batches of 1M ops: 50% inserts of "
                               "new keys + 50% range queries (~32 keys, L=128) at the pre-batch "
                               "snapshot; plain per-batch launch path",
                   "initial_keys": n_initial, "ops_per_batch": args.ops, "node_width": args.node_width,
                   "read_tile": min(args.read_tile or args.node_width, args.node_width),
                   "range_span": w.span, "range_capacity": L},
        "roofline": roofline,
        "breakdown": {k: round(v, 2) for k, v in rates.items()} | {
            "build_Minserts_per_s": round(n_initial / build_s / 1e6, 2),
            "find_hits": hits, "find_node_width": args.find_node_width,
            "find_roofline_frac": round(q.size * BYTES["find"] / (find_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            # the whole step (inserts + ranges, snapshot and reclaim included) by algorithmic bytes
            "step_roofline_frac": round((n_ins * BYTES["insert"] + n_rq * BYTES["range32"]) * args.steps / elapsed
                                        / 1e9 / HBM_PEAK_GBS, 4),
            "mixed_insert_range_Mops": round(mixed_mops, 1) if mixed_mops else None,
            "live_nodes": mem.live_nodes, "peak_live_nodes": mem.peak_live_nodes,
            "reclaimed_nodes": mem.reclaimed_nodes},
    }
    tree.close()
    del dbat, out_pairs, out_counts
    if not args.no_extras:
        # the metric's other forms, measured in the same run: config 4
        # (batched form vs plain per-batch launches) and config 5 (served form
        # vs plain per-request launches, p50/p99 latency) at full size
        w4 = wl.config4(seed=4, n_initial=n_initial, num_batches=4096)
        c4 = batched_vs_plain(w4, w4.range_capacity, dev, args.c4_node_width, args.c4_mode, args.c4_waves,
                              read_tile=args.c4_read_tile)
        del w4
        log(f"config4: batched {c4['batched_Mops']} vs plain {c4['plain_per_batch_Mops']} M ops/s")
        levels = [float(q) for q in args.extras_qps.split(",")]
        c5 = served_vs_plain(args.served_keys, levels, args.extras_requests, args.step_ops, dev, args.served_mode,
                             args.served_node_width, args.served_read_tile)
        result["breakdown"]["plain_config1"] = run_config1(dev)
        c3 = run_config3(args)  # erase + find at the snapshot with reclamation, same key count and steps
        log(f"config3: {c3['Mops_incl_reclaim']} M ops/s, peak {c3['peak_live_MB']} MB")
        result["breakdown"]["erase_find_reclaim"] = {k: v for k, v in c3.items() if k != "timeline"} | {
            "epochs_recorded": len(c3["timeline"]) - 1,
            "last_epoch": c3["timeline"][-1]}
        sweep_args = argparse.Namespace(**vars(args))
        sweep_args.c7_sizes = args.extras_sweep_sizes
        c7 = run_config7(sweep_args)  # the reference's mixed-call sweeps, harmonic means over a subsample
        result["breakdown"]["mixed_sweeps"] = {
            "sizes_M": c7["sizes_M"], "node_width": c7["node_width"],
            "insert_range": [{k: (v["hmean_Mops"] if isinstance(v, dict) else v) for k, v in r.items()}
                             for r in c7["insert_range"]],
            "find_erase": [{k: (v["hmean_Mops"] if isinstance(v, dict) else v) for k, v in r.items()}
                           for r in c7["find_erase"]]}
        c6 = run_config6(args)  # the reclaimer's memory timeline at the reference's Figure-6 shape
        result["breakdown"]["reclaimer_memory"] = {k: v for k, v in c6.items() if k not in ("mixed", "per_batch")} | {
            f: {k: v for k, v in c6[f].items() if k != "timeline"} for f in ("mixed", "per_batch")}
        result["breakdown"]["batched_vs_plain"] = c4
        result["breakdown"]["served_vs_plain"] = {
            "keys": c5["keys"], "requests_per_level": c5["requests_per_level"],
            "levels": [{"qps": lv["qps"],
                        "served": {k: lv["served"][k] for k in ("Mops", "p50_ms", "p99_ms", "steps")},
                        "plain": {k: lv["plain"][k] for k in ("Mops", "p50_ms", "p99_ms")}}
                       for lv in c5["levels"]]}
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline_config2(w, batches[0])
    return result


def run_config1(dev, reps: int = 5):
    """Config 1: plain (unversioned) tree, 1M unique keys inserted in one
    launch, then 1M finds (half present); both launches timed with HIP
    events on their stream (the insert on fresh trees, each find launch on
    the built tree), results checked against the inserted values."""
    import torch

    import mvgpubtree_amd as mv
    from mvgpubtree_amd import workloads as wl

    w = wl.config1(seed=1)
    dk, dv, dq = dev_tensor(w.keys, dev), dev_tensor(w.values, dev), dev_tensor(w.queries, dev)
    out = torch.empty(dq.numel(), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    ins_ms, find_ms = [], []
    for rep in range(reps + 1):
        tree = mv.GpuBlinkTree(pool_nodes=1 << 19)
        a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a.record(stream)
        tree.insert(dk, dv, stream=stream)
        b.record(stream)
        tree.find(dq, out, stream=stream)
        c.record(stream)
        torch.cuda.synchronize()
        if rep:  # the first round loads the code
            ins_ms.append(a.elapsed_time(b))
            find_ms.append(b.elapsed_time(c))
        flags = tree.error_flags()
        tree.close()
        if flags:
            raise RuntimeError(f"device error flags 0x{flags:x} (config 1)")
    got = out.cpu().numpy().view(np.uint32)
    present = np.isin(w.queries, w.keys)
    if not (np.array_equal(got[present], w.queries[present] ^ np.uint32(wl.VALUE_XOR))
            and np.all(got[~present] == 0xFFFFFFFF)):
        raise RuntimeError("config 1 finds disagree with the inserted keys")
    n = w.keys.size
    return {"keys": n, "insert_Mops": round(n / float(np.median(ins_ms)) / 1e3, 1),
            "find_Mops": round(dq.numel() / float(np.median(find_ms)) / 1e3, 1)}


def run_config3(args):
    """Config 3: versioned tree, 40M-key build, then batches of 1M ops - 500K
    erases of live keys and 500K finds (half hits among the keys live at the
    snapshot, half misses) at the pre-batch snapshot - with epoch reclamation
    after every batch. Reports throughput and the per-batch memory timeline
    (live, peak, reclaimed and retired-but-unreclaimed node memory)."""
    import torch

    import mvgpubtree_amd as mv
    from mvgpubtree_amd import workloads as wl

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    w = wl.config3(seed=3, n_initial=args.keys, num_batches=args.steps, ops_per_batch=args.ops)
    ww = wl.config3(seed=33, n_initial=min(args.keys, 1_000_000), num_batches=1, ops_per_batch=args.ops // 10)
    log(f"config3 workload generated in {time.perf_counter() - t0:.1f}s")

    def run(tree, batches):
        out = torch.empty(max(b.find_keys.size for b in batches), dtype=torch.int32, device=dev)
        dev_b = [(dev_tensor(b.find_keys, dev), dev_tensor(b.erase_keys, dev)) for b in batches]
        torch.cuda.synchronize()
        t = time.perf_counter()
        for fk, ek in dev_b:
            s = tree.take_snapshot()
            tree.find(fk, out[:fk.numel()], timestamp=s)
            tree.erase(ek)
            tree.release_snapshot(s)
            tree.reclaim()  # appends one epoch to the device's memory timeline
        torch.cuda.synchronize()
        return time.perf_counter() - t

    def make(n):
        t = mv.GpuVersionedBTree(pool_nodes=int(n / 4 * 16 / args.c3_node_width) + args.steps * args.ops + (1 << 20),
                                 node_width=args.c3_node_width, retire_capacity=1 << 23)
        t.set_read_tile(args.c3_read_tile if args.c3_read_tile < args.c3_node_width else 0)
        return t

    scratch = make(ww.initial_keys.size)
    scratch.insert(dev_tensor(ww.initial_keys, dev), dev_tensor(ww.initial_values, dev))
    run(scratch, ww.batches)
    scratch.close()
    tree = make(args.keys)
    tree.insert(dev_tensor(w.initial_keys, dev), dev_tensor(w.initial_values, dev))
    torch.cuda.synchronize()
    st0 = tree.memory_stats()
    e0 = len(tree.memory_series())
    el = run(tree, w.batches)
    # the per-epoch timeline the reclamations recorded on the device (read once, after timing)
    nb = st0.node_bytes
    timeline = [{"live_MB": round(st0.live_bytes / 1e6, 1), "allocated_total_MB": round(st0.live_bytes / 1e6, 1),
                 "freed_MB": 0.0, "retired_pending": st0.retired_pending}]
    for r in tree.memory_series()[e0:]:
        timeline.append({"epoch": int(r.epoch), "clock": int(r.clock), "live_MB": round(int(r.live_nodes) * nb / 1e6, 1),
                         "allocated_total_MB": round(int(r.allocated_total) * nb / 1e6, 1),
                         "freed_MB": round(int(r.freed_nodes) * nb / 1e6, 1),
                         "retired_pending": int(r.retired_pending)})
    ops = sum(b.num_ops for b in w.batches)
    flags = tree.error_flags()
    st = tree.memory_stats()
    tree.close()
    if flags:
        raise RuntimeError(f"device error flags 0x{flags:x}")

    # The same batches through the reference's one-launch mixed call
    # (concurrent_find_erase: every find at the call's snapshot beside its
    # erases) on a fresh 32-pair tree of the same keys, reclaiming after each
    # call; the first batch warms up.
    mt = mv.GpuVersionedBTree(pool_nodes=int(args.keys / 8) + args.steps * args.ops + (1 << 20), node_width=32,
                              retire_capacity=1 << 23)
    mt.insert(dev_tensor(w.initial_keys, dev), dev_tensor(w.initial_values, dev))
    dev_b = [(dev_tensor(b.find_keys, dev), dev_tensor(b.erase_keys, dev)) for b in w.batches]
    fres = torch.empty(max(b.find_keys.size for b in w.batches), dtype=torch.int32, device=dev)
    mixed = None
    for b, (fk, ek) in enumerate(dev_b):
        if b == 1:
            torch.cuda.synchronize()
            tm = time.perf_counter()
        mt.concurrent_find_erase(fk, ek, find_results=fres[:fk.numel()])
        mt.reclaim()
    torch.cuda.synchronize()
    if len(dev_b) > 1:
        mixed = sum(b.num_ops for b in w.batches[1:]) / (time.perf_counter() - tm) / 1e6
    if mt.error_flags():
        raise RuntimeError(f"device error flags 0x{mt.error_flags():x} (mixed call)")
    mt.close()
    return {"keys": args.keys, "node_width": args.c3_node_width, "batches": len(w.batches), "ops": ops,
            "Mops_incl_reclaim": round(ops / el / 1e6, 1),
            "mixed_find_erase_Mops": round(mixed, 1) if mixed else None,
            "peak_live_MB": round(st.peak_bytes / 1e6, 1), "final_live_MB": round(st.live_bytes / 1e6, 1),
            "reclaimed_MB": round(st.reclaimed_nodes * st.node_bytes / 1e6, 1), "timeline": timeline}


def batched_vs_plain(w, L, dev, nw: int, mode: int, waves: int, plain: bool = True, read_tile: int = 0):
    """Config 4 on a fresh tree per path: the whole stream in one batched
    call through the captured hipGraph, against the plain per-batch launch
    path (C++ loop over the C ABI: snapshot, find, range, insert, erase,
    release per batch). Both start from the same 40M-key tree. plain=False
    runs the batched form only (profiling)."""
    import torch

    import mvgpubtree_amd as mv
    from mvgpubtree_amd import drivers
    from mvgpubtree_amd import workloads as wl

    n_ops = sum(b.num_ops for b in w.batches)
    n0 = w.initial_keys.size
    # spare slots: a fifth of the operation count for the sorted form, which
    # reclaims between its windows; the persistent worker keeps every copy
    pool = int(n0 / 4 * (16 / nw) + (n_ops / 5 if mode == 2 else 2 * n_ops)) + (1 << 20)
    dk, dv = dev_tensor(w.initial_keys, dev), dev_tensor(w.initial_values, dev)

    def fresh_tree():
        t = mv.GpuVersionedBTree(pool_nodes=pool, node_width=nw)
        t.insert(dk, dv)
        t.set_worker_waves(waves)
        t.set_worker_mode(mode)
        t.set_read_tile(read_tile)
        return t

    # warm both paths (code load, graph build) on a small scratch tree
    wb = wl.config4(seed=99, n_initial=1_000_000, num_batches=64)
    scratch = mv.GpuVersionedBTree(pool_nodes=1 << 21, node_width=nw)
    scratch.insert(dev_tensor(wb.initial_keys, dev), dev_tensor(wb.initial_values, dev))
    scratch.set_worker_waves(waves)
    scratch.set_worker_mode(mode)
    scratch.set_read_tile(read_tile)
    drivers.BatchedRunner(wb.batches, L, dev).run(scratch, use_graph=True)
    drivers.PerBatchRunner(wb.batches[:8], L, dev).run(scratch)
    torch.cuda.synchronize()
    scratch.close()

    br = drivers.BatchedRunner(w.batches, L, dev)
    tree = fresh_tree()
    tree.reserve_batches(br.nops, len(w.batches))  # scratch and graph: one-time, outside the timed call
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    br.run(tree, use_graph=True)
    torch.cuda.synchronize()
    batched_s = time.perf_counter() - t0
    flags = tree.error_flags()
    res_b = br.results.cpu().numpy().view(np.uint32)[: br.nops]
    ty = br.ops_host[:, 0]
    failed = int((res_b[(ty == wl.OP_INSERT) | (ty == wl.OP_ERASE)] == 0xFFFFFFFF).sum())
    tree.close()
    del br

    plain_s = float("nan")
    if plain:
        pr = drivers.PerBatchRunner(w.batches, L, dev)
        tree = fresh_tree()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pr.run(tree)
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t0
        flags |= tree.error_flags()
        tree.close()
        del pr
    del dk, dv
    if flags:
        raise RuntimeError(f"device error flags 0x{flags:x}")
    return {
        "batched_Mops": round(n_ops / batched_s / 1e6, 2),
        "plain_per_batch_Mops": round(n_ops / plain_s / 1e6, 2) if plain else None,
        "speedup": round(plain_s / batched_s, 2) if plain else None,
        "batched_ms": round(batched_s * 1e3, 2), "plain_ms": round(plain_s * 1e3, 2) if plain else None,
        "ops": n_ops, "batches": len(w.batches), "initial_keys": int(n0),
        "failed_writes": failed,
        "node_width": nw, "read_tile": read_tile or nw, "worker_mode": mode, "waves_per_cu": waves,
    }


def run_config4(args):
    """Batched form (config 4): 4096 batches, sizes log-uniform in [32, 65536],
    mix 60/20/10/10 find/insert/erase/range (8-128 keys), on a 40M-key tree.
    The persistent work-queue launch (captured hipGraph, autotuned node width,
    worker mode and waves/CU) against the plain per-batch launch path on an
    identical tree."""
    import torch

    from mvgpubtree_amd import autotune
    from mvgpubtree_amd import workloads as wl

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    w = wl.config4(seed=4, n_initial=args.keys, num_batches=args.batches)
    log(f"config4 workload: {sum(b.num_ops for b in w.batches)} ops in {len(w.batches)} batches "
        f"({time.perf_counter() - t0:.1f}s)")
    L = w.range_capacity
    waves = tuple(int(x) for x in args.waves.split(","))
    widths = tuple(int(x) for x in args.widths.split(","))
    modes = tuple(int(x) for x in args.modes.split(","))
    tiles = tuple(int(x) for x in args.tiles.split(","))
    ncfg = len(waves) * len(modes) * len(tiles)
    if len(widths) * ncfg > 1:
        tw = wl.config4(seed=44, n_initial=args.keys, num_batches=2 * args.tune_batches)
        groups = [tw.batches[0::2], tw.batches[1::2]]  # (warm, timed), the same for every configuration
        best, allr = autotune.tune_batched(tw.initial_keys, tw.initial_values, groups, L, dev,
                                           node_widths=widths, waves=waves, log=log, modes=modes, tiles=tiles)
        del tw
    else:
        best, allr = autotune.TuneResult(widths[0], waves[0], 0.0, 0.0, modes[0],
                                         tiles[0] if tiles[0] < widths[0] else 0), []
    log(f"autotune best: node width {best.node_width}, read tile {best.read_tile or best.node_width} lanes, "
        f"worker mode {best.worker_mode}, {best.waves_per_cu} waves/CU")
    out = batched_vs_plain(w, L, dev, best.node_width, best.worker_mode, best.waves_per_cu, plain=not args.no_plain,
                           read_tile=best.read_tile)
    out["tuning"] = [{"nw": r.node_width, "tile": r.read_tile or r.node_width, "mode": r.worker_mode,
                      "waves": r.waves_per_cu, "Mops": round(r.mops, 1)} for r in allr]
    return out


def gpu_unique_keys(n: int, extra: int, seed: int, dev):
    """n + extra distinct uniform keys in [1, 2^32 - 2] (sentinel-collision key
    excluded), random order, generated on the GPU; returned as int32 tensors
    holding the uint32 bit patterns."""
    import torch
    from mvgpubtree_amd import workloads as wl
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    want = n + extra
    u = torch.empty(0, dtype=torch.int64, device=dev)
    while u.numel() < want:
        d = torch.randint(1, 0xFFFFFFFF, (int((want - u.numel()) * 1.03) + 4096,), generator=g,
                          device=dev, dtype=torch.int64)
        u = torch.unique(torch.cat([u, d]))
        u = u[u != wl.SENTINEL_COLLISION]
    u = u[torch.randperm(u.numel(), generator=g, device=dev)[:want]]
    u32 = torch.where(u >= 2**31, u - 2**32, u).to(torch.int32)
    return u32[:n].contiguous(), u32[n:].contiguous()


def served_vs_plain(n: int, levels, n_req: int, step_ops: int, dev, mode: int = 3, nw: int = 16,
                    read_tile: int = 0):
    """Served form (config 5): n-key tree resident in HBM; open-loop Poisson
    arrivals at each QPS level; request sizes log-uniform in [1, 16384] ops,
    80% find / 10% range (~32 keys) / 10% insert, one snapshot per request;
    continuous batching against the plain per-request launch path (C++ loop:
    each request issued at its arrival or when the previous one completes,
    and waited for) on the same requests."""
    import torch

    import mvgpubtree_amd as mv
    from mvgpubtree_amd import serving
    from mvgpubtree_amd import workloads as wl

    t0 = time.perf_counter()
    init, fresh = gpu_unique_keys(n, (len(levels) + 1) * n_req * 2000, seed=5, dev=dev)
    vals = init ^ torch.tensor(wl.VALUE_XOR - 2**32, dtype=torch.int32, device=dev)
    pool = int(n / 6 * (16 / nw)) + (8 << 20)
    tree = mv.GpuVersionedBTree(pool_nodes=pool, node_width=nw)
    tree.set_read_tile(read_tile if read_tile < nw else 0)
    tree.set_worker_mode(mode)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    tree.insert(init, vals)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - tb
    log(f"served: built {n} keys in {build_s:.2f}s (setup {time.perf_counter() - t0:.1f}s)")
    rng = np.random.default_rng(55)
    sample = np.sort(init[torch.randint(0, n, (4_000_000,), device=dev)].cpu().numpy().view(np.uint32))
    fresh_h = fresh.cpu().numpy().view(np.uint32)
    del init, vals, fresh
    fi = 0

    def requests(qps, count):
        nonlocal fi
        r = serving.make_requests(rng, count, qps, sample, fresh_h[fi:], range_capacity=128, n_live=n)
        fi += sum(b.insert_keys.size for b in r.batches)
        return r

    # warmup: both paths, untimed; the served loop keeps one stream throughout
    stream = torch.cuda.Stream()
    w = requests(1e5, 64)
    serving.ContinuousBatchingServer(tree, w, dev, stream=stream).run()
    w = requests(2e3, 16)
    serving.ContinuousBatchingServer(tree, w, dev, stream=stream).run()
    serving.PerRequestServer(tree, w, dev).run()
    torch.cuda.synchronize()
    tree.reclaim()
    out = []
    for q in levels:
        reqs = requests(q, n_req)
        s_sum, _ = serving.ContinuousBatchingServer(tree, reqs, dev, max_step_ops=step_ops, stream=stream).run()
        # the same requests (sizes, mix, arrivals, keys): its inserts then
        # replace values the served run inserted, the same work per key
        p_sum, _ = serving.PerRequestServer(tree, reqs, dev).run()
        tree.reclaim()
        out.append({"qps": q, "served": s_sum, "plain": p_sum})
        log(f"qps {q:.0f}: served {s_sum['Mops']} M ops/s p50 {s_sum['p50_ms']} ms p99 {s_sum['p99_ms']} ms"
            f" (queue p99 {s_sum['queue_p99_ms']} ms, step gpu p50/p99 {s_sum['step_gpu_p50_ms']}/"
            f"{s_sum['step_gpu_p99_ms']} ms, {s_sum['steps']} steps of {s_sum['step_ops_mean']} ops, "
            f"enqueue p50/max {s_sum['enqueue_p50_ms']}/{s_sum['enqueue_max_ms']} ms, worst {s_sum['worst']})"
            f" | plain {p_sum['Mops']} M ops/s p50 {p_sum['p50_ms']} ms p99 {p_sum['p99_ms']} ms")
    flags = tree.error_flags()
    mem = tree.memory_stats()
    tree.close()
    if flags:
        raise RuntimeError(f"device error flags 0x{flags:x}")
    return {"keys": n, "node_width": nw, "read_tile": read_tile or nw, "build_Minserts_per_s": round(n / build_s / 1e6, 1), "requests_per_level": n_req,
            "levels": out, "pool_live_GB": round(mem.live_bytes / 1e9, 2),
            "pool_peak_GB": round(mem.peak_bytes / 1e9, 2)}


def run_config5(args):
    import torch
    levels = [float(q) for q in args.qps.split(",")]
    return served_vs_plain(args.served_keys, levels, args.requests, args.step_ops, torch.device("cuda:0"),
                           args.served_mode, args.served_node_width, args.served_read_tile)


def run_config6(args):
    """The reclaimer's memory timeline at the shape of the reference's
    concurrent_insert_range_reclaimer_bench.cu:50-153 and its archived
    memory_45m_45m_50_16.csv: a 45M-key versioned tree of 16-pair nodes,
    then 22.5M inserts of new keys and 22.5M range queries [k, k + 16) with
    k drawn from the initial keys (rkg.hpp:50-73), L = 16. The reference
    issues them as one concurrent_insert_range launch whose reclaimer
    advanced 440 epochs; here they run as the same number of mixed calls
    (concurrent_insert_range, ranges at the state before the call),
    reclaiming after each, so each call closes one epoch of the timeline,
    and again as per-batch launches (snapshot, ranges at it, inserts,
    release, reclaim). Memory is counted as the reference counts it: live
    nodes (retired but not yet freed included) and nodes ever allocated."""
    import torch

    import mvgpubtree_amd as mv

    dev = torch.device("cuda:0")
    n0, n1, calls, L, span = args.c6_keys, args.c6_keys // 2, args.c6_calls, 16, 16
    init, new = gpu_unique_keys(n0, n1, seed=6, dev=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(66)
    lo = init[torch.randperm(n0, generator=g, device=dev)[:n1]]
    lo64 = lo.to(torch.int64) & 0xFFFFFFFF
    hi64 = torch.clamp(lo64 + span, max=0xFFFFFFFF)
    hi = torch.where(hi64 >= 2**31, hi64 - 2**32, hi64).to(torch.int32)
    bounds = [round(i * n1 / calls) for i in range(calls + 1)]
    pairs = torch.empty(max(b - a for a, b in zip(bounds, bounds[1:])) * L, dtype=torch.int64, device=dev)
    counts = torch.empty(pairs.numel() // L, dtype=torch.int32, device=dev)

    def one(form):
        tree = mv.GpuVersionedBTree(pool_nodes=int((n0 + n1) / 4) + (1 << 22), node_width=16,
                                    retire_capacity=1 << 23)
        tree.insert(init, init)
        torch.cuda.synchronize()
        st0 = tree.memory_stats()
        e0 = len(tree.memory_series())
        t = time.perf_counter()
        for a, b in zip(bounds, bounds[1:]):
            m = b - a
            if form == "mixed":
                tree.concurrent_insert_range(new[a:b], new[a:b], lo[a:b], hi[a:b], L, result=pairs[:m * L],
                                             counts=counts[:m])
            else:
                s = tree.take_snapshot()
                tree.range_query(lo[a:b], hi[a:b], L, result=pairs[:m * L], counts=counts[:m], timestamp=s)
                tree.insert(new[a:b], new[a:b])
                tree.release_snapshot(s)
            tree.reclaim()
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        series = tree.memory_series()[e0:]
        st = tree.memory_stats()
        flags = tree.error_flags() & ~mv.FLAG_RANGE_OVERFLOW
        tree.close()
        if flags:
            raise RuntimeError(f"device error flags 0x{flags:x} ({form})")
        gib = st0.node_bytes / 2**30
        live = [st0.live_nodes] + [int(r.live_nodes) for r in series]
        total = [st0.live_nodes] + [int(r.allocated_total) for r in series]
        rows = [{"epoch": i, "live_nodes": live[i], "allocated_total": total[i], "ratio": round(live[i] / total[i], 4),
                 "GiB": round(live[i] * gib, 4)} for i in range(0, len(live), max(1, (len(live) - 1) // 10))]
        if rows[-1]["epoch"] != len(live) - 1:
            i = len(live) - 1
            rows.append({"epoch": i, "live_nodes": live[i], "allocated_total": total[i],
                         "ratio": round(live[i] / total[i], 4), "GiB": round(live[i] * gib, 4)})
        return {"Mops": round(2 * n1 / el / 1e6, 1), "epochs": len(series), "node_bytes": st0.node_bytes,
                "start_GiB": round(st0.live_nodes * gib, 4), "start_nodes": st0.live_nodes,
                "peak_GiB": round(max(live) * gib, 4), "peak_nodes": max(live),
                "peak_live_before_freeing_GiB": round(st.peak_live_nodes * gib, 4),
                "end_nodes": live[-1], "allocated_total": total[-1], "end_ratio": round(live[-1] / total[-1], 4),
                "timeline": rows}

    out = {"initial_keys": n0, "inserts": n1, "ranges": n1, "range_span": span, "range_capacity": L,
           "node_width": 16, "calls": calls}
    for form in ("mixed", "per_batch"):
        out[form] = one(form)
        log(f"config6 {form}: start {out[form]['start_GiB']} GiB, peak {out[form]['peak_GiB']} GiB, "
            f"live/allocated {out[form]['end_ratio']}")
    return out


def run_config7(args):
    """The reference's mixed-call sweeps (scripts/bench_concurrent_insert_range.sh,
    bench_concurrent_erase_find.sh; benchmarks/concurrent_insert_range_bench.cu,
    concurrent_erase_find_bench.cu): for every configuration, one mixed call
    of num_ops operations on a freshly built tree, num_ops = 1M, 2M, ... 45M,
    rate = num_ops / the call's time (HIP events around the call alone), and
    the harmonic mean over the sizes, on the versioned and the plain tree.
    Keys are a shuffled permutation of 1..N (rkg.hpp:77-97).

    insert_range: initial 1M or 40M keys, update ratio 5/50/90 %, range
    length 8 or 32: inserts of the keys not in the tree, ranges [k, k + len)
    with k from the initial keys (all of them shuffled, then sampled with
    replacement when more ranges are asked for, rkg.hpp:50-73), L = len.
    find_erase: initial 45M keys, update ratio 5/50/90 %: erases of the first
    keys generated, finds of the first keys of a shuffle of them (all present
    before the call, rkg.hpp:37-47)."""
    import torch

    import mvgpubtree_amd as mv

    dev = torch.device("cuda:0")
    sizes = [int(float(x) * 1e6) for x in args.c7_sizes.split(",")] if args.c7_sizes != "all" else \
        [m * 1_000_000 for m in range(1, 46)]
    kinds = args.c7_trees.split(",")
    widths = {"versioned": args.c7_node_width, "plain": args.c7_plain_node_width}
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    s = torch.cuda.current_stream()

    def as_u32(x):  # int64 key values < 2^31 -> int32 tensors of the same bits
        return x.to(torch.int32)

    def make(kind, n_keys, extra):
        cls = mv.GpuVersionedBTree if kind == "versioned" else mv.GpuBlinkTree
        pool = int(n_keys / 4) + int(extra) + (1 << 21)
        return cls(pool_nodes=pool, node_width=widths[kind], retire_capacity=1 << 24)

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        fn()
        b.record(s)
        b.synchronize()
        return a.elapsed_time(b) * 1e-3

    def hmean(v):
        return len(v) / sum(1.0 / x for x in v)

    for kind in kinds:  # loads both mixed kernels before the first timed call
        w = as_u32(torch.randperm(200_000, generator=g, device=dev) + 1)
        tree = make(kind, 200_000, 0)
        tree.insert(w[:100_000], w[:100_000])
        tree.concurrent_insert_range(w[100_000:], w[100_000:], w[:1000], w[:1000] + 8, 8)
        tree.concurrent_find_erase(w[:1000], w[1000:2000])
        torch.cuda.synchronize()
        tree.close()
    out = {"node_width": widths, "sizes_M": [x / 1e6 for x in sizes], "insert_range": [],
           "find_erase": []}
    max_ops = max(sizes)
    pairs = torch.empty(int(max_ops * 0.95) * 32 + 32, dtype=torch.int64, device=dev)
    counts = torch.empty(max_ops + 1, dtype=torch.int32, device=dev)
    for init in (1_000_000, 40_000_000):
        for rl in (8, 32):
            for ratio in (0.05, 0.5, 0.9):
                row = {"initial": init, "range_length": rl, "update_ratio": ratio}
                for kind in kinds:
                    rates = []
                    for n_ops in sizes:
                        n_ins = int(n_ops * ratio)
                        n_rq = n_ops - n_ins
                        perm = torch.randperm(init + n_ins, generator=g, device=dev) + 1
                        keys = as_u32(perm)
                        lo_src = keys[:init][torch.randperm(init, generator=g, device=dev)]
                        if n_rq > init:
                            extra = torch.randint(0, init, (n_rq - init,), generator=g, device=dev)
                            lo_src = torch.cat([lo_src, keys[:init][extra]])
                        lo = lo_src[:n_rq].contiguous()
                        hi = lo + rl
                        tree = make(kind, init + n_ins, 0)
                        tree.insert(keys[:init], keys[:init])
                        ins = keys[init:]
                        el = timed(lambda: tree.concurrent_insert_range(ins, ins, lo, hi, rl, result=pairs[:n_rq * rl],
                                                                        counts=counts[:n_rq]))
                        flags = tree.error_flags() & ~mv.FLAG_RANGE_OVERFLOW
                        tree.close()
                        if flags:
                            raise RuntimeError(f"device error flags 0x{flags:x} (insert_range {row} {kind} {n_ops})")
                        rates.append(n_ops / el / 1e6)
                        del perm, keys, lo_src, lo, hi, ins
                    row[kind] = {"hmean_Mops": round(hmean(rates), 1), "Mops": [round(x, 1) for x in rates]}
                    log(f"config7 insert_range init {init} range {rl} ratio {ratio} {kind}: "
                        f"hmean {row[kind]['hmean_Mops']} M ops/s")
                out["insert_range"].append(row)
    init = 45_000_000
    for ratio in (0.05, 0.5, 0.9):
        row = {"initial": init, "update_ratio": ratio}
        for kind in kinds:
            rates = []
            for n_ops in sizes:
                n_up = int(n_ops * ratio)
                n_q = n_ops - n_up
                keys = as_u32(torch.randperm(init, generator=g, device=dev) + 1)
                er = keys[:n_up].contiguous()
                fk = keys[torch.randperm(init, generator=g, device=dev)[:n_q]].contiguous()
                tree = make(kind, init, 0)
                tree.insert(keys, keys)
                res = counts[:max(n_q, 1)]
                el = timed(lambda: tree.concurrent_find_erase(fk, er, find_results=res))
                flags = tree.error_flags()
                tree.close()
                if flags:
                    raise RuntimeError(f"device error flags 0x{flags:x} (find_erase {row} {kind} {n_ops})")
                rates.append(n_ops / el / 1e6)
                del keys, er, fk
            row[kind] = {"hmean_Mops": round(hmean(rates), 1), "Mops": [round(x, 1) for x in rates]}
            log(f"config7 find_erase ratio {ratio} {kind}: hmean {row[kind]['hmean_Mops']} M ops/s")
        out["find_erase"].append(row)
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_traffic(kernel: str, node_width: int):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, copy-kernel calibrated), when a summary has been recorded for
    this node width."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    return data.get(kernel) if data.get("node_width") == node_width else None


def copy_calibration(nbytes: int = 1 << 30, reps: int = 5):
    """Plain device-to-device copies of known size: under rocprofv3 --pmc
    their FETCH_SIZE / WRITE_SIZE give the counters' bytes-per-unit scale."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--keys", type=int, default=40_000_000)
    ap.add_argument("--ops", type=int, default=1_000_000)
    ap.add_argument("--node-width", type=int, default=32,
                    help="config 2: node width of the measured tree (32 pairs measured fastest for "
                         "insert + range-32 batches)")
    ap.add_argument("--read-tile", type=int, default=16,
                    help="config 2: lanes per read tile of the range and find launches (0: the node width; "
                         "16-lane tiles over 32-pair nodes measured fastest)")
    ap.add_argument("--find-node-width", type=int, default=32,
                    help="config 2: node width of the tree the find breakdown runs on (default: the "
                         "measured tree itself)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--config", type=int, default=2, choices=(2, 3, 4, 5, 6, 7),
                    help="2: headline mixed insert/range batches; 3: erase/find batches with epoch "
                         "reclamation and the memory timeline; 4: batched form vs plain; "
                         "5: served form vs plain per-request path; 6: the reclaimer's memory timeline "
                         "(45M keys, then 22.5M inserts + 22.5M ranges of span 16); 7: the reference's "
                         "mixed-call sweeps (one call of 1M..45M ops per fresh tree, harmonic means)")
    ap.add_argument("--served-keys", type=int, default=500_000_000, help="config 5: tree size")
    ap.add_argument("--qps", default="2000,8000,32000,128000", help="config 5: QPS steps")
    ap.add_argument("--requests", type=int, default=2000, help="config 5: requests per QPS step")
    ap.add_argument("--step-ops", type=int, default=1 << 20, help="config 5: max ops per served step")
    ap.add_argument("--batches", type=int, default=4096, help="config 4: number of batches")
    ap.add_argument("--tune-batches", type=int, default=192,
                    help="config 4: batches in the timed tuning group (and in the warm-up group)")
    ap.add_argument("--modes", default="1,2",
                    help="config 4: worker modes to tune (0 one op per tile, 1 runs, 2 sorted)")
    ap.add_argument("--widths", default="16,32,64", help="config 4: node widths to tune")
    ap.add_argument("--tiles", default="16,32,64",
                    help="config 4: read tile widths (lanes) to tune, each at most the node width")
    ap.add_argument("--no-plain", action="store_true", help="config 4: batched form only (profiling)")
    ap.add_argument("--waves", default="8,16,24,32", help="config 4: waves/CU to tune")
    ap.add_argument("--no-extras", action="store_true",
                    help="default run: skip the config-4 and config-5 legs of the breakdown")
    ap.add_argument("--extras-qps", default="8000,32000,128000,512000", help="default run: config-5 QPS levels")
    ap.add_argument("--extras-requests", type=int, default=1000, help="default run: config-5 requests per level")
    ap.add_argument("--c4-node-width", type=int, default=32, help="default run: config-4 node width (the autotuner's pick)")
    ap.add_argument("--c4-waves", type=int, default=16, help="default run: config-4 worker waves per CU")
    ap.add_argument("--c4-mode", type=int, default=2, help="default run: config-4 worker mode")
    ap.add_argument("--c4-read-tile", type=int, default=16,
                    help="default run: config-4 read tile lanes (0: the node width; the autotuner's pick)")
    ap.add_argument("--served-mode", type=int, default=3, help="config 5: worker mode of the served steps (3: by step size)")
    ap.add_argument("--served-node-width", type=int, default=32, help="config 5: node width of the served tree")
    ap.add_argument("--c3-node-width", type=int, default=16, help="config 3: node width")
    ap.add_argument("--c3-read-tile", type=int, default=0, help="config 3: read tile lanes (0: the node width)")
    ap.add_argument("--served-read-tile", type=int, default=16,
                    help="config 5: read tile lanes of the served tree (0: the node width)")
    ap.add_argument("--c6-keys", type=int, default=45_000_000,
                    help="config 6: initial tree size (half as many inserts and as many ranges follow)")
    ap.add_argument("--c6-calls", type=int, default=440, help="config 6: calls (one reclamation epoch each)")
    ap.add_argument("--c7-sizes", default="all",
                    help="config 7: call sizes in M ops, comma-separated, or 'all' (1..45)")
    ap.add_argument("--c7-trees", default="versioned,plain", help="config 7: tree kinds")
    ap.add_argument("--extras-sweep-sizes", default="1,10,20,30,45",
                    help="default run: config-7 call sizes (M ops) for the sweep harmonic means")
    ap.add_argument("--c7-node-width", type=int, default=32, help="config 7: node width of the versioned tree")
    ap.add_argument("--c7-plain-node-width", type=int, default=16,
                    help="config 7: node width of the plain tree (16 pairs: its faster width for the mixed calls)")
    ap.add_argument("--profile-copy", action="store_true",
                    help="only run 1 GiB device copies (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.profile_copy:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the tree is a single-device structure here "
              "(multi-GPU key-range sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.config == 3:
        print(json.dumps({"config3": run_config3(args)}), flush=True)
        return
    if args.config == 4:
        print(json.dumps({"config4": run_config4(args)}), flush=True)
        return
    if args.config == 5:
        print(json.dumps({"config5": run_config5(args)}), flush=True)
        return
    if args.config == 6:
        print(json.dumps({"config6": run_config6(args)}), flush=True)
        return
    if args.config == 7:
        print(json.dumps({"config7": run_config7(args)}), flush=True)
        return
    result = run_config2(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
