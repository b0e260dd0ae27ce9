// Persistent work-queue form: many batches of mixed operations in one launch.
//
// Semantics (identical to issuing the batches one by one through the plain
// per-batch API with a snapshot before each): batch b reads (find, range) at
// snapshot clock0 + b, i.e. it sees exactly the writes of batches < b; its
// writes (insert, erase) are stamped clock0 + b + 1. The launch reserves the
// nb snapshot ids by advancing the tree clock by nb.
//
// Scheduling: tiles claim operations one at a time from one queue, in order
// (see the worker). An operation of batch b may run once every write of batches < b has
// completed (a device frontier F = last batch whose writes are all done). A
// write of batch b must also wait for that, because leaf version chains are
// stamp-ordered: a late write of batch b-1 must never follow a copy made by
// batch b. Reads of batch b then overlap writes of batch b and later work.
// The lowest unexecuted operation is always runnable, so the queue drains.
// A tile that must wait polls F once per pass of its outer loop and never
// spins inside a branch (the tiles of a wave run in lockstep).
//
// Reference counterpart: the ratio-partitioned mixed kernels
// concurrent_insert_range_kernel / concurrent_find_erase_kernel_versioned
// (btree_kernels.hpp:31-306), which take per-tile snapshots inside the
// kernel and are therefore not deterministic.
#pragma once
#include "mvbt_tree.hpp"

namespace mvbt {

enum : uint32_t { kOpFind = 0, kOpInsert = 1, kOpErase = 2, kOpRange = 3 };

// One operation: {type, key or range lower bound, value or range upper
// bound, output slot of a range (pairs at pairs + slot * L)}.
struct op_record {
  uint32_t type, key, arg, slot;
};

struct alignas(128) batch_job {
  // per-call parameters (written by batch_setup_kernel)
  const op_record* ops;
  unsigned long long nops;  // one past the last op index (= offsets[nb])
  const unsigned long long* offsets;  // nb + 1 absolute op indices into ops
  uint32_t nb;
  uint32_t L;
  uint32_t* results;     // one word per op
  uint64_t* pairs;       // range output
  unsigned long long pair_slots;  // capacity of `pairs` in slots of L pairs
  uint32_t* base_out;    // optional: receives clock0
  uint32_t clock0;
  uint32_t pad0;
  // library-owned scratch, capacity batches each
  uint32_t* nwrites;
  uint32_t* done;
  int32_t* advance;
  unsigned long long pad1[3];
  // hot runtime words, one line each
  unsigned long long next_op;  // next unclaimed operation index
  unsigned long long pad2[15];
  // kFrontierCopies replicas of the frontier, one line each: waiting waves
  // poll the replica of their wave id, so polling spreads over L2 channels.
  // Replicas only grow (atomic max), so no replica ever moves backwards.
  int32_t frontier[8 * 32];
};
constexpr uint32_t kFrontierCopies = 8;

struct batch_params {
  const op_record* ops;
  unsigned long long nops;
  const unsigned long long* offsets;
  uint32_t nb;
  uint32_t L;
  uint32_t* results;
  uint64_t* pairs;
  unsigned long long pair_slots;
  uint32_t* base_out;
};

// 1. Parameters in, clock range reserved, queue reset.
__global__ void batch_setup_kernel(batch_job* job, tree_state* st, batch_params p) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  job->ops = p.ops;
  job->offsets = p.offsets;
  job->nb = p.nb;
  job->L = p.L;
  job->results = p.results;
  job->pairs = p.pairs;
  job->pair_slots = p.pair_slots;
  job->base_out = p.base_out;
  job->clock0 = __hip_atomic_fetch_add(&st->clock, p.nb, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  // Operation indices are absolute: batch b covers ops [offsets[b], offsets[b+1]),
  // results go to the same indices. The caller checked offsets[nb] - offsets[0]
  // equals the op count it passed.
  job->next_op = p.offsets[0];
  job->nops = p.offsets[p.nb];
  for (uint32_t c = 0; c < kFrontierCopies; ++c) job->frontier[c * 32] = -1;
}

// 2. Writes per batch (one block per batch, grid-stride over batches).
__global__ void batch_count_kernel(batch_job* job) {
  __shared__ uint32_t partial[4];
  const uint32_t nb = job->nb;
  for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const unsigned long long lo = job->offsets[b], hi = job->offsets[b + 1];
    uint32_t c = 0;
    for (unsigned long long i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      uint32_t ty = job->ops[i].type;
      c += (ty == kOpInsert || ty == kOpErase) ? 1u : 0u;
    }
    // wave reduction via ballot-free shuffles
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) partial[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
      for (uint32_t w = 0; w < (blockDim.x + 63) / 64; ++w) tot += partial[w];
      job->nwrites[b] = tot;
      job->done[b] = 0;
    }
    __syncthreads();
  }
}

// 3. advance[b] = frontier value once batch b's writes are done: the last
// batch of the run of write-free batches that follows b.
__global__ void batch_frontier_kernel(batch_job* job) {
  const uint32_t nb = job->nb;
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
    uint32_t j = b + 1;
    while (j < nb && job->nwrites[j] == 0) ++j;
    job->advance[b] = (int32_t)j - 1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint32_t j = 0;
    while (j < nb && job->nwrites[j] == 0) ++j;
    for (uint32_t c = 0; c < kFrontierCopies; ++c) job->frontier[c * 32] = (int32_t)j - 1;
  }
}

// 5. Hands clock0 back.
__global__ void batch_finish_kernel(batch_job* job) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && job->base_out) *job->base_out = job->clock0;
}

// Batch of op i for a caller whose previous op was in batch `from` (op
// indices only grow for a tile): gallop forward from `from`, then bisect, so
// the usual case (same or next batch) costs one or two loads instead of
// log2(nb).
MVBT_DEV uint32_t batch_after(const unsigned long long* offsets, uint32_t nb, unsigned long long i, uint32_t from) {
  uint32_t lo = from, step = 1;  // offsets[lo] <= i
  uint32_t hi = lo + step;
  while (hi < nb && offsets[hi] <= i) {
    lo = hi;
    step <<= 1;
    hi = lo + step;
  }
  if (hi > nb) hi = nb;  // offsets[lo] <= i < offsets[hi]
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

MVBT_DEV uint32_t batch_of(const unsigned long long* offsets, uint32_t nb, unsigned long long i) {
  uint32_t lo = 0, hi = nb;  // offsets[lo] <= i < offsets[hi] (i in [offsets[0], offsets[nb]))
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// 4. The persistent worker. Each tile holds one operation at a time, so a
// batch spreads over as many tiles as it has operations and a batch's writes
// finish after about one write latency. The idle tiles of a wave claim the
// next operations of the single queue together, with one atomic (ranked by
// tile order). Claims are in operation order, so the lowest unexecuted
// operation is always held by a running tile and runnable: the queue drains
// without deadlock however many of the launch's blocks are resident (a block
// claims only while it runs). Per-wave home queues would lose this: ops of a
// queue whose waves are not resident could block every resident tile.
template <int NW>
__global__ void __launch_bounds__(256) batch_worker_kernel(tree_ptrs tp, batch_job* job) {
  const tree_view<NW, true> tv(tp);
  using node_t = node<NW, true>;
  tile<NW> t;
  const unsigned long long nops = job->nops;
  const uint32_t nb = job->nb, L = job->L, clock0 = job->clock0;
  const op_record* ops = job->ops;
  uint32_t* results = job->results;
  uint64_t* pairs = job->pairs;
  const unsigned long long pair_slots = job->pair_slots;
  const uint32_t lane = __lane_id();
  const uint32_t my_tile = lane / NW;  // tile index within the wave
  const uint32_t wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int32_t* const my_frontier = job->frontier + (wave % kFrontierCopies) * 32;
  uint32_t backoff = 1;  // s_sleep units while the whole wave waits
  write_ctx ctx = tv.begin_writes(t);
  int32_t seen = -1;      // last frontier value observed (tile-uniform)
  bool have = false;      // tile holds an operation (tile-uniform)
  unsigned long long idx = 0;
  op_record op{};
  uint32_t b = 0;
  uint32_t waits = 0;     // consecutive passes without progress
  for (;;) {
    // Idle tiles of this wave claim operations together.
    const uint64_t idle = __builtin_amdgcn_ballot_w64(!have && t.leader());
    if (idle) {
      const uint32_t n_idle = (uint32_t)__builtin_popcountll(idle);
      const uint32_t first = (uint32_t)__builtin_ctzll(idle);
      unsigned long long base = 0;
      if (lane == first) base = atomic_add64(&job->next_op, (unsigned long long)n_idle);
      base = ((unsigned long long)__builtin_amdgcn_readlane((int)(uint32_t)(base >> 32), (int)first) << 32) |
             (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)base, (int)first);
      if (!have) {
        // rank of this tile among the idle tiles (tile leaders are lanes my_tile*NW)
        const uint64_t below = (my_tile == 0) ? 0ull : (idle & ((1ull << (my_tile * NW)) - 1));
        idx = base + (unsigned long long)__builtin_popcountll(below);
        if (idx >= nops) break;  // queue drained
        op = ops[idx];
        b = batch_after(job->offsets, nb, idx, b);
        have = true;
      }
    }
    // Tiles whose batch is not yet open refresh their view of the frontier:
    // one load per wave (the first such tile leader's), shared by readlane.
    const bool blocked = (int32_t)b - 1 > seen;
    const uint64_t need = __builtin_amdgcn_ballot_w64(blocked && t.leader());
    if (need) {
      const uint32_t poller = (uint32_t)__builtin_ctzll(need);
      int32_t f = 0;
      if (lane == poller) f = __hip_atomic_load(my_frontier, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      f = __builtin_amdgcn_readlane(f, (int)poller);
      if (f > seen) seen = f;
    }
    if ((int32_t)b - 1 > seen) {
      if (++waits > (1u << 21)) {  // seconds: the frontier never moved
        tv.flag(kErrCorrupt, t);
        break;
      }
      // back off only when no tile of the wave can work
      if (__builtin_amdgcn_ballot_w64(!((int32_t)b - 1 > seen)) == 0) {
        for (uint32_t i = 0; i < backoff; ++i) __builtin_amdgcn_s_sleep(8);
        if (backoff < 16) backoff *= 2;
      }
      continue;
    }
    waits = 0;
    backoff = 1;
    // Every operation starts with a root-to-leaf walk, so the tiles of a wave
    // take it together whatever their operation types: their node loads
    // overlap instead of running type by type. Only the short leaf-side parts
    // below diverge.
    uint32_t rkey = op.key;
    if (op.type == kOpRange && rkey == 0) rkey = 1;
    const bool route = valid_key(rkey) && !(op.type == kOpRange && op.arg <= rkey);
    node_t n;
    uint32_t leaf = kNone, par = kNone;
    if (route) leaf = tv.template descend<true, true>(rkey, n, t, &par);
    uint32_t r;
    if (op.type == kOpFind) {
      r = route ? tv.template find_in<true>(op.key, clock0 + b, n, leaf, t) : kEmpty;
    } else if (op.type == kOpRange) {
      uint64_t* out = (pairs && op.slot < pair_slots) ? pairs + (size_t)op.slot * L : nullptr;
      if (pairs && op.slot >= pair_slots) tv.flag(kErrCorrupt, t);
      r = route ? tv.template range_in<true>(rkey, op.arg, clock0 + b, n, leaf, out, L, t) : 0u;
      if (r > L) tv.flag(kErrRangeOverflow, t);
    } else {
      ctx.cur = clock0 + b + 1;
      ctx.want = NW;  // refills grow with demand (write_ctx::grow)
      r = (op.type == kOpInsert) ? tv.template write<0>(op.key, op.arg, ctx, t, leaf, par)
                                 : tv.template write<1>(op.key, 0u, ctx, t, leaf, par);
      order_release();  // this write's stores complete before it is counted
      if (t.leader()) {
        uint32_t prev = atomic_add32(job->done + b, 1u);
        if (prev + 1 == job->nwrites[b]) {
          const int32_t adv = job->advance[b];
          for (uint32_t c = 0; c < kFrontierCopies; ++c)
            __hip_atomic_fetch_max(job->frontier + c * 32, adv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    if (t.leader()) results[idx] = r;
    have = false;
  }
  tv.finish(ctx, t);
}

// 4'. The chunked worker. A tile claims a run of C consecutive operations
// (lane j holds operation start + j) and works through it batch segment by
// batch segment, each segment once its batch is open (the frontier rule of
// the single-op worker). A segment uses the chunk engines of the per-batch
// kernels: all its keys are routed together (descend_chunk), its inserts and
// erases go through the batched leaf phase (write_chunk), its reads start
// from their routed leaves, and its writes are counted with one atomic.
// C follows the size of the batch the wave claimed from last: batch size /
// tiles in the grid, clamped to [1, W], so small batches still spread one
// operation per tile and large ones are taken W at a time. Claims stay in
// operation order and each tile finishes its segments in order, so the
// lowest unfinished operation is always held by a running tile whose earlier
// segments are done: the queue drains as in the single-op worker.
template <int NW>
__global__ void __launch_bounds__(256) batch_chunk_worker_kernel(tree_ptrs tp, batch_job* job) {
  const tree_view<NW, true> tv(tp);
  using node_t = node<NW, true>;
  tile<NW> t;
  const unsigned long long nops = job->nops;
  const uint32_t nb = job->nb, L = job->L, clock0 = job->clock0;
  const op_record* ops = job->ops;
  const unsigned long long* offsets = job->offsets;
  uint32_t* results = job->results;
  uint64_t* pairs = job->pairs;
  const unsigned long long pair_slots = job->pair_slots;
  const uint32_t lane = __lane_id(), r = t.rank();
  const uint32_t my_tile = lane / NW;
  const uint32_t wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int32_t* const my_frontier = job->frontier + (wave % kFrontierCopies) * 32;
  const unsigned long long tiles_total = (unsigned long long)gridDim.x * (blockDim.x / NW);
  uint32_t backoff = 1;
  // The top levels in LDS (as in write_kernel, at 32-pair nodes only), copied
  // once per block of the persistent launch: used when the job gives every
  // tile several operations (the copy delays the block's first one) and the
  // tree's nodes can hold the whole job (its splits then leave the images
  // about one right move stale at most).
  __shared__ route_cache<NW, 16384> rc;
  if constexpr (NW == 32)
    tv.build_route_cache(rc, nops >= 4 * tiles_total ? nops / node_t::kData + 1 : ~0ull);
  write_ctx ctx = tv.begin_writes(t);
  ctx.exact = 1;
  int32_t seen = -1;
  bool have = false;                       // tile holds an unfinished chunk (tile-uniform)
  unsigned long long cstart = 0, seg = 0, end = 0;  // chunk start, next op, chunk end
  uint32_t b = 0;                          // batch of op `seg`
  uint32_t wave_b = 0;                     // batch of the wave's last claim (chunk sizing)
  op_record op{};                          // lane j: operation cstart + j
  uint32_t waits = 0;
  for (;;) {
    const uint64_t idle = __builtin_amdgcn_ballot_w64(!have && t.leader());
    if (idle) {
      const uint32_t n_idle = (uint32_t)__builtin_popcountll(idle);
      const uint32_t first = (uint32_t)__builtin_ctzll(idle);
      const unsigned long long bsz = offsets[wave_b + 1] - offsets[wave_b];
      const unsigned long long cq = bsz / tiles_total;
      const uint32_t C = cq < 1 ? 1u : (cq > (unsigned long long)NW ? (uint32_t)NW : (uint32_t)cq);
      unsigned long long base = 0;
      if (lane == first) base = atomic_add64(&job->next_op, (unsigned long long)n_idle * C);
      base = ((unsigned long long)__builtin_amdgcn_readlane((int)(uint32_t)(base >> 32), (int)first) << 32) |
             (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)base, (int)first);
      if (base < nops) wave_b = batch_after(offsets, nb, base, wave_b);
      if (!have) {
        const uint64_t below = (my_tile == 0) ? 0ull : (idle & ((1ull << (my_tile * NW)) - 1));
        cstart = base + (unsigned long long)__builtin_popcountll(below) * C;
        if (cstart >= nops) break;  // queue drained
        end = cstart + C < nops ? cstart + C : nops;
        seg = cstart;
        if (cstart + r < end) op = ops[cstart + r];
        b = batch_after(offsets, nb, seg, b);
        have = true;
      }
    }
    // the segment's batch must be open (all writes of earlier batches done)
    const bool blocked = (int32_t)b - 1 > seen;
    const uint64_t need = __builtin_amdgcn_ballot_w64(blocked && t.leader());
    if (need) {
      const uint32_t poller = (uint32_t)__builtin_ctzll(need);
      int32_t f = 0;
      if (lane == poller) f = __hip_atomic_load(my_frontier, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      f = __builtin_amdgcn_readlane(f, (int)poller);
      if (f > seen) seen = f;
    }
    if ((int32_t)b - 1 > seen) {
      if (++waits > (1u << 21)) {  // seconds: the frontier never moved
        tv.flag(kErrCorrupt, t);
        break;
      }
      if (__builtin_amdgcn_ballot_w64(!((int32_t)b - 1 > seen)) == 0) {
        for (uint32_t i = 0; i < backoff; ++i) __builtin_amdgcn_s_sleep(8);
        if (backoff < 16) backoff *= 2;
      }
      continue;
    }
    waits = 0;
    backoff = 1;
    // one segment: the chunk's operations of batch b
    const unsigned long long seg_end = end < offsets[b + 1] ? end : offsets[b + 1];
    const bool mine = cstart + r >= seg && cstart + r < seg_end;
    const bool is_w = mine && (op.type == kOpInsert || op.type == kOpErase);
    const bool is_rd = mine && !is_w;
    uint32_t rkey = op.key;
    if (op.type == kOpRange && rkey == 0) rkey = 1;
    const bool route = mine && valid_key(rkey) && !(op.type == kOpRange && op.arg <= rkey);
    uint32_t leaf = kNone, par = kNone, res = 0;
    uint32_t st0 = kNone, sp0 = kNone;
    if constexpr (NW == 32) {
      if (route) st0 = tv.route_lane(rc, rkey, sp0);
    }
    tv.template descend_chunk<true>(rkey, t.ballot(route), leaf, par, t, t.ballot(is_w), st0, sp0);
    // writes: batched leaf phase, then the keys it hands back one by one
    ctx.cur = clock0 + b + 1;
    const uint64_t qw = t.ballot(is_w);
    for (int w = 0; w < 2; ++w) {
      uint64_t q = t.ballot(is_w && op.type == (w == 0 ? kOpInsert : kOpErase));
      if (!q) continue;
      ctx.want = (uint32_t)__builtin_popcountll(q);
      q = (w == 0) ? tv.template write_chunk<0>(op.key, op.arg, q, leaf, res, ctx, t)
                   : tv.template write_chunk<1>(op.key, 0u, q, leaf, res, ctx, t);
      while (q) {
        const uint32_t j = (uint32_t)lowest_bit(q);
        ctx.want = (uint32_t)__builtin_popcountll(q);
        q &= q - 1;
        const uint32_t k = t.shfl(op.key, j);
        const uint32_t v = (w == 0) ? tv.template write<0>(k, t.shfl(op.arg, j), ctx, t, t.shfl(leaf, j), t.shfl(par, j))
                                    : tv.template write<1>(k, 0u, ctx, t, t.shfl(leaf, j), t.shfl(par, j));
        if (r == j) res = v;
      }
    }
    // reads at the batch's snapshot, from their routed leaves: finds together,
    // ranges one by one
    const uint64_t routed = t.ballot(route);
    {
      const uint64_t qf = t.ballot(is_rd && op.type == kOpFind);
      if (qf) {
        const uint32_t v = tv.find_chunk_conc(op.key, leaf, clock0 + b, qf & routed, t);
        if ((qf >> r) & 1) res = v;
      }
    }
    uint64_t qr = t.ballot(is_rd && op.type != kOpFind);
    while (qr) {
      const uint32_t j = (uint32_t)lowest_bit(qr);
      qr &= qr - 1;
      const uint32_t ty = t.shfl(op.type, j), k = t.shfl(op.key, j), a = t.shfl(op.arg, j);
      const uint32_t lf = t.shfl(leaf, j);
      const bool ok = (routed >> j) & 1;
      uint32_t v;
      if (ty == kOpFind) {
        v = kEmpty;
        if (ok) {
          node_t n;
          n.load(tv.at(lf), t);
          v = tv.template find_in<true>(k, clock0 + b, n, lf, t);
        }
      } else {
        v = 0;
        const uint32_t sl = t.shfl(op.slot, j);
        uint64_t* out = (pairs && sl < pair_slots) ? pairs + (size_t)sl * L : nullptr;
        if (pairs && sl >= pair_slots) tv.flag(kErrCorrupt, t);
        if (ok) {
          node_t n;
          n.load(tv.at(lf), t);
          v = tv.template range_in<true>(k == 0 ? 1u : k, a, clock0 + b, n, lf, out, L, t);
        }
        if (v > L) tv.flag(kErrRangeOverflow, t);
      }
      if (r == j) res = v;
    }
    if (mine) results[cstart + r] = res;
    // count the segment's writes: one atomic; the batch's last one opens the next
    if (qw) {
      order_release();  // the segment's stores complete before they are counted
      if (t.leader()) {
        const uint32_t cnt = (uint32_t)__builtin_popcountll(qw);
        const uint32_t prev = atomic_add32(job->done + b, cnt);
        if (prev + cnt == job->nwrites[b]) {
          const int32_t adv = job->advance[b];
          for (uint32_t c = 0; c < kFrontierCopies; ++c)
            __hip_atomic_fetch_max(job->frontier + c * 32, adv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    seg = seg_end;
    if (seg >= end) have = false;
    else b = batch_after(offsets, nb, seg, b);
  }
  tv.finish(ctx, t);
}

}  // namespace mvbt
