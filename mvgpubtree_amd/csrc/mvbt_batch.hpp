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
// A batch's writes must name distinct keys: two writes of one key inside a
// batch would make the result depend on which tile runs first. The count
// pass checks this with a hash table over (batch, key) of the call's write
// operations and rejects the whole call (kErrBadBatch; nothing is applied,
// results are left as they were) when a key repeats, an operation type is
// unknown, or the offsets disagree with the operation count.
//
// Reference counterpart: the ratio-partitioned mixed kernels
// concurrent_insert_range_kernel / concurrent_find_erase_kernel_versioned
// (btree_kernels.hpp:31-306), which take per-tile snapshots inside the
// kernel and are therefore not deterministic.
#pragma once
#include "mvbt_kernels.hpp"
#include "mvbt_sort.hpp"
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
  uint32_t hash_mask;    // conflict table size - 1: a power of two >= 2 * writes (batch_mask_kernel)
  unsigned long long first_op;  // offsets[0]
  // library-owned scratch: capacity batches each, and the conflict table
  uint32_t* nwrites;
  uint32_t* done;
  int32_t* advance;
  uint32_t* hash;        // hash_mask + 1 entries: 0 empty, else 1 + op index - first_op
  uint32_t* op_batch;    // batch of op first_op + i, written by the count pass
  uint32_t bad;          // nonzero: the call was rejected and runs nothing
  uint32_t pad1[1];
  // hot runtime words, one line each
  unsigned long long next_op;  // next unclaimed operation index
  unsigned long long pad2[15];
  // kFrontierCopies replicas of the frontier, one line each: waiting waves
  // poll the replica of their wave id, so polling spreads over L2 channels.
  // Replicas only grow (atomic max), so no replica ever moves backwards.
  int32_t frontier[8 * 32];
  // Sorted form (worker mode 2): library-owned scratch, one entry per
  // operation unless stated.
  uint32_t* w_rel;       // write operations (index - first_op), in operation order
  uint32_t* w_par;       // parent hint of each write (indexed like w_rel)
  uint32_t* sort_k[2];   // leaf hints (sort keys), two buffers
  uint32_t* sort_v[2];   // positions in w_rel (sort payloads), two buffers
  uint32_t* rd_find;     // find operations (index - first_op), any order
  uint32_t* rd_range;    // range operations (index - first_op), any order
  uint32_t* seg;         // per class (writes, finds, ranges) kSortSegments segment counts
  uint32_t* hist;        // kRadix * kSortSegments histogram + kRadix digit totals
  // the sorted writes' fields in sorted order (batch_gather_kernel): key,
  // value, index - first_op (bit 31: erase), batch, parent hint, copy slot
  uint32_t* gat[6];
  // counts of the current call (zeroed by batch_setup_kernel), one line
  uint32_t n_writes;
  uint32_t n_finds;
  uint32_t n_ranges;
  uint32_t n_epochs;     // (leaf, batch) groups of the sorted writes: one planned copy slot each
  uint32_t cp_valid;     // planned slots that exist (the rest: pool exhausted)
  uint32_t cp_got;       // of them, taken from the free stack
  uint32_t total_writes;  // write operations of the call (sizes the conflict table)
  uint32_t hash_cap_mask;  // the conflict table's capacity - 1
  long long cp_top;                 // free-stack top before the reservation
  unsigned long long cp_bump;       // first never-used index of the reservation
  unsigned long long cp_ring;       // first retire-ring position of the reservation
  // windows of the sorted form: the call's batches in `windows` consecutive
  // groups of about equal operation count; [w_lo, w_hi) = the current
  // window's operations (index - first_op)
  uint32_t windows;
  uint32_t reclaim;
  uint32_t w_lo;
  uint32_t w_hi;
  unsigned long long pad4[7];
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
  uint32_t hash_mask;
  uint32_t windows;  // sorted form: windows of batches the call runs in (>= 1)
  uint32_t reclaim;  // sorted form: free every retired copy after each window
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
  job->hash_cap_mask = p.hash_mask;
  job->hash_mask = 1023u;
  job->total_writes = 0;
  // Operation indices are absolute: batch b covers ops [offsets[b], offsets[b+1]),
  // results go to the same indices; offsets[nb] - offsets[0] must equal the op
  // count the caller passed (the size its buffers were checked against).
  const unsigned long long first = p.offsets[0], last = p.offsets[p.nb];
  const bool ok = last >= first && last - first == p.nops;
  job->first_op = first;
  job->next_op = first;
  job->nops = ok ? last : first;
  job->bad = ok ? 0u : 1u;
  job->n_writes = 0;
  job->n_finds = 0;
  job->n_ranges = 0;
  job->windows = p.windows > 0 ? p.windows : 1u;
  job->reclaim = p.reclaim;
  job->w_lo = 0;
  job->w_hi = ok ? (uint32_t)(last - first) : 0u;
  if (!ok) atomic_or32(&st->error_flags, kErrBadBatch);
  for (uint32_t c = 0; c < kFrontierCopies; ++c) job->frontier[c * 32] = -1;
}

// 0. Points a job descriptor at its scratch arrays from the device side (the
// allocation-free entry point: the job and its scratch all lie in one static
// device array, so no host copy is needed).
struct scratch_layout {
  uint32_t* nwrites;
  uint32_t* done;
  int32_t* advance;
  uint32_t* hash;
  uint32_t* op_batch;
  uint32_t* sorted[16];  // w_rel, w_par, sort_k[0..1], sort_v[0..1], rd_find, rd_range, seg, hist, gat[0..5]
};
__global__ void batch_bind_kernel(batch_job* job, scratch_layout l) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  job->nwrites = l.nwrites;
  job->done = l.done;
  job->advance = l.advance;
  job->hash = l.hash;
  job->op_batch = l.op_batch;
  job->w_rel = l.sorted[0];
  job->w_par = l.sorted[1];
  job->sort_k[0] = l.sorted[2];
  job->sort_k[1] = l.sorted[3];
  job->sort_v[0] = l.sorted[4];
  job->sort_v[1] = l.sorted[5];
  job->rd_find = l.sorted[6];
  job->rd_range = l.sorted[7];
  job->seg = l.sorted[8];
  job->hist = l.sorted[9];
  for (int k = 0; k < 6; ++k) job->gat[k] = l.sorted[10 + k];
}

// 1'. Clears the conflict table (a fill kernel, so the launch graph holds no
// memset node).
__global__ void batch_clear_kernel(batch_job* job) {
  const uint32_t n = job->hash_mask + 1;
  uint32_t* h = job->hash;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) h[i] = 0;
}

MVBT_DEV uint32_t batch_of(const unsigned long long* offsets, uint32_t nb, unsigned long long i);

MVBT_DEV uint32_t conflict_hash(uint32_t key, uint32_t b) {
  uint32_t h = key * 0x9E3779B1u ^ (b * 0x7FEB352Du + 0x165667B1u);
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}

// Records write op i (batch b, valid key) in the conflict table; false when
// batch b already writes `key`. Linear probing; the table has at least twice
// as many entries as the call has operations, so a free entry always exists.
MVBT_DEV bool claim_write(const batch_job* job, unsigned long long i, uint32_t b, uint32_t key) {
  const uint32_t mask = job->hash_mask;
  uint32_t* table = job->hash;
  const uint32_t me = (uint32_t)(i - job->first_op) + 1u;
  uint32_t h = conflict_hash(key, b) & mask;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    uint32_t cur = ld32_coherent(table + h);
    if (cur == 0) {
      cur = atomic_cas32(table + h, 0u, me);
      if (cur == 0) return true;
    }
    const unsigned long long j = job->first_op + cur - 1u;
    if (job->ops[j].key == key && job->op_batch[cur - 1u] == b) return false;
    h = (h + 1) & mask;
  }
  return true;
}

// 2. Writes per batch (one block per batch, grid-stride over batches), each
// operation's batch, the call's write count, and the first checks: known op
// types, ordered offsets.
__global__ void batch_count_kernel(batch_job* job, tree_state* st) {
  __shared__ uint32_t partial[4];
  const uint32_t nb = job->nb;
  if (ld32_coherent(&job->bad)) return;  // offsets disagree with the op count
  for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const unsigned long long lo = job->offsets[b], hi = job->offsets[b + 1];
    uint32_t c = 0;
    bool bad = hi < lo;
    for (unsigned long long i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      const op_record o = job->ops[i];
      const bool w = o.type == kOpInsert || o.type == kOpErase;
      c += w ? 1u : 0u;
      job->op_batch[i - job->first_op] = b;
      if (o.type > kOpRange) bad = true;
    }
    if (bad) {
      atomic_or32(&job->bad, 1u);
      atomic_or32(&st->error_flags, kErrBadBatch);
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
      if (tot) atomic_add32(&job->total_writes, tot);
    }
    __syncthreads();
  }
}

// 2'. The conflict table's size: the smallest power of two >= twice the
// call's writes (at least 1024, at most its capacity), so a table that fits
// the caches serves a call whose operations are mostly reads.
__global__ void batch_mask_kernel(batch_job* job) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long e = 1024;
  while (e < 2ull * job->total_writes) e *= 2;
  const unsigned long long cap = (unsigned long long)job->hash_cap_mask + 1;
  job->hash_mask = (uint32_t)((e < cap ? e : cap) - 1);
}

// 2''. (after batch_clear_kernel) Distinct write keys per batch: every write
// of a valid key claims (batch, key) in the conflict table; a second claim
// rejects the call.
__global__ void batch_conflict_kernel(batch_job* job, tree_state* st) {
  if (ld32_coherent(&job->bad)) return;
  const unsigned long long first = job->first_op, n = job->nops - first;
  bool bad = false;
  for (unsigned long long r = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; r < n;
       r += (unsigned long long)gridDim.x * blockDim.x) {
    const op_record o = job->ops[first + r];
    if ((o.type == kOpInsert || o.type == kOpErase) && valid_key(o.key) &&
        !claim_write(job, first + r, job->op_batch[r], o.key))
      bad = true;
  }
  if (bad) {
    atomic_or32(&job->bad, 1u);
    atomic_or32(&st->error_flags, kErrBadBatch);
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

// Batch of op i (offsets[0] <= i < offsets[nb]), by bisection.
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
  if (ld32_coherent(&job->bad)) return;  // rejected call: nothing runs
  const tree_view<NW, true> tv(tp);
  using node_t = node<NW, true>;
  tile<NW> t;
  const unsigned long long nops = job->nops;
  const uint32_t L = job->L, clock0 = job->clock0;
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
  ctx.track = 0;          // stamps are fixed per batch: clock0 + b + 1
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
        b = job->op_batch[idx - job->first_op];
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
      r = route ? tv.template find_in<true, false>(op.key, clock0 + b, n, leaf, t) : kEmpty;
    } else if (op.type == kOpRange) {
      uint64_t* out = (pairs && op.slot < pair_slots) ? pairs + (size_t)op.slot * L : nullptr;
      if (pairs && op.slot >= pair_slots) tv.flag(kErrCorrupt, t);
      r = route ? tv.template range_in<true, false>(rkey, op.arg, clock0 + b, n, leaf, out, L, t) : 0u;
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
// (lane j holds operation start + j) and routes all of them at once, before
// their batches open: every key stops one level above the leaves and takes
// its leaf from the parent's entry (routing reads only the latest internal
// nodes, which need no frontier). The run is then worked through batch
// segment by batch segment, each once its batch is open (the frontier rule of
// the single-op worker), with only the leaf side left to do: inserts and
// erases go through the batched leaf phase (write_chunk) from the leaf hints,
// finds run a pass engine that starts at the hints and moves right or down
// as the tree has changed since (find_chunk_from), ranges start there one at
// a time; the segment's writes are counted with one atomic. So the walk from
// the root, about half of an operation's dependent loads, overlaps the wait
// for earlier batches instead of following it.
// C follows the size of the batch the wave claimed from last: batch size /
// tiles in the grid, clamped to [1, W], so small batches still spread one
// operation per tile and large ones are taken W at a time. Claims stay in
// operation order and each tile finishes its segments in order, so the
// lowest unfinished operation is always held by a running tile whose earlier
// segments are done: the queue drains as in the single-op worker.
template <int NW>
__global__ void __launch_bounds__(256) batch_chunk_worker_kernel(tree_ptrs tp, batch_job* job) {
  if (ld32_coherent(&job->bad)) return;  // rejected call: nothing runs
  const tree_view<NW, true> tv(tp);
  using node_t = node<NW, true>;
  tile<NW> t;
  const unsigned long long nops = job->nops, first_op = job->first_op;
  const uint32_t L = job->L, clock0 = job->clock0;
  const op_record* ops = job->ops;
  const unsigned long long* offsets = job->offsets;
  const uint32_t* op_batch = job->op_batch;
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
    tv.build_route_cache(rc, nops - first_op >= 4 * tiles_total ? (nops - first_op) / node_t::kData + 1 : ~0ull);
  write_ctx ctx = tv.begin_writes(t);
  ctx.exact = 1;
  ctx.track = 0;  // stamps are fixed per batch: clock0 + b + 1
  int32_t seen = -1;
  bool have = false;          // tile holds an unfinished run (tile-uniform)
  uint64_t pending = 0;       // lanes of the run not yet done (tile-uniform)
  unsigned long long cstart = 0;
  uint32_t wave_b = 0;        // batch of the wave's last claim (run sizing)
  op_record op{};             // lane j: operation cstart + j
  uint32_t ob = 0;            // lane j: its batch
  uint32_t hint = kNone, hint_par = kNone;  // lane j: leaf hint and its parent
  bool routed = false;        // lane j: its key was routed
  uint32_t waits = 0;
  for (;;) {
    const uint64_t idle = __builtin_amdgcn_ballot_w64(!have && t.leader());
    bool fresh = false;  // tile-uniform: claimed a run in this pass
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
      if (!have) {
        const uint64_t below = (my_tile == 0) ? 0ull : (idle & ((1ull << (my_tile * NW)) - 1));
        cstart = base + (unsigned long long)__builtin_popcountll(below) * C;
        if (cstart >= nops) break;  // queue drained
        const unsigned long long end = cstart + C < nops ? cstart + C : nops;
        const bool mine = cstart + r < end;
        if (mine) {
          op = ops[cstart + r];
          ob = op_batch[cstart + r - first_op];
        }
        pending = t.ballot(mine);
        have = true;
        fresh = true;
      }
      // size the wave's next runs from the batch its claim started in
      if (base < nops) wave_b = op_batch[base - first_op];
    }
    // route the operations just claimed, whatever their batch
    if (__builtin_amdgcn_ballot_w64(fresh)) {
      uint32_t rkey = op.key;
      if (op.type == kOpRange && rkey == 0) rkey = 1;
      const bool route = fresh && ((pending >> r) & 1) && valid_key(rkey) && !(op.type == kOpRange && op.arg <= rkey);
      const uint64_t todo = t.ballot(route);
      uint32_t st0 = kNone, sp0 = kNone;
      if constexpr (NW == 32) {
        if (route) st0 = tv.route_lane(rc, rkey, sp0);
      }
      uint32_t lf = kNone, pr = kNone;
      tv.template descend_chunk<true>(rkey, todo, lf, pr, t, todo, st0, sp0);
      if (fresh) {
        hint = lf;
        hint_par = pr;
        routed = route && lf < tv.cap_;
      }
    }
    if (!have) continue;
    // the run's next segment: its pending operations of the lowest pending batch
    const uint32_t b = t.shfl(ob, (uint32_t)lowest_bit(pending));
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
      if (__builtin_amdgcn_ballot_w64(have && !((int32_t)b - 1 > seen)) == 0) {
        for (uint32_t i = 0; i < backoff; ++i) __builtin_amdgcn_s_sleep(8);
        if (backoff < 16) backoff *= 2;
      }
      continue;
    }
    waits = 0;
    backoff = 1;
    const uint32_t ts = clock0 + b;
    const bool mine = ((pending >> r) & 1) && ob == b;
    const uint64_t seg = t.ballot(mine);
    const bool is_w = mine && (op.type == kOpInsert || op.type == kOpErase);
    const bool is_rd = mine && !is_w;
    uint32_t res = 0;
    // writes: batched leaf phase from the hints, then the keys it hands back one by one
    ctx.cur = ts + 1;
    const uint64_t qw = t.ballot(is_w);
    for (int w = 0; w < 2; ++w) {
      uint64_t q = t.ballot(is_w && op.type == (w == 0 ? kOpInsert : kOpErase));
      if (!q) continue;
      ctx.want = (uint32_t)__builtin_popcountll(q);
      q = (w == 0) ? tv.template write_chunk<0>(op.key, op.arg, q, hint, res, ctx, t)
                   : tv.template write_chunk<1>(op.key, 0u, q, hint, res, ctx, t);
      while (q) {
        const uint32_t j = (uint32_t)lowest_bit(q);
        ctx.want = (uint32_t)__builtin_popcountll(q);
        q &= q - 1;
        const uint32_t k = t.shfl(op.key, j);
        const uint32_t lf = t.shfl(hint, j), pr = t.shfl(hint_par, j);
        const uint32_t v = (w == 0) ? tv.template write<0>(k, t.shfl(op.arg, j), ctx, t, lf, pr)
                                    : tv.template write<1>(k, 0u, ctx, t, lf, pr);
        if (r == j) res = v;
      }
    }
    // reads at the batch's snapshot, from the hints: finds together, ranges one by one
    const uint64_t ok = t.ballot(routed);
    {
      const uint64_t qf = t.ballot(is_rd && op.type == kOpFind);
      if (qf) {
        const uint32_t v = tv.find_chunk_from(op.key, hint, ts, qf & ok, t);
        if ((qf >> r) & 1) res = v;
      }
    }
    uint64_t qr = t.ballot(is_rd && op.type == kOpRange);
    while (qr) {
      const uint32_t j = (uint32_t)lowest_bit(qr);
      qr &= qr - 1;
      const uint32_t k = t.shfl(op.key, j), a = t.shfl(op.arg, j), lf = t.shfl(hint, j);
      const uint32_t sl = t.shfl(op.slot, j);
      uint64_t* out = (pairs && sl < pair_slots) ? pairs + (size_t)sl * L : nullptr;
      if (pairs && sl >= pair_slots) tv.flag(kErrCorrupt, t);
      uint32_t v = 0;
      if ((ok >> j) & 1) v = tv.template range<true, false>(k == 0 ? 1u : k, a, ts, out, L, t, lf);
      if (v > L) tv.flag(kErrRangeOverflow, t);
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
    pending &= ~seg;
    if (!pending) have = false;
  }
  tv.finish(ctx, t);
}

// ---- The sorted form (worker mode 2) ------------------------------------------
// The whole stream is known when the call starts, so it need not run batch
// after batch. Because every read names its snapshot and the tree keeps
// versions, a read of batch b gives the same answer after every write of the
// call as before the writes of batch b: it reads the version stamped <=
// clock0 + b. So the call applies all its writes first and then answers all
// its reads in one parallel pass. The writes only have to keep, per leaf,
// the order of their batches (version chains are stamp-ordered); writes to
// different leaves are independent. The pipeline:
//   split  the writes, in operation order, into one list; finds and ranges
//          into two more (any order);
//   route  every write to its leaf (the tree is quiescent: no writer runs);
//   sort   the writes by leaf, stably (mvbt_sort.hpp), so each leaf's
//          writes are contiguous and in operation order, i.e. batch order;
//   write  each tile owns the leaf groups that start in its window of the
//          sorted list (a group that runs past the window stays with the
//          tile that holds its start), and applies each group's writes
//          batch by batch: one round writes, for every group, the writes of
//          its earliest remaining batch through the batched leaf phase with
//          that batch's stamp (write_chunk_at); keys it hands back (split,
//          moved right) go one by one. No other tile writes a group's leaves
//          or the leaves split from them, so no leaf lock is ever contended;
//   read   finds and ranges, each at its batch's snapshot, as plain
//          snapshot reads (no writer runs any more).
// No frontier and no per-batch wait: a call costs a few bulk passes however
// many batches it holds. Results are identical to the per-batch path.

// 5'. Window k of the sorted form: the batches whose first operation lies in
// the k-th of `windows` equal slices of the call's operations (a batch is
// never split). Windows run one after another; a read of a window sees the
// writes of the earlier windows, which have been applied, and of its own
// earlier batches, as in the whole-call form.
MVBT_DEV uint32_t window_start(const batch_job* job, uint32_t k) {
  const uint32_t nb = job->nb, K = job->windows;
  if (k == 0) return 0;
  if (k >= K) return nb;
  const unsigned long long n = job->nops - job->first_op;
  const unsigned long long target = job->first_op + (n * k) / K;
  uint32_t lo = 0, hi = nb;  // the first batch starting at or after target
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (job->offsets[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void batch_window_begin_kernel(batch_job* job, uint32_t k) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (ld32_coherent(&job->bad)) return;
  const unsigned long long first = job->first_op;
  job->w_lo = (uint32_t)(job->offsets[window_start(job, k)] - first);
  job->w_hi = (uint32_t)(job->offsets[window_start(job, k + 1)] - first);
  job->n_writes = 0;
  job->n_finds = 0;
  job->n_ranges = 0;
  job->n_epochs = 0;
}

// End of a window: every read of the call at a snapshot below the next
// window's first batch is done, so - when the call may reclaim (no snapshot
// outside it is live) - every retired copy is garbage: the window's own copies
// hold contents that only those reads could see. Frees them (appending an
// epoch to the memory timeline) and merges parked slots either way; the
// grid part is reclaim_copy_kernel.
__global__ void batch_window_end_kernel(batch_job* job, tree_state* st, epoch_stat* series) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (ld32_coherent(&job->bad)) {
    st->scratch[1] = 0;  // reclaim_copy_kernel then copies nothing
    st->scratch[3] = 0;
    return;
  }
  const bool r = job->reclaim != 0;
  reclaim_prep(st, r ? ~0ull : 0ull, r ? series : nullptr);
}

// Operation classes of the split: writes (insert, erase), finds, ranges.
MVBT_DEV uint32_t op_class(uint32_t ty) {
  return (ty == kOpInsert || ty == kOpErase) ? 0u : ty == kOpFind ? 1u : ty == kOpRange ? 2u : 3u;
}
constexpr uint32_t kOpClasses = 3;

// 6a. Operations of each class per segment (segments as in the sort):
// seg[class * segments + segment]. A rejected call writes zero counts: the
// scan and the radix passes that follow do not look at `bad`, and their
// sizes come from these counts.
__global__ void __launch_bounds__(kSortBlock) batch_split_count_kernel(batch_job* job) {
  __shared__ uint32_t part[kOpClasses][kSortBlock / 64];
  const uint32_t n = ld32_coherent(&job->bad) ? 0u : job->w_hi - job->w_lo;
  const op_record* ops = job->ops + job->first_op + job->w_lo;
  uint32_t a, b;
  segment_of(n, blockIdx.x, gridDim.x, a, b);
  uint32_t c[kOpClasses] = {0, 0, 0};
  for (uint32_t i = a + threadIdx.x; i < b; i += kSortBlock) {
    const uint32_t k = op_class(ops[i].type);
#pragma unroll
    for (uint32_t q = 0; q < kOpClasses; ++q) c[q] += k == q ? 1u : 0u;
  }
#pragma unroll
  for (uint32_t q = 0; q < kOpClasses; ++q) {
    for (int o = 32; o > 0; o >>= 1) c[q] += __shfl_xor(c[q], o, 64);
    if ((threadIdx.x & 63) == 0) part[q][threadIdx.x >> 6] = c[q];
  }
  __syncthreads();
  if (threadIdx.x < kOpClasses) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < kSortBlock / 64; ++w) tot += part[threadIdx.x][w];
    job->seg[threadIdx.x * gridDim.x + blockIdx.x] = tot;
  }
}

// 6b (after radix_scan_kernel turned the counts into offsets and the class
// totals into n_writes, n_finds, n_ranges). Each class goes to its list in
// operation order (a stable scatter per segment; no shared counter): writes
// to w_rel, finds to rd_find, ranges to rd_range.
__global__ void __launch_bounds__(kSortBlock) batch_split_kernel(batch_job* job) {
  constexpr uint32_t kWaves = kSortBlock / 64;
  __shared__ uint32_t wsum[kOpClasses][kWaves];
  if (ld32_coherent(&job->bad)) return;
  const uint32_t w_lo = job->w_lo, n = job->w_hi - w_lo;
  const op_record* ops = job->ops + job->first_op + w_lo;
  uint32_t* const lists[kOpClasses] = {job->w_rel, job->rd_find, job->rd_range};
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6, G = gridDim.x;
  const uint64_t below = (1ull << lane) - 1;
  uint32_t a, b;
  segment_of(n, blockIdx.x, G, a, b);
  uint32_t run[kOpClasses];
#pragma unroll
  for (uint32_t q = 0; q < kOpClasses; ++q) run[q] = job->seg[q * G + blockIdx.x];
  for (uint32_t c = a; c < b; c += kSortBlock) {
    const uint32_t i = c + tid;
    const uint32_t k = i < b ? op_class(ops[i].type) : 3u;
    uint64_t m[kOpClasses];
#pragma unroll
    for (uint32_t q = 0; q < kOpClasses; ++q) {
      m[q] = __builtin_amdgcn_ballot_w64(k == q);
      if (lane == 0) wsum[q][w] = (uint32_t)__builtin_popcountll(m[q]);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kOpClasses; ++q) {
      uint32_t pos = run[q] + (uint32_t)__builtin_popcountll(m[q] & below), tot = 0;
#pragma unroll
      for (uint32_t v = 0; v < kWaves; ++v) {
        if (v < w) pos += wsum[q][v];
        tot += wsum[q][v];
      }
      if (k == q) lists[q][pos] = w_lo + i;
      run[q] += tot;
    }
    __syncthreads();
  }
}

// 7. Routes every write to its leaf: sort key = leaf (cap_ for a key that is
// not routed: reserved keys, which fail in the write pass), payload = the
// write's position in w_rel; the parent goes to w_par. T: routing tile
// width (T < NW: NW / T pairs per lane, mvbt_wide.hpp).
template <int NW, int T = NW>
__global__ void __launch_bounds__(256) batch_route_kernel(tree_ptrs tp, batch_job* job) {
  if (ld32_coherent(&job->bad)) return;
  const tree_view<NW, true> tv(tp);
  tile<T> t;
  chunking<T> c;
  __shared__ route_cache<NW, kRangeCacheBytes> rc;
  tv.build_route_cache(rc);
  const uint32_t n = job->n_writes;
  const op_record* ops = job->ops + job->first_op;
  const uint32_t* w_rel = job->w_rel;
  uint32_t* hint = job->sort_k[0];
  uint32_t* pos = job->sort_v[0];
  uint32_t* w_par = job->w_par;
  for (uint64_t base = c.first; base < n; base += c.stride) {
    const uint64_t i = base + t.rank();
    const bool act = i < n;
    const uint32_t key = act ? ops[w_rel[i]].key : 0u;
    const bool go = act && valid_key(key);
    const uint64_t q = t.ballot(go);
    uint32_t st0 = kNone, sp0 = kNone, leaf = kNone, par = kNone;
    if (go) st0 = tv.route_lane(rc, key, sp0);
    if constexpr (T < NW)
      descend_chunk_wide<T, false, kWideRouteGroup<NW / T>>(tv, key, q, leaf, par, t, q, st0, sp0);
    else
      tv.template descend_chunk<false>(key, q, leaf, par, t, q, st0, sp0);
    if (act) {
      const bool routed = go && leaf < tv.cap_;
      hint[i] = routed ? leaf : tv.cap_;
      pos[i] = (uint32_t)i;
      w_par[i] = routed ? par : kNone;
    }
  }
}

// 7'. Copy planning. Every (leaf, batch) group of the sorted writes may copy
// its leaf once, so the groups are counted and numbered here (a stable
// count / scan / assign over the sorted list) and one reservation takes a
// slot for each from the free stack and the never-used frontier, and enters
// them all in the retire ring: the write pass then copies into its group's
// slot without touching the shared allocation and retirement counters. A
// slot its group does not use (no change, or the per-key path took over)
// is simply freed by the next reclamation.
MVBT_DEV bool epoch_start(const batch_job* job, const uint32_t* hint, const uint32_t* pos, uint32_t i) {
  if (i == 0) return true;
  if (hint[i] != hint[i - 1]) return true;
  const uint32_t* op_batch = job->op_batch;
  const uint32_t* w_rel = job->w_rel;
  return op_batch[w_rel[pos[i]]] != op_batch[w_rel[pos[i - 1]]];
}

__global__ void __launch_bounds__(kSortBlock) batch_epoch_count_kernel(batch_job* job, uint32_t which) {
  __shared__ uint32_t part[kSortBlock / 64];
  const uint32_t n = ld32_coherent(&job->bad) ? 0u : job->n_writes;
  const uint32_t* hint = job->sort_k[which];
  const uint32_t* pos = job->sort_v[which];
  uint32_t a, b;
  segment_of(n, blockIdx.x, gridDim.x, a, b);
  uint32_t c = 0;
  for (uint32_t i = a + threadIdx.x; i < b; i += kSortBlock) c += epoch_start(job, hint, pos, i) ? 1u : 0u;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < kSortBlock / 64; ++w) tot += part[w];
    job->seg[blockIdx.x] = tot;
  }
}

// (after radix_scan_kernel: seg holds each segment's first group number and
// n_epochs the total) ep[i] = the group number of sorted write i.
__global__ void __launch_bounds__(kSortBlock) batch_epoch_assign_kernel(batch_job* job, uint32_t which) {
  constexpr uint32_t kWaves = kSortBlock / 64;
  __shared__ uint32_t wsum[kWaves];
  if (ld32_coherent(&job->bad)) return;
  const uint32_t n = job->n_writes;
  const uint32_t* hint = job->sort_k[which];
  const uint32_t* pos = job->sort_v[which];
  uint32_t* ep = job->sort_k[which ^ 1u];
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint64_t below = (1ull << lane) - 1;
  uint32_t a, b;
  segment_of(n, blockIdx.x, gridDim.x, a, b);
  uint32_t run = job->seg[blockIdx.x];  // groups started before this segment
  for (uint32_t c = a; c < b; c += kSortBlock) {
    const uint32_t i = c + tid;
    const bool st = i < b && epoch_start(job, hint, pos, i);
    const uint64_t m = __builtin_amdgcn_ballot_w64(st);
    if (lane == 0) wsum[w] = (uint32_t)__builtin_popcountll(m);
    __syncthreads();
    uint32_t before = run + (uint32_t)__builtin_popcountll(m & below), tot = 0;
#pragma unroll
    for (uint32_t v = 0; v < kWaves; ++v) {
      if (v < w) before += wsum[v];
      tot += wsum[v];
    }
    // the group number is the count of starts at or before i, minus one
    if (i < b) ep[i] = before + (st ? 1u : 0u) - 1u;
    run += tot;
    __syncthreads();
  }
}

// One thread: reserves n_epochs slots (free stack first, then the
// frontier) and as many retire-ring positions as slots exist.
__global__ void batch_copy_reserve_kernel(batch_job* job, tree_ptrs tp) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (ld32_coherent(&job->bad)) return;
  tree_state* st = tp.st;
  const uint32_t E = job->n_epochs;
  long long top = atomic_add_s64(&st->free_top, -(long long)E);
  const uint32_t got = (uint32_t)(top < 0 ? 0 : (top > (long long)E ? (long long)E : top));
  unsigned long long b = 0;
  if (got < E) b = atomic_add64(&st->bump, (unsigned long long)(E - got));
  const unsigned long long fresh = b < (unsigned long long)tp.cap_ ? (unsigned long long)tp.cap_ - b : 0ull;
  const uint32_t valid = got + (uint32_t)((unsigned long long)(E - got) < fresh ? (E - got) : fresh);
  if (valid < E) atomic_or32(&st->error_flags, kErrPoolExhausted);
  const unsigned long long ring = atomic_add64(&st->retire_tail, (unsigned long long)valid);
  job->cp_top = top;
  job->cp_got = got;
  job->cp_bump = b;
  job->cp_valid = valid;
  job->cp_ring = ring;
}

// Writes the slot of every group (kNone past the pool) to cslot and the
// existing ones to the retire ring.
__global__ void batch_copy_slots_kernel(batch_job* job, tree_ptrs tp, uint32_t which) {
  if (ld32_coherent(&job->bad)) return;
  const uint32_t E = job->n_epochs, got = job->cp_got, valid = job->cp_valid;
  const long long top = job->cp_top;
  const unsigned long long b = job->cp_bump, ring = job->cp_ring;
  const unsigned long long mask = tp.retire_cap - 1;
  uint32_t* cslot = job->sort_v[which ^ 1u];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < E; k += gridDim.x * blockDim.x) {
    uint32_t slot = kNone;
    if (k < got) slot = ld32_coherent(tp.free_stack + (top - 1 - (long long)k));
    else if (k < valid) slot = (uint32_t)(b + (k - got));
    cslot[k] = slot;
    if (k < valid) tp.retire_ring[(ring + k) & mask] = slot;
  }
}

// 7''. Gathers each sorted write's fields into sorted order, so the write
// pass reads them with coalesced loads instead of chasing the position, the
// operation and its batch one dependent load after another.
__global__ void __launch_bounds__(256) batch_gather_kernel(batch_job* job, uint32_t which) {
  if (ld32_coherent(&job->bad)) return;
  const uint32_t n = job->n_writes;
  const uint32_t* pos = job->sort_v[which];
  const uint32_t* ep = job->sort_k[which ^ 1u];
  const uint32_t* cslot = job->sort_v[which ^ 1u];
  const op_record* ops = job->ops + job->first_op;
  uint32_t* const* g = job->gat;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t pv = pos[i];
    const uint32_t rel = job->w_rel[pv];
    const op_record o = ops[rel];
    g[0][i] = o.key;
    g[1][i] = o.arg;
    g[2][i] = rel | (o.type == kOpErase ? 0x80000000u : 0u);
    g[3][i] = job->op_batch[rel];
    g[4][i] = job->w_par[pv];
    g[5][i] = cslot[ep[i]];
  }
}

// 8. The write pass over the sorted writes (buffers sort_k/v[which]).
template <int NW>
__global__ void __launch_bounds__(256) batch_sorted_write_kernel(tree_ptrs tp, batch_job* job, uint32_t which) {
  if (ld32_coherent(&job->bad)) return;
  const tree_view<NW, true> tv(tp);
  tile<NW> t;
  chunking<NW> c;
  const uint32_t r = t.rank();
  const uint32_t n = job->n_writes;
  const uint32_t* hint = job->sort_k[which];
  const uint32_t* const* g = job->gat;  // the writes' fields in sorted order
  uint32_t* results = job->results + job->first_op;
  const uint32_t clock0 = job->clock0;
  write_ctx ctx = tv.begin_writes(t);
  ctx.exact = 1;
  ctx.track = 0;  // stamps are fixed per batch: clock0 + b + 1
  for (uint64_t p = c.first; p < n; p += c.stride) {
    uint64_t i = p + r;
    bool start = i < n && (i == 0 || hint[i] != hint[i - 1]);
    const uint64_t starts = t.ballot(start);
    if (!starts) continue;  // the window lies inside a group that started earlier
    // sub-windows of NW positions from the first group start in the window to
    // the first group start at or past its end
    const uint64_t wend = p + NW;
    for (uint64_t a = p + (uint64_t)lowest_bit(starts);; a += NW) {
      i = a + r;
      const bool in = i < n;
      const uint32_t h = in ? hint[i] : kNone;
      start = in && (i == 0 || h != hint[i - 1]);
      uint64_t mine = t.ballot(in);
      const uint64_t foreign = t.ballot(start && i >= wend);
      bool last = a + NW >= n;
      if (foreign) {
        mine &= (1ull << lowest_bit(foreign)) - 1;
        last = true;
      }
      if (!mine) break;
      const bool act = (mine >> r) & 1;
      op_record op{};
      uint32_t rel = 0, b = 0, par = kNone, planned = kNone;
      if (act) {
        op.key = g[0][i];
        op.arg = g[1][i];
        const uint32_t rt = g[2][i];
        rel = rt & 0x7FFFFFFFu;
        op.type = (rt >> 31) ? kOpErase : kOpInsert;
        b = g[3][i];
        par = g[4][i];
        planned = g[5][i];
      }
      const uint32_t cur = clock0 + b + 1;
      const uint32_t leaf = (act && h < tv.cap_) ? h : kNone;
      // lane r's group begins at glead (lane 0 or a group start of this sub-window)
      const uint64_t gs = t.ballot(start) | 1ull;
      const uint64_t upto = (r == 63) ? ~0ull : ((2ull << r) - 1);
      const uint64_t from_lead = ~((1ull << highest_bit(gs & upto)) - 1);
      uint32_t res = 0;
      // rounds: every group writes its earliest remaining batch (a group's
      // lanes are in batch order, so that is its lowest remaining lane's)
      for (uint64_t R = mine; R;) {
        const uint64_t ahead = R & from_lead;
        const uint32_t front = ahead ? (uint32_t)lowest_bit(ahead) : r;
        const uint32_t fb = t.shfl(b, front);
        const bool sel = ((R >> r) & 1) && b == fb;
        const uint64_t S = t.ballot(sel);
        for (int w = 0; w < 2; ++w) {
          uint64_t q = t.ballot(sel && op.type == (w == 0 ? kOpInsert : kOpErase));
          if (!q) continue;
          ctx.want = (uint32_t)__builtin_popcountll(q);
          q = (w == 0) ? tv.template write_chunk_at<0, true>(op.key, op.arg, q, leaf, cur, res, ctx, t, planned)
                       : tv.template write_chunk_at<1, true>(op.key, 0u, q, leaf, cur, res, ctx, t, planned);
          while (q) {
            const uint32_t j = (uint32_t)lowest_bit(q);
            ctx.want = (uint32_t)__builtin_popcountll(q);
            q &= q - 1;
            ctx.cur = t.shfl(cur, j);
            const uint32_t k = t.shfl(op.key, j), lf = t.shfl(leaf, j), pr = t.shfl(par, j);
            const uint32_t v = (w == 0) ? tv.template write<0>(k, t.shfl(op.arg, j), ctx, t, lf, pr)
                                        : tv.template write<1>(k, 0u, ctx, t, lf, pr);
            if (r == j) res = v;
          }
        }
        R &= ~S;
      }
      if (act) results[rel] = res;
      if (last) break;
    }
  }
  tv.finish(ctx, t);
}

// 9. Finds, each at its batch's snapshot (no writer runs). T: read tile
// width (T < NW: P = NW / T pairs per lane, mvbt_wide.hpp).
template <int NW, int T = NW>
__global__ void __launch_bounds__(kFindBlock) __attribute__((amdgpu_waves_per_eu(8)))
batch_sorted_find_kernel(tree_ptrs tp, const batch_job* job) {
  if (ld32_coherent(&job->bad)) return;
  const tree_view<NW, true> tv(tp);
  tile<T> t;
  chunking<T> c;
  __shared__ route_cache<NW, kFindCacheBytes> rc;
  tv.build_route_cache(rc);
  const uint32_t n = job->n_finds;
  const uint32_t* list = job->rd_find;
  const uint32_t* op_batch = job->op_batch;
  const op_record* ops = job->ops + job->first_op;
  uint32_t* results = job->results + job->first_op;
  const uint32_t clock0 = job->clock0;
  for (uint64_t base = c.first; base < n; base += c.stride) {
    const uint64_t i = base + t.rank();
    const bool act = i < n;
    const uint32_t rel = act ? list[i] : 0u;
    const uint32_t key = act ? ops[rel].key : 0u;
    const uint32_t ts = act ? clock0 + op_batch[rel] : 0u;
    uint32_t result = kEmpty;
    if constexpr (T < NW) {
      result = read_chunk_wide<T, false, true, kWideFindGroup<NW / T>>(tv, key, 0u, ts, nullptr, 0u, t.ballot(act), t,
                                                                        &rc);
    } else if constexpr (NW >= 32) {
      // as in find_kernel: the pass-based engine keeps more walks in flight
      result = tv.template read_chunk<false, route_cache<NW, kFindCacheBytes>, true, tree_view<NW, true>::kFindGroup>(
          key, 0u, ts, nullptr, 0u, t.ballot(act), t, &rc);
    } else {
      uint32_t start = kNone, unused;
      if (act && valid_key(key)) start = tv.route_lane(rc, key, unused);
      for (uint64_t q = t.ballot(act); q;) {
        const uint32_t j = (uint32_t)lowest_bit(q);
        q &= q - 1;
        const uint32_t v = tv.template find<false>(t.shfl(key, j), t.shfl(ts, j), t, t.shfl(start, j));
        if (t.rank() == j) result = v;
      }
    }
    if (act) results[rel] = result;
  }
}

// 10. Ranges, each at its batch's snapshot: the pass-based reader with a
// snapshot and an output slot per lane (T: read tile width, as above).
template <int NW, int T = NW>
__global__ void __launch_bounds__(256) batch_sorted_range_kernel(tree_ptrs tp, const batch_job* job) {
  if (ld32_coherent(&job->bad)) return;
  using rc_t = route_cache<NW, kRangeCacheBytes>;
  const tree_view<NW, true> tv(tp);
  tile<T> t;
  chunking<T> c;
  __shared__ rc_t rc;
  tv.build_route_cache(rc);
  const uint32_t n = job->n_ranges, L = job->L;
  const uint32_t* list = job->rd_range;
  const uint32_t* op_batch = job->op_batch;
  const op_record* ops = job->ops + job->first_op;
  uint32_t* results = job->results + job->first_op;
  uint64_t* pairs = job->pairs;
  const unsigned long long pair_slots = job->pair_slots;
  const uint32_t clock0 = job->clock0;
  bool overflow = false, bad_slot = false;
  for (uint64_t base = c.first; base < n; base += c.stride) {
    const uint64_t i = base + t.rank();
    const bool act = i < n;
    const uint32_t rel = act ? list[i] : 0u;
    op_record op{};
    if (act) op = ops[rel];
    const uint32_t ts = act ? clock0 + op_batch[rel] : 0u;
    uint64_t* out = nullptr;
    if (act && pairs) {
      if (op.slot < pair_slots) out = pairs + (size_t)op.slot * L;
      else bad_slot = true;
    }
    uint32_t cnt;
    if constexpr (T < NW)
      cnt = read_chunk_wide<T, true, true, kWideRangeGroup<NW / T>>(tv, op.key == 0 ? 1u : op.key, op.arg, ts, out, L,
                                                                    t.ballot(act), t, &rc);
    else
      cnt = tv.template read_chunk<true, rc_t, true>(op.key == 0 ? 1u : op.key, op.arg, ts, out, L, t.ballot(act), t,
                                                     &rc);
    if (act) results[rel] = cnt;
    overflow |= act && cnt > L;
  }
  const uint32_t err = (t.any(overflow) ? kErrRangeOverflow : 0u) | (t.any(bad_slot) ? kErrCorrupt : 0u);
  if (err && t.leader()) atomic_or32(&tv.st->error_flags, err);
}

}  // namespace mvbt
