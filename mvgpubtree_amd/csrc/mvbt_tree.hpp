// Device-side multiversion B-link tree over uint32 keys and values.
//
// Reference behaviour restated here (all paths under the reference's include/):
//   * cooperative_find(key, ts)            gpu_versioned_blink_tree.hpp:439-474,
//                                          plain gpu_blink_tree.hpp:278-303
//   * concurrent_cooperative_range_query   gpu_versioned_blink_tree.hpp:505-559
//   * cooperative_insert_versioned_in_place gpu_versioned_blink_tree.hpp:825-1086,
//     plain cooperative_insert_blink       gpu_blink_tree.hpp:523-733
//   * cooperative_erase                    gpu_versioned_blink_tree.hpp:356-436
//   * take_snapshot(tile)                  gpu_versioned_blink_tree.hpp:333-344
//   * traverse_version_list / side links   gpu_versioned_blink_tree.hpp:1480-1603
//
// Design (differs from the reference on purpose):
//   * Only leaves carry versions. Internal nodes are routing data: every reader,
//     whatever its snapshot, routes through the latest internal nodes and then
//     walks the leaf's version chain to the newest version stamped <= snapshot.
//     Because keys only ever move right (splits, never merges) and a leaf made
//     by a split inherits the version chain of the leaf it came from, the
//     version found is the leaf that held the key at the snapshot. A split
//     therefore copies at most the split leaf, never its parent.
//   * The root is internal from creation and never moves, so no node ever
//     changes between leaf and internal: a torn read of an internal node mixes
//     {pivot, child} pairs of two valid versions and can only route too far
//     left, which the B-link right move repairs.
//   * Writers take locks with try_lock only and restart on failure; no tile
//     ever spins on a lock inside a loop of its own, because the 64/W tiles of
//     a wave execute in lockstep and a spinning tile would starve its holder.
//   * A leaf modified in a new epoch is first copied (the copy keeps the old
//     content and stamp), then its version word {cur, copy} is published, and
//     only then are its data lanes rewritten, each step behind a release fence.
//     A concurrent snapshot reader that re-reads the version word after reading
//     the leaf therefore detects any overlapping rewrite.
#pragma once
#include "mvbt_node.hpp"

namespace mvbt {

// Device-resident state of one tree (one per handle). Counters that many
// tiles update are each on their own 128-byte line: a line shared by several
// hot counters serialises every access to any of them.
struct alignas(128) tree_state {
  uint32_t root;           // index of the root node (fixed after creation)
  uint32_t pad_root[31];
  uint32_t clock;          // write stamp; take_snapshot returns clock++
  uint32_t pad_clock[31];
  unsigned long long bump;  // next never-used node index
  unsigned long long pad_bump[15];
  long long free_top;       // recycled indices available in free_stack
  unsigned long long pad_free[15];
  unsigned long long spill_count;  // unused magazine slots parked in `spill`
  unsigned long long pad_spill[15];
  unsigned long long retire_tail;  // total retirements appended to the ring
  unsigned long long pad_tail[15];
  unsigned long long retire_head;  // first retirement not yet reclaimed
  unsigned long long pad_head[15];
  uint32_t error_flags;            // kErr* bits
  uint32_t pad_flags[31];
  // cold
  unsigned long long capacity;     // pool capacity in nodes
  unsigned long long retire_cap;   // ring capacity (power of two)
  unsigned long long reclaimed;    // total nodes returned to the free stack
  unsigned long long peak_live;    // max live nodes observed at reclaim/stats
  unsigned long long scratch[4];   // host<->device hand-off words
  uint32_t call_snapshot;          // snapshot reserved by the current mixed call
  uint32_t pad_call;
  unsigned long long series_count;  // reclamation epochs recorded in the memory timeline
};

// One entry of the memory timeline (mvbt_epoch_stat in include/mvbt.h),
// appended by every reclamation: the {usage, total allocated} pair the
// reference records at each epoch advance (memory_reclaimer.hpp:387-398),
// with the tree clock and the reclamation's own figures.
struct epoch_stat {
  unsigned long long epoch, clock, live_nodes, allocated_total, freed_nodes, retired_pending, reclaimed_total,
      capacity_nodes;
};
constexpr uint32_t kSeriesCap = 4096;  // entries kept (a ring of the latest)

// Result of a snapshot taken on the device: the id and the retire-ring
// position at the epoch boundary (used by epoch reclamation).
struct snapshot_record {
  uint32_t id;
  uint32_t pad;
  unsigned long long ring_pos;
};

// Per-tile writer context held in registers for the life of a tile's work:
// an allocation magazine (lane i holds a free node index, lanes < mag_n valid)
// and a retirement buffer (lanes < ret_n valid). One atomic on the shared
// counters serves W allocations or W retirements. finish() must run before
// the tile exits: it flushes retirements and parks unused slots in the spill
// list, which the next reclamation returns to the free stack.
struct write_ctx {
  uint32_t mag;    // per lane
  uint32_t ret;    // per lane
  uint32_t mag_n;  // tile-uniform
  uint32_t ret_n;  // tile-uniform
  uint32_t cur;    // write stamp (tile-uniform)
  uint32_t track;  // nonzero: the stamp follows the clock, read again under each leaf lock;
                   // zero: a launch-fixed stamp (batched worker, mixed calls)
  uint32_t want;   // writes this tile still has queued (caps refills)
  uint32_t grow;   // next refill size: starts at 2, doubles up to W
  uint32_t exact;  // nonzero: `want` is exact, version-copy refills take want + 2
  uint32_t backoff;  // sleep rounds after the next failed lock (tile-uniform)
};

// Device pointers of one tree, independent of the node width (passed to
// kernels by value; tree_view adds the width-specific operations).
struct tree_ptrs {
  uint64_t* nodes;         // pool: node i occupies nodes[i*NW .. i*NW+NW)
  tree_state* st;
  uint32_t* free_stack;    // recycled node indices, [0, free_top)
  uint32_t* spill;         // parked unused magazine slots, [0, spill_count)
  uint32_t* retire_ring;   // retired node indices, ring of retire_cap
  unsigned long long retire_cap;
  uint32_t root_;
  uint32_t cap_;           // pool capacity: every followed index must be below it
};

// The top levels of the tree, copied into LDS by each workgroup of the find
// and range kernels (and of the insert kernel and the chunked batched worker
// at 32-pair nodes) when it
// starts (tree_view::build_route_cache) and used for routing only
// (tree_view::route_lane): an operation then reads from global memory only
// the levels below them. Levels are cached breadth-first, whole,
// while they hold internal nodes only and fit kBytes; the children of a
// cached node occupy consecutive slots in entry order. An image may go stale
// while the kernel runs (a concurrent writer's splits); it then routes to a
// node left of the right one, which the B-link right moves of the first
// uncached level repair. The find kernel runs 1024-thread workgroups so that
// two of them (32 waves) share a CU with 72 KB of node images each: at 40M
// keys and 16-pair nodes that is the root and the four levels below it
// (about 520 nodes). The range kernel (4 waves per SIMD) keeps 256-thread
// workgroups, whose 16 KB hold the root and two levels (about 54 nodes).
template <int NW, uint32_t BYTES>
struct route_cache {
  static constexpr uint32_t kBytes = BYTES;
  // Node images are NW + 1 pairs apart: lanes routing different keys read
  // the same entry of different nodes, and an unpadded stride of NW pairs
  // (a multiple of the LDS bank count) would put all of them on one bank.
  static constexpr uint32_t kStride = NW + 1;
  static constexpr uint32_t kMaxNodes = kBytes / (kStride * 8);
  static constexpr uint16_t kNoChild = 0xFFFF;
  uint64_t nodes[kMaxNodes * kStride];
  MVBT_DEV uint64_t* img(uint32_t s) { return nodes + s * kStride; }
  MVBT_DEV const uint64_t* img(uint32_t s) const { return nodes + s * kStride; }
  uint32_t index[kMaxNodes];        // global index of the node in slot s
  uint16_t first_child[kMaxNodes];  // slot of its first child, or kNoChild
  uint32_t count;                   // slots routed through (internal node images)
  uint32_t stop;                    // build: the last level has been reached
};

template <int NW, bool VER>
struct tree_view : tree_ptrs {
  using node_t = node<NW, VER>;
  using tile_t = tile<NW>;
  static constexpr int kWidth = NW;
  static constexpr bool kVersioned = VER;

  MVBT_HD tree_view() = default;
  MVBT_HD explicit tree_view(const tree_ptrs& p) : tree_ptrs(p) {}

  MVBT_DEV uint64_t* at(uint32_t idx) const { return nodes + (size_t)idx * NW; }
  MVBT_DEV uint32_t* lock_word(uint32_t idx) const {
    return reinterpret_cast<uint32_t*>(at(idx) + node_t::kMetaLane) + 1;
  }
  MVBT_DEV void flag(uint32_t bits, const tile_t& t) const {
    if (t.leader()) atomic_or32(&st->error_flags, bits);
  }

  template <bool CONC>
  MVBT_DEV void load(node_t& n, uint32_t idx, const tile_t& t) const {
    if constexpr (CONC) n.load(at(idx), t);
    else n.load_weak(at(idx), t);
  }

  // ---- writer context --------------------------------------------------------
  MVBT_DEV write_ctx begin_writes(const tile_t& t) const {
    write_ctx c;
    c.mag = kNone;
    c.ret = kNone;
    c.mag_n = 0;
    c.ret_n = 0;
    c.cur = refresh_stamp(t);
    c.track = 1;
    c.want = NW;
    c.grow = 2;
    c.exact = 0;
    c.backoff = 1;
    return c;
  }
  // The clock (leader load, broadcast).
  MVBT_DEV uint32_t refresh_stamp(const tile_t& t) const {
    uint32_t c = 0;
    if (VER && t.leader()) c = ld32_coherent(&st->clock);
    return t.shfl(c, 0);
  }
  // Stamp of a write to leaves this tile has just locked. A tile that tracks
  // the clock reads it only once it holds the locks: every snapshot taken
  // before then is older than the stamp (the write is invisible to it, its
  // leaf gets a version copy), and a snapshot taken later was taken after
  // the lock. A leaf is therefore rewritten in place (stamp == cur) only by
  // a writer that locked it before any snapshot that can read this content
  // existed; concurrent snapshot readers rely on that (to_version).
  MVBT_DEV void stamp_under_lock(write_ctx& c, const tile_t& t) const {
    if (VER && c.track) c.cur = refresh_stamp(t);
  }
  // True when a leaf with version word `ver` must be copied out before a
  // write stamped c.cur changes it: its content belongs to an older epoch.
  // A leaf stamped after c.cur (a launch-fixed stamp overlapping another
  // writer) is written in place under its newer stamp, so every version
  // chain stays ordered by stamp, and the overlap is flagged.
  MVBT_DEV bool needs_copy(uint64_t ver, const write_ctx& c, const tile_t& t) const {
    return needs_copy_at(ver, c.cur, t);
  }
  MVBT_DEV bool needs_copy_at(uint64_t ver, uint32_t cur, const tile_t& t) const {
    const uint32_t s = node_t::stamp(ver);
    if (s > cur) flag(kErrStampOrder, t);
    return s < cur;
  }

  // ---- allocation (roles of the reference's device_bump_allocator.hpp:45-54
  // and slab_alloc.hpp:131-205). Refills take W recycled indices from the free
  // stack, the rest from the never-used frontier. Returns kNone when the pool
  // is exhausted.
  // Refill size. A version copy in a write call that knows its queued writes
  // exactly (c.exact) takes want + 2 at once: in a new epoch nearly every
  // write copies its leaf, and one atomic on the shared counter then serves
  // the whole chunk. Otherwise (splits, and callers that hold one operation
  // at a time) refills grow geometrically (2, 4, 8, ... W) with demand, so a
  // tile never parks much more than it has used: with many tiles, parked
  // slots could otherwise exhaust a nearly full pool. A plain tree writes in
  // place and allocates only to split (2 for the root). Capped at W.
  MVBT_DEV uint32_t refill_size(write_ctx& c, bool copy) const {
    uint32_t k = VER ? c.want + 2 : 2;
    if (!(copy && c.exact)) {
      if (k > c.grow) k = c.grow;
      if (c.grow < (uint32_t)NW) c.grow *= 2;
    }
    return k < (uint32_t)NW ? k : (uint32_t)NW;
  }
  MVBT_DEV void refill(write_ctx& c, bool copy, const tile_t& t) const {
    const uint32_t k = refill_size(c, copy);
    long long top = 0;
    unsigned long long b = 0;
    if (t.leader()) {
      top = atomic_add_s64(&st->free_top, -(long long)k);
      long long got = top < 0 ? 0 : (top > (long long)k ? (long long)k : top);
      if (got < (long long)k) b = atomic_add64(&st->bump, (unsigned long long)(k - got));
    }
    top = (long long)t.shfl64((uint64_t)top, 0);
    b = t.shfl64(b, 0);
    const uint32_t got = (uint32_t)(top < 0 ? 0 : (top > (long long)k ? (long long)k : top));
    const uint32_t r = t.rank();
    uint32_t idx = kNone;
    if (r < got) {
      idx = ld32_coherent(free_stack + (top - 1 - r));
    } else if (r < k) {
      unsigned long long f = b + (r - got);
      idx = f < (unsigned long long)cap_ ? (uint32_t)f : kNone;
    }
    uint64_t ok = t.ballot(idx != kNone);
    // valid lanes are a prefix: stack entries first, then fresh indices
    c.mag = idx;
    c.mag_n = (uint32_t)__builtin_popcountll(ok);
    if (c.mag_n < k) flag(kErrPoolExhausted, t);
  }
  MVBT_DEV uint32_t alloc(write_ctx& c, const tile_t& t, bool copy = false) const {
    if (c.mag_n == 0) refill(c, copy, t);
    if (c.mag_n == 0) return kNone;
    c.mag_n -= 1;
    return t.shfl(c.mag, c.mag_n);
  }
  // Epoch retirement of a version copy (reference DEBR_device::retire,
  // memory_reclaimer.hpp:146-228). The copy becomes garbage once no snapshot
  // older than the epoch that created it is live.
  MVBT_DEV void flush_retired(write_ctx& c, const tile_t& t) const {
    if (c.ret_n == 0) return;
    unsigned long long pos = 0, head = 0;
    if (t.leader()) {
      pos = atomic_add64(&st->retire_tail, (unsigned long long)c.ret_n);
      head = __hip_atomic_load(&st->retire_head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    pos = t.shfl64(pos, 0);
    head = t.shfl64(head, 0);
    const uint32_t r = t.rank();
    bool over = false;
    if (r < c.ret_n) {
      unsigned long long p = pos + r;
      if (p - head < retire_cap) st32_coherent(retire_ring + (p & (retire_cap - 1)), c.ret);
      else over = true;  // the slot still holds an unreclaimed entry: this copy leaks
    }
    if (t.any(over)) flag(kErrRetireOverflow, t);
    c.ret_n = 0;
  }
  MVBT_DEV void retire(write_ctx& c, uint32_t idx, const tile_t& t) const {
    if (t.rank() == c.ret_n) c.ret = idx;
    c.ret_n += 1;
    if (c.ret_n == (uint32_t)NW) flush_retired(c, t);
  }
  // Ends a tile's writes: flush retirements, park unused slots.
  MVBT_DEV void finish(write_ctx& c, const tile_t& t) const {
    flush_retired(c, t);
    if (c.mag_n) {
      unsigned long long s = 0;
      if (t.leader()) s = atomic_add64(&st->spill_count, (unsigned long long)c.mag_n);
      s = t.shfl64(s, 0);
      if (t.rank() < c.mag_n) st32_coherent(spill + s + t.rank(), c.mag);
      c.mag_n = 0;
    }
  }

  // ---- locks (reference versioned_node.hpp:528-568; here try-lock only) ----
  // Test-and-test-and-set: a held lock costs a waiter one coherent load, which
  // the L2 serves beside other loads, instead of an atomic, which the L2
  // serialises on the line the holder itself needs.
  MVBT_DEV bool try_lock(uint32_t idx, const tile_t& t) const {
    uint32_t old = kLockBit;
    if (t.leader()) {
      old = ld32_coherent(lock_word(idx));
      if (!(old & kLockBit)) old = atomic_or32(lock_word(idx), kLockBit);
    }
    return (t.shfl(old, 0) & kLockBit) == 0;
  }
  // The same, with the test taken from metadata the tile has just loaded
  // coherently (no extra round trip on the uncontended path).
  MVBT_DEV bool try_lock_seen(uint32_t idx, uint64_t meta, const tile_t& t) const {
    if (hi32(meta) & kLockBit) return false;
    uint32_t old = 0;
    if (t.leader()) old = atomic_or32(lock_word(idx), kLockBit);
    return (t.shfl(old, 0) & kLockBit) == 0;
  }
  // After a failed lock: sleep, doubling the time per consecutive failure up
  // to kMaxBackoff rounds (the whole wave sleeps, so the cap stays small).
  // Without it, hundreds of tiles re-descending onto the few nodes of a small
  // tree keep the lock holders' lines busy.
  static constexpr uint32_t kMaxBackoff = 32;
  MVBT_DEV static void back_off(write_ctx& c) {
    for (uint32_t i = 0; i < c.backoff; ++i) __builtin_amdgcn_s_sleep(2);
    if (c.backoff < kMaxBackoff) c.backoff <<= 1;
  }
  MVBT_DEV void unlock(uint32_t idx, const tile_t& t) const {
    order_release();
    if (t.leader()) atomic_and32(lock_word(idx), ~kLockBit);
  }
  // Both locks after one wait for the tile's earlier stores.
  MVBT_DEV void unlock2(uint32_t a, uint32_t b, const tile_t& t) const {
    order_release();
    if (t.leader()) {
      atomic_and32(lock_word(a), ~kLockBit);
      atomic_and32(lock_word(b), ~kLockBit);
    }
  }

  // ---- snapshots ------------------------------------------------------------
  // Takes a snapshot: returns the id s; writes after this call are stamped > s.
  MVBT_DEV uint32_t take_snapshot(const tile_t& t) const {
    uint32_t s = 0;
    if (t.leader()) {
      fence_release();
      s = __hip_atomic_fetch_add(&st->clock, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    }
    return t.shfl(s, 0);
  }

  // Re-reads a leaf's version word (leader load, broadcast).
  MVBT_DEV uint64_t reread_version(uint32_t idx, const tile_t& t) const {
    uint64_t v = 0;
    if (t.leader()) v = ld_coherent(at(idx) + node_t::kVerLane);
    return t.shfl64(v, 0);
  }

  // Moves `n` (the leaf stored at `idx`) to its newest version stamped <= ts.
  // Returns false if the key range did not exist at ts (version gap).
  // CONC (writers may run): a live leaf whose stamp is visible is accepted
  // only if its version word is unchanged when read again after the leaf
  // read returned. Writers that copy publish the new version word before
  // they touch the data, so any such write overlapping the read is seen.
  // GUARD (writers of other launches may run): a leaf may also be rewritten
  // in place, without a new version word, by a writer that locked it before
  // the snapshot existed (stamp_under_lock); its content is trusted only
  // when read after a completed read showed the leaf unlocked. Without
  // GUARD the caller guarantees that no in-place writer of a visible epoch
  // runs (the batched worker's frontier, the mixed calls' single snapshot).
  // ts == kLatest is handled by the callers (see find_in).
  template <bool CONC, bool GUARD = CONC>
  MVBT_DEV bool to_version(node_t& n, uint32_t idx, uint32_t ts, const tile_t& t) const {
    if constexpr (!VER) {
      return true;
    } else {
      if (ts == kLatest) return true;
      uint64_t ver = n.version(t);
      if (node_t::stamp(ver) <= ts) {
        if constexpr (!CONC) return true;
        bool clear = !GUARD;  // GUARD: `n` was read after the leaf was seen unlocked
        for (uint32_t step = 0;; ++step) {
          if (step >= kMaxReadSteps) {
            flag(kErrCorrupt, t);
            return false;
          }
          if (!clear) {
            clear = !node_t::is_locked(n.meta(t));
            if (!clear) __builtin_amdgcn_s_sleep(2);
          } else {
            const uint64_t ver2 = reread_version(idx, t);
            if (ver2 == ver) return true;
            clear = !GUARD;
          }
          // read the leaf again (after the observation above returned)
          load<true>(n, idx, t);
          ver = n.version(t);
          if (node_t::stamp(ver) > ts) break;
        }
      }
      for (uint32_t step = 0; node_t::stamp(ver) > ts; ++step) {
        uint32_t o = node_t::older(ver);
        if (o == kNone) {
          flag(kErrVersionGap, t);
          return false;
        }
        if (o >= cap_ || step >= kMaxReadSteps) {
          flag(kErrCorrupt, t);
          return false;
        }
        load<CONC>(n, o, t);  // version copies are immutable once published
        ver = n.version(t);
      }
      return true;
    }
  }

  // Descends (latest routing, B-link right moves) to the leaf covering `key`.
  // Returns kNone (and flags kErrCorrupt) if the walk leaves the pool.
  // `parent` (optional) receives the internal node the leaf was reached from
  // (kNone after a right move at the leaf level). CACHED_ROUTE (with CONC):
  // internal nodes are read through the L1, where the top levels stay hot
  // (routing only: a stale image routes too far left and the right moves
  // repair it), and a node found to be a leaf is read again coherently.
  template <bool CONC, bool CACHED_ROUTE = false>
  MVBT_DEV uint32_t descend(uint32_t key, node_t& n, const tile_t& t, uint32_t* parent = nullptr,
                            uint32_t start = kNone, uint32_t start_parent = kNone) const {
    uint32_t idx = root_, par = kNone;
    if (start < cap_) {
      idx = start;
      par = start_parent;
    }
    for (uint32_t step = 0; step < kMaxReadSteps; ++step) {
      if constexpr (CONC && CACHED_ROUTE) {
        n.load_weak(at(idx), t);
        if (node_t::is_leaf(n.meta(t))) n.load(at(idx), t);
      } else {
        load<CONC>(n, idx, t);
      }
      uint64_t m = n.meta(t);
      uint32_t next;
      if (key >= node_t::high_key(m)) {
        next = node_t::sibling(m);
        par = kNone;
      } else if (node_t::is_leaf(m)) {
        if (parent) *parent = par;
        return idx;
      } else {
        next = n.child_for(key, t);
        par = idx;
      }
      if (next >= cap_) break;
      idx = next;
    }
    flag(kErrCorrupt, t);
    return kNone;
  }

  // Descends the keys of a whole chunk together, one tree level per pass, in
  // groups of kGroup keys (8 measured faster than 16 on MI355X at 40M keys).
  // Lane j holds key j; `todo` (tile-uniform) marks the lanes to route. Each
  // pass issues the node loads of every unfinished key of a group before it
  // uses any of them, so a group of G keys waits about one dependent latency
  // per level instead of G (the descent is latency-bound: one random node
  // per level and key). On return lane j holds the leaf covering key j and
  // its parent (kNone when the leaf was reached by a right move), or kNone as
  // leaf when the walk left the pool (kErrCorrupt is flagged). Routing only:
  // writers re-check both under lock, readers validate leaf versions.
  // `start` / `start_par` (per lane): where lane j's walk begins (a node from
  // the LDS route cache) and the node it was reached from; kNone: the root.
  // Lanes in `early` (writes) stop at the level above the leaves and take the
  // leaf from their parent's entry without reading it: a writer locks the
  // leaf before it reads it anyway, and checks under the lock that it is a
  // leaf covering the key (a torn read of a root being split can name an
  // internal node here). Readers (not in `early`) always end on a node read
  // as a leaf.
  static constexpr int kGroup = NW < 8 ? NW : 8;
  template <bool CONC>
  MVBT_DEV void descend_chunk(uint32_t key, uint64_t todo, uint32_t& leaf, uint32_t& parent,
                              const tile_t& t, uint64_t early = 0, uint32_t start = kNone,
                              uint32_t start_par = kNone) const {
    const uint32_t r = t.rank();
    uint32_t cur = root_, par = kNone;
    if (start < cap_) {
      cur = start;
      par = start_par;
    }
    bool corrupt = false;
    todo &= t.ballot(valid_key(key));  // reserved keys are not routed (callers reject them)
    for (int base = 0; base < NW; base += kGroup) {
      uint64_t active = (todo >> base) & ((kGroup == 64) ? ~0ull : ((1ull << kGroup) - 1));
      for (uint32_t pass = 0; active != 0; ++pass) {
        if (pass >= kMaxReadSteps) {
          corrupt = true;
          if (r >= (uint32_t)base && r < (uint32_t)(base + kGroup) && ((active >> (r - base)) & 1)) cur = kNone;
          break;
        }
        uint64_t p[kGroup];
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
          const uint32_t idx = t.shfl(cur, (uint32_t)(base + j));
          if ((active >> j) & 1) p[j] = CONC ? ld_coherent(at(idx) + r) : ld_weak(at(idx) + r);
        }
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
          if (!((active >> j) & 1)) continue;
          const uint32_t idx = t.shfl(cur, (uint32_t)(base + j));  // not yet advanced
          node_t n;
          n.p = p[j];
          const uint32_t k = t.shfl(key, (uint32_t)(base + j));
          const uint64_t m = n.meta(t);
          uint32_t next = kNone, npar = kNone;
          bool done = false;
          bool arrived = false;  // next is the leaf: writers lock it before reading it
          if (k >= node_t::high_key(m)) {
            next = node_t::sibling(m);
          } else if (node_t::is_leaf(m)) {
            done = true;
          } else {
            next = n.child_for(k, t);
            npar = idx;
            arrived = node_t::above_leaves(m) && ((early >> (base + j)) & 1);
          }
          if (!done && next >= cap_) {
            corrupt = true;
            done = true;
            next = kNone;
            if (r == (uint32_t)(base + j)) cur = kNone;
          } else if (!done && r == (uint32_t)(base + j)) {
            cur = next;
            par = npar;
          }
          if (done || arrived) active &= ~(1ull << j);
        }
      }
    }
    if (corrupt) flag(kErrCorrupt, t);
    leaf = cur;
    parent = par;
  }

  // Pass-based reads of a chunk (for launches in which no writer runs): each
  // key of lane j is a small state machine (route -> leaf -> version chain ->
  // leaf scan), and every pass loads one node for each unfinished key of a
  // group before it uses any of them. A group of kGroup keys then waits about
  // one latency per level, version hop and scanned leaf, instead of one per
  // key and step. RANGE = false: find of key `lo`; returns the value visible
  // at `ts` or kEmpty. RANGE = true: range [lo, hi) with lo >= 1; pairs of the
  // key in lane j go to out + j * cap (at most cap of them; out may be null)
  // and the true count is returned. Keys outside `todo` return kEmpty / 0.
  // LANES: `ts` is lane j's own snapshot and `out` lane j's own output (the
  // sorted batched form's reads, each at its batch's snapshot).
  // RG: walks advanced per pass (finds at 32- and 64-pair nodes take 4: the
  // smaller register footprint lets two 1024-thread workgroups share a CU).
  static constexpr int kReadGroup = NW < 8 ? NW : 8;
  static constexpr int kFindGroup = NW < 8 ? NW : (NW >= 32 ? 4 : 8);
  template <bool RANGE, class RC = route_cache<NW, 1024>, bool LANES = false, int RG = kReadGroup>
  MVBT_DEV uint32_t read_chunk(uint32_t lo, uint32_t hi, uint32_t ts, uint64_t* out, uint32_t cap,
                               uint64_t todo, const tile_t& t, const RC* rc = nullptr) const {
    enum : uint32_t { kRoute = 0, kArrive = 1, kChain = 2 };
    const uint32_t r = t.rank();
    uint32_t cur = root_, phase = kRoute, cnt = 0;
    uint32_t res = RANGE ? 0u : kEmpty;
    uint32_t err = 0;
    todo &= t.ballot(valid_key(lo) && (!RANGE || hi > lo));
    if (rc && ((todo >> r) & 1)) {  // the cached top levels
      uint32_t unused;
      cur = route_lane(*rc, lo, unused);
    }
    for (int base = 0; base < NW; base += RG) {
      uint64_t active = (todo >> base) & ((RG == 64) ? ~0ull : ((1ull << RG) - 1));
      for (uint32_t pass = 0; active != 0; ++pass) {
        if (pass >= kMaxReadSteps) {
          err |= kErrCorrupt;
          break;
        }
        uint64_t p[RG];
#pragma unroll
        for (int j = 0; j < RG; ++j) {
          const uint32_t idx = t.shfl(cur, (uint32_t)(base + j));
          if ((active >> j) & 1) p[j] = ld_weak(at(idx) + r);
        }
#pragma unroll
        for (int j = 0; j < RG; ++j) {
          if (!((active >> j) & 1)) continue;
          const uint32_t lj = (uint32_t)(base + j);
          node_t n;
          n.p = p[j];
          const uint64_t m = n.meta(t);
          const uint32_t k = t.shfl(lo, lj);
          const uint32_t tsj = LANES ? t.shfl(ts, lj) : ts;
          uint32_t ph = t.shfl(phase, lj);
          uint32_t next = kNone;
          bool visit = false, done = false;
          if (ph == kRoute) {
            if (k >= node_t::high_key(m)) next = node_t::sibling(m);
            else if (!node_t::is_leaf(m)) next = n.child_for(k, t);
            else ph = kArrive;
          }
          if (ph == kArrive || ph == kChain) {
            uint64_t ver = 0;
            bool later = false;
            if constexpr (VER) {
              ver = n.version(t);
              later = tsj != kLatest && node_t::stamp(ver) > tsj;
            }
            if (later) {
              next = node_t::older(ver);
              ph = kChain;
              if (next == kNone) {  // no version old enough
                err |= kErrVersionGap;
                done = true;
              }
            } else {
              visit = true;
            }
          }
          uint32_t c = 0;
          if (visit) {
            if constexpr (RANGE) {
              const uint32_t h = t.shfl(hi, lj);
              c = t.shfl(cnt, lj);
              uint64_t* oj = LANES ? reinterpret_cast<uint64_t*>(t.shfl64(reinterpret_cast<uint64_t>(out), lj))
                                   : (out ? out + (size_t)lj * cap : nullptr);
              c += n.collect(k, h, oj, c, cap, t);
              if (h <= node_t::high_key(m) || node_t::sibling(m) == kNoSibling) {
                done = true;
              } else {
                next = node_t::sibling(m);
                ph = kArrive;
              }
            } else {
              const uint32_t v = n.lookup(k, t);
              if (r == lj) res = v;
              done = true;
            }
          }
          if (!done && next >= cap_) {
            err |= kErrCorrupt;
            done = true;
          }
          if (r == lj) {
            if (RANGE && visit) cnt = c;
            if (!done) {
              cur = next;
              phase = ph;
            }
          }
          if (done) active &= ~(1ull << j);
        }
      }
    }
    if (RANGE) res = cnt;
    if (err) flag(err, t);
    return res;
  }

  // Concurrent finds of a chunk at snapshot `ts` (versioned trees; writers
  // stamped after ts may run, as inside the batched form), each key starting
  // from a node that is at or left of its leaf, or above it (lane j: key,
  // start): a leaf hint taken from a parent entry before the batch opened,
  // possibly stale. Pass-based like read_chunk: every pass loads one node for
  // each unfinished key of a group. On the latest tree a key moves right
  // while it is at or above the node's high key and descends through internal
  // nodes, which finds the leaf covering it now; that leaf's version at ts
  // covers it at ts (keys only move right). A live leaf whose stamp is
  // visible gives a provisional answer that the next pass confirms by reading
  // the leaf again and comparing the version words (writers that copy publish
  // a new version word before touching the data); a leaf stamped after ts
  // hands over to its version chain, whose copies are immutable once
  // published. Returns the value or kEmpty.
  MVBT_DEV uint32_t find_chunk_from(uint32_t key, uint32_t start, uint32_t ts, uint64_t todo,
                                    const tile_t& t) const {
    enum : uint32_t { kWalk = 0, kCheck = 1, kChain = 2 };
    const uint32_t r = t.rank();
    uint32_t cur = start, phase = kWalk, res = kEmpty, seen_ver_lo = 0, seen_ver_hi = 0, prov = kEmpty;
    uint32_t err = 0;
    todo &= t.ballot(valid_key(key) && start < cap_);
    for (int base = 0; base < NW; base += kReadGroup) {
      uint64_t active = (todo >> base) & ((kReadGroup == 64) ? ~0ull : ((1ull << kReadGroup) - 1));
      for (uint32_t pass = 0; active != 0; ++pass) {
        if (pass >= kMaxReadSteps) {
          err |= kErrCorrupt;
          break;
        }
        uint64_t p[kReadGroup];
#pragma unroll
        for (int j = 0; j < kReadGroup; ++j) {
          const uint32_t idx = t.shfl(cur, (uint32_t)(base + j));
          if ((active >> j) & 1) p[j] = ld_coherent(at(idx) + r);
        }
#pragma unroll
        for (int j = 0; j < kReadGroup; ++j) {
          if (!((active >> j) & 1)) continue;
          const uint32_t lj = (uint32_t)(base + j);
          node_t n;
          n.p = p[j];
          const uint64_t ver = n.version(t);
          const uint64_t m = n.meta(t);
          const uint32_t k = t.shfl(key, lj);
          uint32_t ph = t.shfl(phase, lj);
          bool done = false;
          uint32_t next = t.shfl(cur, lj), out = kEmpty, keep_prov = kEmpty;
          if (ph == kCheck) {
            const uint64_t was = pack(t.shfl(seen_ver_lo, lj), t.shfl(seen_ver_hi, lj));
            if (ver == was) {
              out = t.shfl(prov, lj);
              done = true;
            } else {
              ph = kWalk;  // the leaf moved on: treat this read as a fresh one
            }
          }
          if (!done && ph == kWalk) {
            if (k >= node_t::high_key(m)) {
              next = node_t::sibling(m);
            } else if (!node_t::is_leaf(m)) {
              next = n.child_for(k, t);
            } else if (node_t::stamp(ver) > ts) {
              next = node_t::older(ver);
              ph = kChain;
              if (next == kNone) {
                err |= kErrVersionGap;
                done = true;
              }
            } else {  // visible live leaf: answer provisionally, confirm next pass
              keep_prov = n.lookup(k, t);
              ph = kCheck;
            }
          } else if (!done && ph == kChain) {
            if (node_t::stamp(ver) > ts) {
              next = node_t::older(ver);
              if (next == kNone) {
                err |= kErrVersionGap;
                done = true;
              }
            } else {
              out = n.lookup(k, t);
              done = true;
            }
          }
          if (!done && next >= cap_) {
            err |= kErrCorrupt;
            done = true;
          }
          if (r == lj) {
            if (done) {
              res = out;
            } else {
              cur = next;
              phase = ph;
              if (ph == kCheck) {
                prov = keep_prov;
                seen_ver_lo = lo32(ver);
                seen_ver_hi = hi32(ver);
              }
            }
          }
          if (done) active &= ~(1ull << j);
        }
      }
    }
    if (err) flag(err, t);
    return res;
  }

  // Block-wide (every thread of the workgroup calls it, before any early
  // return; blockDim.x >= 64): copies the top levels into `rc`. The cache
  // stays empty (routing starts at the root) while fewer than `min_nodes`
  // nodes have been handed out.
  template <class RC>
  MVBT_DEV void build_route_cache(RC& rc, unsigned long long min_nodes = 0) const {
    using rc_t = RC;
    const uint32_t tid = threadIdx.x, nthreads = blockDim.x;
    if (tid == 0) {
      const bool use = min_nodes == 0 || ld_coherent((const uint64_t*)&st->bump) >= min_nodes;
      rc.index[0] = root_;
      rc.count = use ? 1u : 0u;
      rc.stop = use ? 0u : 1u;
    }
    __syncthreads();
    if (rc.stop) return;
    uint32_t b = 0, e = 1;
    // every level that does not stop the build adds at least one slot, so
    // the build stops (rc.stop) within kMaxNodes levels
    for (uint32_t level = 0; level < rc_t::kMaxNodes; ++level) {
      for (uint32_t w = tid; w < (e - b) * NW; w += nthreads) {
        const uint32_t s = b + w / NW;
        rc.img(s)[w % NW] = ld_coherent(at(rc.index[s]) + w % NW);
      }
      __syncthreads();
      if (tid < 64) {  // wave 0: child counts, prefix sum, child slots
        const uint32_t lane = tid;
        uint32_t next = e;
        bool all_internal = true;
        for (uint32_t s0 = b; s0 < e; s0 += 64) {
          const uint32_t s = s0 + lane;
          uint32_t cnt = 0;
          bool internal = true;
          if (s < e) {  // leading entries only: an image read during a split may have gaps
            internal = !(hi32(rc.img(s)[NW - 1]) & kLeafBit);
            while (cnt < node_t::kData && lo32(rc.img(s)[cnt]) != kEmpty) ++cnt;
          }
          all_internal = all_internal && __builtin_amdgcn_ballot_w64(!internal) == 0;
          uint32_t incl = cnt;  // inclusive scan over the wave
          for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t up = (uint32_t)__builtin_amdgcn_ds_bpermute((int)((lane >= d ? lane - d : lane) << 2), (int)incl);
            if (lane >= d) incl += up;
          }
          if (s < e) rc.first_child[s] = (uint16_t)(next + incl - cnt);
          next += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        }
        const bool fits = all_internal && next <= rc_t::kMaxNodes && next > e;
        for (uint32_t s = b + lane; s < e; s += 64) {
          if (!fits) {
            rc.first_child[s] = rc_t::kNoChild;
            continue;
          }
          const uint32_t fc = rc.first_child[s];
          for (uint32_t l = 0; l < node_t::kData; ++l) {
            const uint64_t pr = rc.img(s)[l];
            if (lo32(pr) == kEmpty) break;
            rc.index[fc + l] = hi32(pr);
          }
        }
        if (lane == 0) {
          if (fits) {
            rc.count = next;
          } else {
            rc.stop = 1;
            if (!all_internal) rc.count = b;  // a level with leaves is not routed through
          }
        }
      }
      __syncthreads();
      if (rc.stop) break;
      b = e;
      e = rc.count;
    }
  }

  // Routes `key` through the cached levels, one key per lane (no cross-lane
  // operation: any subset of lanes may call it). Returns the first node to
  // read from global memory and, in `parent`, the node it was reached from
  // (kNone after a right move). A torn or stale image only routes left of the
  // right node; the root is returned if an image names no node of the pool.
  template <class RC>
  MVBT_DEV uint32_t route_lane(const RC& rc, uint32_t key, uint32_t& parent) const {
    using rc_t = RC;
    uint32_t s = 0;
    parent = kNone;
    if (rc.count == 0) return root_;
    for (uint32_t step = 0; step < 2 * rc_t::kMaxNodes; ++step) {
      const uint64_t* img = rc.img(s);
      const uint64_t m = img[node_t::kMetaLane];
      uint32_t next;
      if (key >= node_t::high_key(m)) {
        next = node_t::sibling(m);
        parent = kNone;
        if (s + 1 < rc.count && rc.index[s + 1] == next) {
          ++s;
          continue;
        }
      } else {
        uint32_t l = 0;  // last entry whose pivot is <= key (entry 0 routes from the node's low bound)
#pragma unroll
        for (uint32_t e = 1; e < node_t::kData; ++e) {
          const uint32_t k = lo32(img[e]);
          if (k != kEmpty && key >= k) l = e;
        }
        next = hi32(img[l]);
        parent = rc.index[s];
        const uint32_t fc = rc.first_child[s];
        if (fc != rc_t::kNoChild && fc + l < rc.count && rc.index[fc + l] == next) {
          s = fc + l;
          continue;
        }
      }
      if (next < cap_) return next;
      break;
    }
    parent = kNone;
    return root_;
  }

  // ---- find -----------------------------------------------------------------
  template <bool CONC, bool GUARD = CONC>
  MVBT_DEV uint32_t find(uint32_t key, uint32_t ts, const tile_t& t, uint32_t start = kNone) const {
    if (!valid_key(key)) return kEmpty;
    node_t n;
    uint32_t idx = descend<CONC>(key, n, t, nullptr, start);
    return find_in<CONC, GUARD>(key, ts, n, idx, t);
  }
  // A leaf's latest content is rewritten in place (pairs shift on insert and
  // erase; a versioned leaf keeps one version word per epoch), a multi-lane
  // store a concurrent reader can observe half done, and no version word
  // tells the reader. So a concurrent reader of the latest state (every read
  // of a plain tree, ts == kLatest on a versioned one) reads a leaf under its
  // lock: one attempt per pass of the caller's loop, never a spin of its own,
  // because the holder may be another tile of the same wave, which releases
  // within the same pass. Snapshot reads are validated instead (to_version).
  MVBT_DEV bool lock_and_load(uint32_t idx, node_t& n, const tile_t& t) const {
    if (!try_lock(idx, t)) {
      for (uint32_t i = 0; i < 4; ++i) __builtin_amdgcn_s_sleep(2);
      return false;
    }
    n.load(at(idx), t);
    return true;
  }

  // The find once `key` has been routed to leaf `idx`, loaded in `n`.
  template <bool CONC, bool GUARD = CONC>
  MVBT_DEV uint32_t find_in(uint32_t key, uint32_t ts, node_t& n, uint32_t idx, const tile_t& t) const {
    if (CONC && (!VER || ts == kLatest)) {
      for (uint32_t step = 0; idx < cap_ && step < kMaxReadSteps; ++step) {
        if (!lock_and_load(idx, n, t)) continue;  // busy: try again next pass
        const uint64_t m = n.meta(t);
        if (key >= node_t::high_key(m)) {  // split meanwhile: the key moved right
          unlock(idx, t);
          idx = node_t::sibling(m);
          continue;
        }
        const uint32_t v = n.lookup(key, t);
        unlock(idx, t);
        return v;
      }
      if (idx != kNone) flag(kErrCorrupt, t);
      return kEmpty;
    } else {
      if (idx == kNone || !to_version<CONC, GUARD>(n, idx, ts, t)) return kEmpty;
      return n.lookup(key, t);
    }
  }

  // ---- range query [lo, hi) ------------------------------------------------
  // Writes min(count, cap) pairs to out[0..) and returns the true count.
  template <bool CONC, bool GUARD = CONC>
  MVBT_DEV uint32_t range(uint32_t lo, uint32_t hi, uint32_t ts, uint64_t* out, uint32_t cap,
                          const tile_t& t, uint32_t start = kNone) const {
    if (lo == 0) lo = 1;
    if (hi <= lo) return 0;
    node_t n;
    uint32_t idx = descend<CONC>(lo, n, t, nullptr, start);
    return range_in<CONC, GUARD>(lo, hi, ts, n, idx, out, cap, t);
  }
  // The range query once `lo` (>= 1, below hi) has been routed to leaf `idx`,
  // loaded in `n`.
  template <bool CONC, bool GUARD = CONC>
  MVBT_DEV uint32_t range_in(uint32_t lo, uint32_t hi, uint32_t ts, node_t& n, uint32_t idx, uint64_t* out,
                             uint32_t cap, const tile_t& t) const {
    if (CONC && (!VER || ts == kLatest)) {
      // latest leaves read under lock, one at a time (see lock_and_load)
      uint32_t count = 0;
      bool first = true;
      for (uint32_t step = 0; idx < cap_ && step < kMaxReadSteps; ++step) {
        if (!lock_and_load(idx, n, t)) continue;  // busy: try again next pass
        const uint64_t m = n.meta(t);
        if (first && lo >= node_t::high_key(m)) {  // split meanwhile: lo moved right
          unlock(idx, t);
          idx = node_t::sibling(m);
          continue;
        }
        first = false;
        count += n.collect(lo, hi, out, count, cap, t);
        unlock(idx, t);
        if (hi <= node_t::high_key(m) || node_t::sibling(m) == kNoSibling) return count;
        idx = node_t::sibling(m);
      }
      if (idx != kNone) flag(kErrCorrupt, t);
      return count;
    }
    if (idx == kNone || !to_version<CONC, GUARD>(n, idx, ts, t)) return 0;
    uint32_t count = 0;
    for (uint32_t step = 0;; ++step) {
      count += n.collect(lo, hi, out, count, cap, t);
      uint64_t m = n.meta(t);
      uint32_t sib = node_t::sibling(m);
      if (hi <= node_t::high_key(m) || sib == kNoSibling) break;
      if (sib >= cap_ || step >= kMaxReadSteps) {
        flag(kErrCorrupt, t);
        break;
      }
      idx = sib;
      load<CONC>(n, idx, t);
      if (!to_version<CONC, GUARD>(n, idx, ts, t)) break;
    }
    return count;
  }

  // ---- writers ----------------------------------------------------------------
  enum step : uint32_t { kAtRoot = 0, kAtParent = 1, kOutOfMemory = 2, kInserted = 3 };
  static constexpr uint32_t kWriteSame = 0, kWriteNew = 1, kWriteFailed = kEmpty;

  // Splits the full root in place: its pairs move to two new children and the
  // root keeps {pivot0 -> left, pivot -> right} (reference split_as_root,
  // versioned_node.hpp:185-212, with the root never a leaf here).
  MVBT_DEV uint32_t split_root(uint64_t seen_meta, write_ctx& c, const tile_t& t) const {
    uint32_t r = root_;
    if (!try_lock_seen(r, seen_meta, t)) {
      back_off(c);
      return kAtRoot;
    }
    node_t n;
    n.load(at(r), t);
    if (!n.is_full(t)) {
      unlock(r, t);
      return kAtRoot;
    }
    uint32_t left = alloc(c, t);
    uint32_t right = (left == kNone) ? kNone : alloc(c, t);
    if (right == kNone) {
      unlock(r, t);
      return kOutOfMemory;
    }
    uint64_t m = n.meta(t);
    uint32_t pivot = n.split_pivot(t);
    const uint64_t internal_ver = pack(0u, kNone);
    node_t rn = n.upper_half(m, internal_ver, t);
    node_t ln = n;
    ln.keep_lower_half(pivot, right, t);
    ln.set_lock_in_registers(false, t);
    if (VER) ln.set_lane(node_t::kVerLane, internal_ver, t);
    ln.store(at(left), t);
    rn.store(at(right), t);
    order_release();
    uint32_t k0 = lo32(n.lane_pair(0, t));
    uint32_t rk = t.rank();
    if (node_t::data_lane(rk)) {
      n.p = (rk == 0) ? pack(k0, left) : (rk == 1) ? pack(pivot, right) : node_t::kEmptyPair;
    }
    if (node_t::above_leaves(m)) {
      // the root is now above internal nodes: its metadata loses the flag
      // (stored whole while the lock is held; the lock bit stays set)
      if (rk == node_t::kMetaLane) n.p = pack(lo32(m), hi32(m) & ~kAboveLeaves);
      if (!VER || rk != node_t::kVerLane) st_coherent(at(r) + rk, n.p);
    } else {
      n.store_body(at(r), t);
    }
    unlock(r, t);
    return kAtRoot;
  }

  // Splits full internal node `idx` whose parent is `parent` (reference
  // split, versioned_node.hpp:149-178; parent locked and re-checked as in
  // gpu_versioned_blink_tree.hpp:934-1009, but with try_lock and restart).
  MVBT_DEV uint32_t split_internal(uint32_t idx, uint32_t parent, uint32_t key, uint64_t seen_meta,
                                   write_ctx& c, const tile_t& t) const {
    if (!try_lock_seen(idx, seen_meta, t)) {
      back_off(c);
      return kAtParent;
    }
    node_t n;
    n.load(at(idx), t);
    uint64_t m = n.meta(t);
    if (!n.is_full(t) || key >= node_t::high_key(m)) {
      unlock(idx, t);
      return kAtParent;
    }
    if (!try_lock(parent, t)) {
      unlock(idx, t);
      back_off(c);
      return kAtParent;
    }
    node_t pn;
    pn.load(at(parent), t);
    if (pn.is_full(t) || !pn.has_child(idx, t)) {
      unlock(parent, t);
      unlock(idx, t);
      return kAtRoot;
    }
    uint32_t y = alloc(c, t);
    if (y == kNone) {
      unlock(parent, t);
      unlock(idx, t);
      return kOutOfMemory;
    }
    uint32_t pivot = n.split_pivot(t);
    node_t yn = n.upper_half(m, pack(0u, kNone), t);
    yn.store(at(y), t);
    order_release();  // y is complete before anything links to it
    // The left half and the parent's new entry go out in one round: a reader
    // that sees either one reaches the complete y.
    n.keep_lower_half(pivot, y, t);
    n.store(at(idx), t);  // metadata lane keeps the lock bit
    pn.insert(pivot, y, t);
    pn.store_body(at(parent), t);
    unlock2(parent, idx, t);
    return kAtParent;
  }

  // Publishes a new epoch for locked leaf `idx` if its stamp is older than
  // `cur`: copies the current content out, publishes {cur, copy} in the
  // version word, fenced on both sides. Returns false when out of memory.
  MVBT_DEV bool open_epoch(node_t& n, uint32_t idx, write_ctx& c, const tile_t& t) const {
    if constexpr (VER) {
      uint64_t ver = n.version(t);
      if (needs_copy(ver, c, t)) {
        uint32_t cp = alloc(c, t, true);
        if (cp == kNone) return false;
        n.store_unlocked_copy(at(cp), t);
        retire(c, cp, t);
        order_release();
        uint64_t nv = pack(c.cur, cp);
        if (t.leader()) st_coherent(at(idx) + node_t::kVerLane, nv);
        n.set_lane(node_t::kVerLane, nv, t);
        order_release();
      }
    }
    return true;
  }

  // Splits locked full leaf `idx` (loaded in `n`) under locked-and-checked
  // parent; always unlocks `idx`.
  // When the split succeeds, `key` (absent from the full leaf) goes into the
  // half that covers it before either half is stored, and kInserted is
  // returned: the insert that needed the split costs no second descent, lock
  // and store round.
  // Store rounds (each ends with one wait): {epoch copy, right half y};
  // {version word} when a copy was made; {left half, parent's new entry};
  // then both unlocks. The orders that readers rely on hold: the copy before
  // the version word naming it, the version word before the data it guards,
  // y before any link to it.
  MVBT_DEV uint32_t split_leaf(node_t& n, uint32_t idx, uint32_t parent, write_ctx& c, const tile_t& t,
                               uint32_t key, uint32_t value) const {
    if (parent == kNone) {
      unlock(idx, t);
      return kAtRoot;
    }
    if (!try_lock(parent, t)) {
      unlock(idx, t);
      back_off(c);
      return kAtParent;
    }
    node_t pn;
    pn.load(at(parent), t);
    if (pn.is_full(t) || !pn.has_child(idx, t)) {
      unlock(parent, t);
      unlock(idx, t);
      return kAtRoot;
    }
    uint32_t y = alloc(c, t);
    bool copied = false;
    if constexpr (VER) {
      if (y != kNone && needs_copy(n.version(t), c, t)) {  // new epoch (as open_epoch)
        const uint32_t cp = alloc(c, t, true);
        if (cp == kNone) {
          y = kNone;  // an allocated y is simply never linked; counted as leaked
        } else {
          n.store_unlocked_copy(at(cp), t);
          retire(c, cp, t);
          n.set_lane(node_t::kVerLane, pack(c.cur, cp), t);
          copied = true;
        }
      }
    }
    if (y == kNone) {
      unlock2(parent, idx, t);
      return kOutOfMemory;
    }
    uint64_t m = n.meta(t);
    uint32_t pivot = n.split_pivot(t);
    uint64_t ver = VER ? n.version(t) : 0ull;
    node_t yn = n.upper_half(m, ver, t);
    n.keep_lower_half(pivot, y, t);
    if (key >= pivot) yn.insert(key, value, t);
    else n.insert(key, value, t);
    yn.store(at(y), t);
    order_release();
    if constexpr (VER) {
      if (copied) {
        if (t.leader()) st_coherent(at(idx) + node_t::kVerLane, ver);
        order_release();
      }
      if (t.rank() != node_t::kVerLane) st_coherent(at(idx) + t.rank(), n.p);
    } else {
      n.store(at(idx), t);
    }
    pn.insert(pivot, y, t);
    pn.store_body(at(parent), t);
    unlock2(parent, idx, t);
    return kInserted;
  }

  // Shared writer loop. OP = 0 insert/update, 1 erase. Returns kWriteNew
  // (insert of a new key / erase of a present key), kWriteSame (insert that
  // replaced a value / erase of an absent key) or kWriteFailed (reserved key,
  // pool exhausted, corrupted structure; see the error flags).
  // `start` / `start_parent`: a leaf covering the key (or left of it) and its
  // parent from an earlier routing pass (descend_chunk); kNone starts at the
  // root. Both are re-checked under lock like any node reached by descent.
  // A hint can be arbitrarily stale (the tree may have grown by millions of
  // leaves since it was taken, all to be crossed by right moves), so after
  // kHintHops right moves the write descends afresh from the root.
  static constexpr uint32_t kHintHops = 2;
  template <int OP>
  MVBT_DEV uint32_t write(uint32_t key, uint32_t value, write_ctx& c, const tile_t& t,
                          uint32_t start = kNone, uint32_t start_parent = kNone) const {
    if (!valid_key(key)) {
      flag(kErrInvalidKey, t);
      return kWriteFailed;
    }
    uint32_t idx = root_, parent = kNone;
    uint32_t hint_hops = ~0u;  // right moves still allowed from a hint
    if (start < cap_) {
      idx = start;
      parent = start_parent;
      hint_hops = kHintHops;
    }
    node_t n;
    for (uint32_t step = 0;; ++step) {
      if (idx >= cap_ || step >= kMaxWriteSteps) {  // no lock is held here
        flag(kErrCorrupt, t);
        return kWriteFailed;
      }
      // Coherent (L1-bypassing) read: a writer acts on what it reads (split
      // decisions), and a stale L1 image of a hot node would repeat forever.
      n.load(at(idx), t);
      uint64_t m = n.meta(t);
      if (key >= node_t::high_key(m)) {
        idx = node_t::sibling(m);
        parent = kNone;
        if (hint_hops != ~0u && hint_hops-- == 0) {
          idx = root_;
          hint_hops = ~0u;
        }
        continue;
      }
      if (!node_t::is_leaf(m)) {
        if (OP == 0 && n.is_full(t)) {  // proactive split on the way down
          uint32_t s = (idx == root_) ? split_root(m, c, t)
                       : (parent == kNone) ? (uint32_t)kAtRoot
                                           : split_internal(idx, parent, key, m, c, t);
          if (s == kOutOfMemory) return kWriteFailed;
          if (s == kAtParent && parent != kNone) idx = parent;
          else idx = root_;
          parent = kNone;
          continue;
        }
        parent = idx;
        idx = n.child_for(key, t);
        continue;
      }
      if (!try_lock_seen(idx, m, t)) {
        back_off(c);
        continue;  // re-read the same leaf
      }
      n.load(at(idx), t);
      stamp_under_lock(c, t);
      m = n.meta(t);
      if (key >= node_t::high_key(m)) {
        unlock(idx, t);
        idx = node_t::sibling(m);
        parent = kNone;
        if (hint_hops != ~0u && hint_hops-- == 0) {
          idx = root_;
          hint_hops = ~0u;
        }
        continue;
      }
      if (OP == 1) {
        if (!n.contains(key, t)) {
          unlock(idx, t);
          return kWriteSame;
        }
      } else if (n.is_full(t) && !n.contains(key, t)) {
        uint32_t s = split_leaf(n, idx, parent, c, t, key, value);
        if (s == kOutOfMemory) return kWriteFailed;
        if (s == kInserted) {
          c.backoff = c.backoff > 1 ? c.backoff >> 1 : 1u;
          return kWriteNew;
        }
        idx = (s == kAtParent && parent != kNone) ? parent : root_;
        parent = kNone;
        continue;
      }
      if (!open_epoch(n, idx, c, t)) {
        unlock(idx, t);
        return kWriteFailed;
      }
      bool fresh = true;
      if (OP == 0) fresh = n.insert(key, value, t);
      else n.erase(key, t);
      uint32_t rk = t.rank();
      if (node_t::data_lane(rk)) st_coherent(at(idx) + rk, n.p);
      unlock(idx, t);
      c.backoff = c.backoff > 1 ? c.backoff >> 1 : 1u;
      return fresh ? kWriteNew : kWriteSame;
    }
  }

  // Batched leaf phase for a routed chunk (lane j: key j, value j, routed
  // leaf j; `q` the keys to write). Every leaf that receives the chunk's keys
  // is locked at once (one lane-parallel atomic). The locked leaves are then
  // handled kLeafBatch at a time, each step for all of them before the next:
  // reload, apply the keys in registers in lane order, store the version
  // copies, one wait, store the version words (lane-parallel), one wait,
  // store the data lanes; the next batch's reloads are issued before the wait
  // that completes those stores, and then this batch is unlocked
  // (lane-parallel). Per leaf this is the order write() uses (copy -> version
  // word -> data -> unlock), so snapshot readers validate the same way; a
  // chunk pays about three waits per batch of leaves instead of six per key.
  // No lock is waited for here and every lock is released before returning,
  // so the tiles of a wave never wait on each other's locks inside it. Keys
  // whose leaf was not locked, would overflow, or no longer covers the key
  // are returned as a mask for write(); lane j's `result` receives
  // kWriteNew / kWriteSame for the others.
  // 32-pair nodes take 32-key chunks: batches of 12 leaves cover them in
  // three rounds within the same register budget (4 waves per SIMD).
  static constexpr int kLeafBatch = NW == 32 ? 12 : 8;
  template <int OP>
  MVBT_DEV uint64_t write_chunk(uint32_t key, uint32_t value, uint64_t q, uint32_t leaf,
                                uint32_t& result, write_ctx& c, const tile_t& t) const {
    return write_chunk_at<OP>(key, value, q, leaf, c.cur, result, c, t);
  }
  // The same with a stamp per lane (`cur`, equal for the lanes of one leaf):
  // the sorted batched form writes several batches' epochs in one chunk.
  // A tile that tracks the clock (c.track) uses c.cur for every lane.
  // PLANNED: `planned` (per lane, equal for the lanes of one leaf) is the
  // version-copy slot reserved for the leaf's epoch ahead of the launch and
  // already entered in the retire ring, so no copy allocates or retires here
  // (kNone: the pool had no slot for it). PLANNED also takes no leaf locks:
  // in the sorted form's write pass every leaf has one writing tile (the one
  // that owns its group of sorted writes; a split's new right node holds
  // only keys of that group) and no reader runs, while the per-key path and
  // the parents it updates still lock as everywhere else.
  // The sorted form (PLANNED) takes batches of 8 leaves: its chunks hold few
  // leaves, and the smaller batch lifts its write pass to 4 waves per SIMD.
  template <int OP, bool PLANNED = false>
  MVBT_DEV uint64_t write_chunk_at(uint32_t key, uint32_t value, uint64_t q, uint32_t leaf, uint32_t cur,
                                   uint32_t& result, write_ctx& c, const tile_t& t,
                                   uint32_t planned = kNone) const {
    constexpr int LB = PLANNED ? 8 : kLeafBatch;
    if constexpr (NW > 32) {
      return q;  // 64-pair nodes: register budget; per-key path
    } else {
      const uint32_t r = t.rank();
      const uint64_t todo = q & t.ballot(valid_key(key) && leaf < cap_);
      uint64_t left = q & ~todo;
      // group keys by leaf: my_lead = lowest lane routed to the same leaf
      uint64_t lead = 0;
      uint32_t my_lead = 0;
      for (uint64_t m = todo; m;) {
        const uint32_t j = (uint32_t)lowest_bit(m);
        const uint64_t same = todo & t.ballot(leaf == t.shfl(leaf, j));
        lead |= 1ull << j;
        if ((same >> r) & 1) my_lead = j;
        m &= ~same;
      }
      // 1. lock every leaf of the chunk at once (each leader lane its own leaf)
      uint32_t old = kLockBit;
      if (PLANNED) old = 0;  // see above: these leaves have no other writer
      else if ((lead >> r) & 1) old = atomic_or32(lock_word(leaf), kLockBit);
      const uint64_t locked = lead & t.ballot(!(old & kLockBit));
      for (uint64_t m = lead & ~locked; m; m &= m - 1)
        left |= todo & t.ballot(my_lead == (uint32_t)lowest_bit(m));
      if (locked) stamp_under_lock(c, t);  // one clock read covers every leaf locked above
      if (c.track) cur = c.cur;
      uint64_t pending = locked, prev = 0;
      for (;;) {
        uint64_t got = 0;
        {
          uint64_t m = pending;
          for (int i = 0; i < LB && m; ++i) {
            got |= m & (~m + 1);
            m &= m - 1;
          }
        }
        pending &= ~got;
        // 2. reload this batch's leaves
        uint64_t img[LB];
        uint32_t gl[LB];
        {
          uint64_t m = got;
#pragma unroll
          for (int i = 0; i < LB; ++i) {
            gl[i] = m ? (uint32_t)lowest_bit(m) : 0u;
            if (m) img[i] = ld_coherent(at(t.shfl(leaf, gl[i])) + r);
            m &= m - 1;
          }
        }
        // 6 (of the previous batch). unlock its leaves once its data stores
        // are complete: that wait also covers the reloads just issued
        if (prev) {
          order_release();
          if (!PLANNED && ((prev >> r) & 1)) atomic_and32(lock_word(leaf), ~kLockBit);
        }
        if (!got) break;
        const int ng = __builtin_popcountll(got);
        // 3. plan: apply each leaf's keys in registers; a leaf whose keys do
        // not all fit (or that no longer covers one of them) goes per key
        uint32_t ok = 0, changed = 0, need_copy = 0;  // bit i: slot i
#pragma unroll
        for (int i = 0; i < LB; ++i) {
          if (i >= ng) continue;
          node_t n;
          n.p = img[i];
          const uint64_t meta = n.meta(t);
          const uint64_t mem = todo & t.ballot(my_lead == gl[i]);
          bool good = node_t::is_leaf(meta);
          bool chg = false;
          for (uint64_t mm = mem; mm && good; mm &= mm - 1) {
            const uint32_t j = (uint32_t)lowest_bit(mm);
            const uint32_t kj = t.shfl(key, j);
            if (kj >= node_t::high_key(meta)) {
              good = false;
            } else if (OP == 0) {
              if (!n.contains(kj, t) && n.is_full(t)) {
                good = false;
              } else {
                const bool fresh = n.insert(kj, t.shfl(value, j), t);
                if (r == j) result = fresh ? kWriteNew : kWriteSame;
                chg = true;
              }
            } else {
              const bool had = n.erase(kj, t);
              if (r == j) result = had ? kWriteNew : kWriteSame;
              chg |= had;
            }
          }
          if (!good) {
            left |= mem;
          } else {
            ok |= 1u << i;
            if (chg) changed |= 1u << i;
            node_t o;
            o.p = img[i];
            if (VER && chg && needs_copy_at(o.version(t), t.shfl(cur, gl[i]), t)) need_copy |= 1u << i;
          }
        }
        // 4. version copies (new epoch), then their version words
        uint32_t cp = kNone;  // leader lane: copy slot of its leaf
        if constexpr (VER) {
#pragma unroll
          for (int i = 0; i < LB; ++i) {
            if (!((need_copy >> i) & 1)) continue;
            const uint32_t slot = PLANNED ? t.shfl(planned, gl[i]) : alloc(c, t, true);
            if (slot == kNone) {  // pool exhausted: leave the leaf unchanged
              if (PLANNED) flag(kErrPoolExhausted, t);
              ok &= ~(1u << i);
              changed &= ~(1u << i);
              need_copy &= ~(1u << i);
              continue;
            }
            node_t n;
            n.p = img[i];
            n.store_unlocked_copy(at(slot), t);
            if (!PLANNED) retire(c, slot, t);
            if (r == gl[i]) cp = slot;
          }
          if (need_copy) {
            order_release();
            if (cp != kNone) st_coherent(at(leaf) + node_t::kVerLane, pack(cur, cp));
            order_release();
          }
        }
        // 5. data lanes (the plan re-applied to the loaded image)
#pragma unroll
        for (int i = 0; i < LB; ++i) {
          if (!((changed >> i) & 1)) continue;
          node_t n;
          n.p = img[i];
          const uint64_t mem = todo & t.ballot(my_lead == gl[i]);
          for (uint64_t mm = mem; mm; mm &= mm - 1) {
            const uint32_t j = (uint32_t)lowest_bit(mm);
            if (OP == 0) n.insert(t.shfl(key, j), t.shfl(value, j), t);
            else n.erase(t.shfl(key, j), t);
          }
          // The whole node goes out as one full 128-byte line: the metadata
          // lane as loaded under the lock (unchanged while it is held) and
          // the version word as just published.
          // (the shuffles must run on every lane of the tile: a permute reads
          // nothing from a disabled source lane, and the leader may be one)
          const uint32_t target = t.shfl(leaf, gl[i]);
          if constexpr (VER) {
            const uint32_t cpi = t.shfl(cp, gl[i]), curi = t.shfl(cur, gl[i]);
            if (((need_copy >> i) & 1) && r == node_t::kVerLane) n.p = pack(curi, cpi);
          }
          st_coherent(at(target) + r, n.p);
        }
        // keys of leaves that failed for lack of memory go per key (they fail there too)
        for (int i = 0; i < ng; ++i)
          if (!((ok >> i) & 1)) left |= todo & t.ballot(my_lead == gl[i]);
        prev = got;
      }
      return left;
    }
  }

  MVBT_DEV uint32_t insert(uint32_t key, uint32_t value, write_ctx& c, const tile_t& t) const {
    return write<0>(key, value, c, t);
  }
  MVBT_DEV uint32_t erase(uint32_t key, write_ctx& c, const tile_t& t) const {
    return write<1>(key, 0u, c, t);
  }
};

}  // namespace mvbt
