// Batch kernels (reference btree_kernels.hpp:308-665). Each tile takes a chunk
// of W consecutive operations (lane j holds operation j, so inputs and
// outputs move coalesced), then runs them one at a time cooperatively
// (ballot -> lowest lane -> shuffle -> cooperative op), the reference's
// warp-cooperative work sharing (btree_kernels.hpp:507-517). Chunks are taken
// grid-stride so a launch can be sized to residency instead of to n.
#pragma once
#include "mvbt_tree.hpp"
#include "mvbt_wide.hpp"

namespace mvbt {

// The find kernel: 1024-thread workgroups sharing one large route cache. The
// range kernel keeps 256-thread workgroups: at 4 waves per SIMD a CU then
// holds one 1024-thread workgroup at a time and idles while its slowest wave
// finishes (measured 1.5% slower with the large cache).
constexpr int kFindBlock = 1024;
constexpr uint32_t kFindCacheBytes = 72 * 1024;
constexpr uint32_t kRangeCacheBytes = 16 * 1024;

template <int W>
struct chunking {
  uint64_t first, stride;
  MVBT_DEV chunking() {
    const uint64_t tiles_per_block = blockDim.x / W;
    first = ((uint64_t)blockIdx.x * tiles_per_block + threadIdx.x / W) * W;
    stride = (uint64_t)gridDim.x * tiles_per_block * W;
  }
};

// Creates the root (internal, pivot 0 -> first leaf) and the first, empty leaf.
// Reference initialize_kernel / allocate_root_node (btree_kernels.hpp:525-534,
// gpu_versioned_blink_tree.hpp:1454-1475); here the root is never a leaf.
template <int NW, bool VER>
__global__ void init_kernel(tree_view<NW, VER> tv) {
  using node_t = node<NW, VER>;
  tile<NW> t;
  if (blockIdx.x != 0 || threadIdx.x >= (unsigned)NW) return;
  const uint32_t root = 0, leaf = 1;
  uint32_t r = t.rank();
  uint64_t rp, lp;
  if (r == node_t::kMetaLane) {
    rp = pack(kEmpty, kNoSibling | kAboveLeaves);
    lp = pack(kEmpty, kNoSibling | kLeafBit);
  } else if (VER && r == node_t::kVerLane) {
    rp = pack(0u, kNone);
    lp = pack(tv.st->clock, kNone);
  } else {
    rp = (r == 0) ? pack(0u, leaf) : ~0ull;
    lp = ~0ull;
  }
  tv.at(root)[r] = rp;
  tv.at(leaf)[r] = lp;
}

// Finds and ranges with read tiles of T < NW lanes (P = NW / T pairs per
// lane, mvbt_wide.hpp), for launches in which no writer runs: the same
// pass-based reader, with 64 / T walks per wave. Walks advanced per pass:
// kWideFindGroup at the find kernel's 8 waves per SIMD, kWideRangeGroup in
// the range kernel (whose occupancy is not pinned).
template <int P>
constexpr int kWideFindGroup = 4 / P;
template <int P>
constexpr int kWideRangeGroup = 8 / P;
// Writers' routing with 16-lane sub-tiles: keys advanced per pass per sub-tile.
template <int P>
constexpr int kWideRouteGroup = 8 / P;
// Reads of the mixed calls (each walk also keeps one image for its check):
// walks advanced per pass per sub-tile, within the write phase's registers.
template <int P>
constexpr int kWideConcGroup = 4 / P;

// Plain-tree reads of a chunk beside writers, after their routing. A plain
// leaf is rewritten in place, so it is read under its lock, and RG walks of
// the tile advance together: a pass takes the locks of the walks' leaves
// (test, then set, each walk by its own lane), loads the leaves it holds,
// scans them and releases them. A walk whose leaf is busy tries again on the
// next pass; no walk waits while it holds a lock. Lane j: key `key` (find)
// or range [key, hi) with pairs to its own `out` (at most cap), from leaf
// `leaf`; returns the value or kEmpty, or the count.
constexpr int kLockedReadGroup = 4;
template <bool RANGE, int NW>
MVBT_DEV uint32_t read_locked_chunk(const tree_view<NW, false>& tv, uint32_t key, uint32_t hi, uint32_t leaf,
                                    uint64_t todo, uint64_t* out, uint32_t cap, const tile<NW>& t) {
  using node_t = node<NW, false>;
  constexpr int RG = kLockedReadGroup;
  const uint32_t r = t.rank();
  uint32_t cur = leaf, cnt = 0, first = 1, res = RANGE ? 0u : kEmpty;
  uint32_t err = 0;
  todo &= t.ballot(leaf < tv.cap_);  // a walk whose routing left the pool has flagged it already
  for (int base = 0; base < NW; base += RG) {
    uint64_t act = (todo >> base) & ((1ull << RG) - 1);
    for (uint32_t pass = 0; act != 0; ++pass) {
      if (pass >= kMaxReadSteps) {
        err |= kErrCorrupt;
        break;
      }
      const bool mine = r >= (uint32_t)base && r < (uint32_t)(base + RG) && ((act >> (r - base)) & 1);
      bool got = false;
      if (mine) {
        uint32_t old = ld32_coherent(tv.lock_word(cur));
        if (!(old & kLockBit)) old = atomic_or32(tv.lock_word(cur), kLockBit);
        got = !(old & kLockBit);
      }
      const uint64_t held = (t.ballot(got) >> base) & ((1ull << RG) - 1);
      if (held == 0) {  // every leaf busy: back off before the next pass
        for (uint32_t i = 0; i < 4; ++i) __builtin_amdgcn_s_sleep(2);
        continue;
      }
      node_t n[RG];
#pragma unroll
      for (int j = 0; j < RG; ++j)
        if ((held >> j) & 1) n[j].load(tv.at(t.shfl(cur, (uint32_t)(base + j))), t);
      const uint32_t locked = cur;  // the leaf this lane's walk holds, if got
#pragma unroll
      for (int j = 0; j < RG; ++j) {
        if (!((held >> j) & 1)) continue;
        const uint32_t lj = (uint32_t)(base + j);
        const uint64_t m = n[j].meta(t);
        const uint32_t k = t.shfl(key, lj);
        uint32_t next = kNone;
        bool done = false;
        if (t.shfl(first, lj) && k >= node_t::high_key(m)) {
          next = node_t::sibling(m);  // split meanwhile: the key moved right
        } else if constexpr (RANGE) {
          const uint32_t h = t.shfl(hi, lj);
          uint64_t* oj = reinterpret_cast<uint64_t*>(t.shfl64(reinterpret_cast<uint64_t>(out), lj));
          const uint32_t c = t.shfl(cnt, lj) + n[j].collect(k, h, oj, t.shfl(cnt, lj), cap, t);
          if (r == lj) {
            cnt = c;
            first = 0;
          }
          if (h <= node_t::high_key(m) || node_t::sibling(m) == kNoSibling) done = true;
          else next = node_t::sibling(m);
        } else {
          const uint32_t v = n[j].lookup(k, t);
          if (r == lj) res = v;
          done = true;
        }
        if (!done && next >= tv.cap_) {
          err |= kErrCorrupt;
          done = true;
        }
        if (r == lj && !done) cur = next;
        if (done) act &= ~(1ull << j);
      }
      order_release();  // the loads above have completed before a lock is released
      if (got) atomic_and32(tv.lock_word(locked), ~kLockBit);
    }
  }
  if (RANGE) res = cnt;
  if (err && t.leader()) atomic_or32(&tv.st->error_flags, err);
  return res;
}

// Routing of a chunk's ranges beside writers (MVBT_CONCURRENT): lane j's key
// from `start` (its route-cache node, kNone: the root) to a node read as a
// leaf, the keys of the chunk together (the writers' pass-based routing,
// coherent loads; 16-lane sub-tiles at 32- and 64-pair nodes). Lanes not in
// `go` get an unspecified node and must not use it.
template <int NW, bool VER>
MVBT_DEV uint32_t route_reads(const tree_view<NW, VER>& tv, uint32_t key, bool go, uint32_t start,
                              const tile<NW>& t) {
  uint32_t leaf = kNone, unused_par;
  if constexpr (NW >= 32) {
    tile<16> h;
    descend_chunk_wide<16, true, kWideRouteGroup<NW / 16>>(tv, key, h.ballot(go), leaf, unused_par, h, 0ull, start,
                                                           kNone);
  } else {
    tv.template descend_chunk<true>(key, t.ballot(go), leaf, unused_par, t, 0ull, start, kNone);
  }
  return leaf;
}

template <int NW, bool VER, bool CONC>
__global__ void __launch_bounds__(kFindBlock) __attribute__((amdgpu_waves_per_eu(8)))
find_kernel(tree_view<NW, VER> tv, const uint32_t* keys,
                                                          uint32_t* values, uint64_t n, uint32_t ts) {
  tile<NW> t;
  chunking<NW> c;
  __shared__ route_cache<NW, kFindCacheBytes> rc;
  tv.build_route_cache(rc);
  for (uint64_t base = c.first; base < n; base += c.stride) {
    uint64_t i = base + t.rank();
    bool active = i < n;
    uint32_t key = active ? keys[i] : 0u;
    uint32_t result = kEmpty;
    uint64_t q = t.ballot(active);
    if constexpr (!CONC && NW >= 32) {
      // 32- and 64-pair tiles: a wave holds only two (one) of them, so one
      // walk per tile leaves too few loads in flight; the pass-based engine
      // advances kReadGroup walks of the tile per pass instead.
      using rc_t = route_cache<NW, kFindCacheBytes>;
      result = tv.template read_chunk<false, rc_t, false, tree_view<NW, VER>::kFindGroup>(key, 0u, ts, nullptr, 0u,
                                                                                           q, t, &rc);
    } else {
      uint32_t start = kNone, unused;  // each lane routes its key through the cached levels
      if (active && valid_key(key)) start = tv.route_lane(rc, key, unused);
      // Key by key: a find is one walk to a leaf whose upper levels stay in
      // the L1, and the small register footprint keeps 8 waves per SIMD in
      // flight; at 16-pair tiles the pass-based engine measured slower.
      while (q) {
        uint32_t j = (uint32_t)lowest_bit(q);
        q &= q - 1;
        uint32_t r = tv.template find<CONC>(t.shfl(key, j), ts, t, t.shfl(start, j));
        if (t.rank() == j) result = r;
      }
    }
    if (active) values[i] = result;
  }
}

template <int NW, bool VER, bool CONC>
__global__ void __launch_bounds__(256)
    range_kernel(tree_view<NW, VER> tv, const uint32_t* lower, const uint32_t* upper,
                 uint64_t* out, uint32_t* counts, uint32_t L, uint64_t n, uint32_t ts) {
  tile<NW> t;
  chunking<NW> c;
  __shared__ route_cache<NW, kRangeCacheBytes> rc;
  tv.build_route_cache(rc);
  for (uint64_t base = c.first; base < n; base += c.stride) {
    uint64_t i = base + t.rank();
    bool active = i < n;
    uint32_t lo = active ? lower[i] : 0u;
    uint32_t hi = active ? upper[i] : 0u;
    uint32_t result = 0;
    uint64_t q = t.ballot(active);
    if constexpr (!CONC) {
      // no writer runs: every query of the chunk advances one node per pass
      result = tv.template read_chunk<true>(lo == 0 ? 1u : lo, hi, ts, out ? out + base * (uint64_t)L : nullptr,
                                            L, q, t, &rc);
    } else if constexpr (NW == 32) {
      // beside writers at 32-pair nodes: the chunk's ranges are routed
      // together, then each range reads its leaves (under their locks at the
      // latest state, validated against their version words at a snapshot).
      // 11-13% faster here; at 16-pair nodes the routing's registers cost
      // more than it saves (plain trees 25% slower), so those keep the walk
      // below, range by range.
      const uint32_t k0 = lo == 0 ? 1u : lo;
      const bool go = active && valid_key(k0) && hi > k0;
      uint32_t start = kNone, unused;
      if (go) start = tv.route_lane(rc, k0, unused);
      const uint32_t leaf = route_reads(tv, k0, go, start, t);
      if constexpr (!VER) {  // plain leaves under their locks, four walks a pass
        result = read_locked_chunk<true, NW>(tv, k0, hi, leaf, t.ballot(go), (go && out) ? out + i * (uint64_t)L : nullptr,
                                             L, t);
      } else {
        q = t.ballot(go);
        while (q) {
          uint32_t j = (uint32_t)lowest_bit(q);
          q &= q - 1;
          uint64_t* slot = out ? out + (base + j) * (uint64_t)L : nullptr;
          const uint32_t idx = t.shfl(leaf, j);
          node<NW, VER> n;
          if (ts != kLatest && idx < tv.cap_) tv.template load<true>(n, idx, t);
          uint32_t cnt = tv.template range_in<true>(t.shfl(k0, j), t.shfl(hi, j), ts, n, idx, slot, L, t);
          if (t.rank() == j) result = cnt;
        }
      }
    } else {
      uint32_t start = kNone, unused;
      if (active && valid_key(lo == 0 ? 1u : lo)) start = tv.route_lane(rc, lo == 0 ? 1u : lo, unused);
      while (q) {
        uint32_t j = (uint32_t)lowest_bit(q);
        q &= q - 1;
        uint64_t* slot = out ? out + (base + j) * (uint64_t)L : nullptr;
        uint32_t cnt = tv.template range<CONC>(t.shfl(lo, j), t.shfl(hi, j), ts, slot, L, t, t.shfl(start, j));
        if (t.rank() == j) result = cnt;
      }
    }
    if (active && counts) counts[i] = result;
    if (t.any(active && result > L) && t.leader()) atomic_or32(&tv.st->error_flags, kErrRangeOverflow);
  }
}

template <int NW, bool VER, int T>
__global__ void __launch_bounds__(kFindBlock) __attribute__((amdgpu_waves_per_eu(8)))
find_wide_kernel(tree_view<NW, VER> tv, const uint32_t* keys, uint32_t* values, uint64_t n, uint32_t ts) {
  using rc_t = route_cache<NW, kFindCacheBytes>;
  tile<T> t;
  chunking<T> c;
  __shared__ rc_t rc;
  tv.build_route_cache(rc);
  for (uint64_t base = c.first; base < n; base += c.stride) {
    const uint64_t i = base + t.rank();
    const bool active = i < n;
    const uint32_t key = active ? keys[i] : 0u;
    const uint32_t result =
        read_chunk_wide<T, false, false, kWideFindGroup<NW / T>>(tv, key, 0u, ts, nullptr, 0u, t.ballot(active), t, &rc);
    if (active) values[i] = result;
  }
}

template <int NW, bool VER, int T>
__global__ void __launch_bounds__(256)
    range_wide_kernel(tree_view<NW, VER> tv, const uint32_t* lower, const uint32_t* upper, uint64_t* out,
                      uint32_t* counts, uint32_t L, uint64_t n, uint32_t ts) {
  using rc_t = route_cache<NW, kRangeCacheBytes>;
  tile<T> t;
  chunking<T> c;
  __shared__ rc_t rc;
  tv.build_route_cache(rc);
  for (uint64_t base = c.first; base < n; base += c.stride) {
    const uint64_t i = base + t.rank();
    const bool active = i < n;
    const uint32_t lo = active ? lower[i] : 0u;
    const uint32_t hi = active ? upper[i] : 0u;
    const uint32_t result = read_chunk_wide<T, true, false, kWideRangeGroup<NW / T>>(
        tv, lo == 0 ? 1u : lo, hi, ts, out ? out + base * (uint64_t)L : nullptr, L, t.ballot(active), t, &rc);
    if (active && counts) counts[i] = result;
    if (t.any(active && result > L) && t.leader()) atomic_or32(&tv.st->error_flags, kErrRangeOverflow);
  }
}

template <int NW, bool VER, int OP>
__global__ void __launch_bounds__(256) write_kernel(tree_view<NW, VER> tv, const uint32_t* keys,
                                                    const uint32_t* values, uint64_t n) {
  tile<NW> t;
  chunking<NW> c;
  using node_t = node<NW, VER>;
  // The top levels in LDS at 16- and 32-pair nodes: the cache saves each
  // routing group the top passes (at 16-pair nodes, one chunk per tile, it
  // measured 2-4% on 500K-key inserts at 40M keys, no loss on erases). The
  // images go stale as this launch's splits land, so the cache is used only
  // when the tree's nodes can hold four times the batch (a stale image then
  // costs a right move or two); a launch that grows a small tree routes
  // from the root.
  __shared__ route_cache<NW, 16384> rc;
  if constexpr (NW <= 32) tv.build_route_cache(rc, 4 * n / node_t::kData + 1);
  if (c.first >= n) return;
  write_ctx ctx = tv.begin_writes(t);
  ctx.exact = 1;  // ctx.want below is the chunk's exact remaining count
  for (uint64_t base = c.first; base < n; base += c.stride) {
    uint64_t i = base + t.rank();
    bool active = i < n;
    uint32_t key = active ? keys[i] : 0u;
    uint32_t val = (active && OP == 0) ? values[i] : 0u;
    uint64_t q = t.ballot(active);
    // Route the whole chunk first (one latency per level for all its keys),
    // then write key by key starting from the routed leaf.
    uint32_t leaf = kNone, par = kNone;
    uint32_t st0 = kNone, sp0 = kNone;
    if constexpr (NW <= 32) {
      if (active && valid_key(key)) st0 = tv.route_lane(rc, key, sp0);
    }
    if constexpr (NW >= 32) {
      // routed by the tile's 16-lane halves (quarters), NW / 16 pairs per
      // lane: lane j of sub-tile s is lane 16 s + j of the tile, so each
      // key's route lands in its own lane; the write below uses all NW lanes
      tile<16> h;
      const uint64_t qh = h.ballot((q >> t.rank()) & 1);
      descend_chunk_wide<16, true, kWideRouteGroup<NW / 16>>(tv, key, qh, leaf, par, h, qh, st0, sp0);
    } else {
      tv.template descend_chunk<true>(key, q, leaf, par, t, q, st0, sp0);
    }
    // Leaves that take keys of this chunk are written together; the keys it
    // hands back (lock busy, split needed, leaf moved) go one at a time.
    // ctx.want is the chunk's exact count: refills for version copies take
    // the whole chunk's need at once.
    uint32_t res = 0;
    ctx.want = (uint32_t)__builtin_popcountll(q);
    q = tv.template write_chunk<OP>(key, val, q, leaf, res, ctx, t);
    while (q) {
      uint32_t j = (uint32_t)lowest_bit(q);
      ctx.want = (uint32_t)__builtin_popcountll(q);
      q &= q - 1;
      tv.template write<OP>(t.shfl(key, j), OP == 0 ? t.shfl(val, j) : 0u, ctx, t, t.shfl(leaf, j),
                            t.shfl(par, j));
    }
  }
  tv.finish(ctx, t);
}

// Mixed launch (reference concurrent_insert_range_kernel btree_kernels.hpp:150-306,
// concurrent_find_erase_kernel_versioned :31-143, and their blink variants
// :672-860): the writes and reads of one call run in one grid, chunks taken
// grid-stride over [writes | reads]. Versioned trees: the reads see the
// snapshot reserved by mixed_setup_kernel (every write of the call is stamped
// after it), i.e. the linearization in which every read of the call precedes
// every write; the reference's reads take snapshots inside the kernel instead
// and are not deterministic. Plain trees: reads see the latest state, any
// linearization, as in the reference.
// IR = true: writes insert (wkeys, wvals), reads are ranges [rlo, rhi) with
// pairs at pairs + j*L and counts. IR = false: writes erase (wkeys), reads
// find (rlo) into find_out.
__global__ void mixed_setup_kernel(tree_state* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    st->call_snapshot = __hip_atomic_fetch_add(&st->clock, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NW, bool VER, bool IR>
__global__ void __launch_bounds__(256)
    mixed_kernel(tree_view<NW, VER> tv, const uint32_t* wkeys, const uint32_t* wvals, uint64_t nw,
                 const uint32_t* rlo, const uint32_t* rhi, uint64_t nr, uint32_t* find_out, uint64_t* pairs,
                 uint32_t* counts, uint32_t L) {
  tile<NW> t;
  chunking<NW> c;
  const uint64_t n = nw + nr;
  // The top levels in LDS, as in write_kernel at 32-pair nodes: writes and
  // reads start below them (at every node width: on 1M-op calls it saved
  // 6-9% at 16, 32 and 64 pairs per node). The images go stale as
  // the call's splits land, and a stale image routes only left of a key's
  // node (right moves follow), so the cache is used only when the tree's
  // nodes can hold four times the call's writes.
  __shared__ route_cache<NW, 16384> rc;
  tv.build_route_cache(rc, 4 * nw / node<NW, VER>::kData + 1);
  if (c.first >= n) return;
  constexpr int OP = IR ? 0 : 1;
  const uint32_t ts = VER ? ld32_coherent(&tv.st->call_snapshot) : kLatest;
  write_ctx ctx = tv.begin_writes(t);
  ctx.exact = 1;
  if (VER) ctx.cur = ts + 1;
  ctx.track = 0;  // the call's writes are stamped right after its snapshot
  bool overflow = false;
  for (uint64_t base = c.first; base < n; base += c.stride) {
    const uint64_t i = base + t.rank();
    const bool is_w = i < nw, is_r = i >= nw && i < n;
    // writes of the chunk: routed together, leaves written together
    uint64_t q = t.ballot(is_w);
    if (q) {
      const uint32_t key = is_w ? wkeys[i] : 0u;
      const uint32_t val = (IR && is_w) ? wvals[i] : 0u;
      uint32_t leaf = kNone, par = kNone, res = 0;
      uint32_t st0 = kNone, sp0 = kNone;
      if (is_w && valid_key(key)) st0 = tv.route_lane(rc, key, sp0);
      if constexpr (NW >= 32) {  // routed by 16-lane sub-tiles, as in write_kernel
        tile<16> h;
        const uint64_t qh = h.ballot((q >> t.rank()) & 1);
        descend_chunk_wide<16, true, kWideRouteGroup<NW / 16>>(tv, key, qh, leaf, par, h, qh, st0, sp0);
      } else {
        tv.template descend_chunk<true>(key, q, leaf, par, t, q, st0, sp0);
      }
      ctx.want = (uint32_t)__builtin_popcountll(q);
      q = tv.template write_chunk<OP>(key, val, q, leaf, res, ctx, t);
      while (q) {
        const uint32_t j = (uint32_t)lowest_bit(q);
        ctx.want = (uint32_t)__builtin_popcountll(q);
        q &= q - 1;
        tv.template write<OP>(t.shfl(key, j), IR ? t.shfl(val, j) : 0u, ctx, t, t.shfl(leaf, j), t.shfl(par, j));
      }
    }
    // reads of the chunk, beside the writes of other tiles
    q = t.ballot(is_r);
    if (q) {
      const uint64_t ri = is_r ? i - nw : 0;
      const uint32_t lo = is_r ? rlo[ri] : 0u;
      const uint32_t hi = (IR && is_r) ? rhi[ri] : 0u;
      uint32_t result = IR ? 0u : kEmpty;
      const uint32_t k0 = (IR && lo == 0) ? 1u : lo;
      uint32_t start = kNone;  // each lane routes its key through the cached levels
      {
        uint32_t unused;
        if (is_r && valid_key(k0)) start = tv.route_lane(rc, k0, unused);
      }
      if constexpr (VER) {
        // pass-based, by the tile's 16-lane sub-tiles (NW / 16 pairs per
        // lane; at 16-pair nodes the tile itself); a visible live leaf is
        // confirmed by a later read of its version word before it is scanned
        tile<16> hsub;
        const uint64_t qh = hsub.ballot(is_r);
        uint64_t* slot = (IR && is_r && pairs) ? pairs + ri * (uint64_t)L : nullptr;
        result = read_chunk_wide_conc<16, IR, kWideConcGroup<NW / 16>>(tv, k0, hi, ts, slot, L, qh, hsub, start);
      } else {
        // plain tree: the chunk's keys are routed together (RG keys a pass,
        // each walk ending on a node read as a leaf), then they read their
        // leaves under the leaves' locks, a group of walks per pass
        const bool go = is_r && valid_key(k0) && (!IR || hi > k0);
        uint32_t leaf = kNone, unused_par;
        if constexpr (NW >= 32) {
          tile<16> h;
          descend_chunk_wide<16, true, kWideRouteGroup<NW / 16>>(tv, k0, h.ballot(go), leaf, unused_par, h, 0ull,
                                                                 start, kNone);
        } else {
          tv.template descend_chunk<true>(k0, t.ballot(go), leaf, unused_par, t, 0ull, start, kNone);
        }
        uint64_t* slot = (IR && is_r && pairs) ? pairs + ri * (uint64_t)L : nullptr;
        result = read_locked_chunk<IR, NW>(tv, k0, hi, leaf, t.ballot(go), slot, L, t);
      }
      if (IR) {
        if (is_r && counts) counts[ri] = result;
        overflow |= is_r && result > L;
      } else if (is_r) {
        find_out[ri] = result;
      }
    }
  }
  if (t.any(overflow)) tv.flag(kErrRangeOverflow, t);
  tv.finish(ctx, t);
}

// take_snapshot_kernel (btree_kernels.hpp:473-478): one lane reads the retire
// ring position, then advances the clock. Synchronous form: both go to the
// host record `rec`. Stream-ordered form (rec null): the ring position goes
// to the device slot `pos_slot`, read later by reclaim_prep_kernel, and the
// id the host already returned (`expected`) is checked against the clock.
__global__ void snapshot_kernel(tree_state* st, snapshot_record* rec, unsigned long long* pos_slot,
                                uint32_t expected) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  fence_release();
  unsigned long long pos =
      __hip_atomic_load(&st->retire_tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t s = __hip_atomic_fetch_add(&st->clock, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  if (rec) {
    __hip_atomic_store(&rec->ring_pos, pos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&rec->id, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  } else {
    __hip_atomic_store(pos_slot, pos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (s != expected) atomic_or32(&st->error_flags, kErrSnapshotId);
  }
}

// Reclamation, phase 1 (one lane): frees retire-ring entries [head, end) by
// moving them onto the free stack. Records the live count before freeing as a
// peak candidate and, with `series`, appends the epoch to the memory timeline.
MVBT_DEV void reclaim_prep(tree_state* st, unsigned long long end_pos, epoch_stat* series) {
  unsigned long long head = st->retire_head, tail = st->retire_tail;
  unsigned long long end = end_pos < tail ? end_pos : tail;
  if (end < head) end = head;
  unsigned long long count = end - head;
  // The ring holds at least as many entries as the pool has nodes, so it
  // cannot fill; were it ever to, the positions past head + cap were never
  // stored and their slots still name nodes already freed: free nothing
  // from the ring then (those copies leak instead of being freed twice).
  if (count > st->retire_cap || (st->error_flags & kErrRetireOverflow)) count = 0;
  long long top = st->free_top;
  unsigned long long base = top > 0 ? (unsigned long long)top : 0ull;
  unsigned long long spilled = st->spill_count;
  unsigned long long used = st->bump < st->capacity ? st->bump : st->capacity;
  unsigned long long live = used - base - spilled;
  if (live > st->peak_live) st->peak_live = live;
  st->free_top = (long long)(base + count + spilled);
  st->spill_count = 0;
  st->scratch[0] = head;  // first ring position to copy
  st->scratch[1] = count;
  st->scratch[2] = base;
  st->scratch[3] = spilled;
  st->retire_head = end;
  st->reclaimed += count;
  if (series) {
    const unsigned long long e = st->series_count;
    epoch_stat& x = series[e % kSeriesCap];
    x.epoch = e;
    x.clock = st->clock;
    x.live_nodes = live - count;
    x.allocated_total = live - count + st->reclaimed;
    x.freed_nodes = count;
    x.retired_pending = tail - end;
    x.reclaimed_total = st->reclaimed;
    x.capacity_nodes = st->capacity;
    st->series_count = e + 1;
  }
}

__global__ void reclaim_prep_kernel(tree_state* st, unsigned long long end_pos, const unsigned long long* end_from,
                                    epoch_stat* series) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  // the edge recorded on the device by a stream-ordered snapshot
  if (end_from) end_pos = __hip_atomic_load(end_from, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  reclaim_prep(st, end_pos, series);
}

// Stream-ordered memory statistics (same fields as mvbt_memstats).
__global__ void stats_kernel(const tree_state* st, unsigned long long node_bytes,
                             unsigned long long* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long on_stack = (st->free_top > 0 ? (unsigned long long)st->free_top : 0ull) +
                                st->spill_count;
  unsigned long long bump = st->bump < st->capacity ? st->bump : st->capacity;
  unsigned long long live = bump - on_stack;
  out[0] = node_bytes;
  out[1] = st->capacity;
  out[2] = live;
  out[3] = st->peak_live > live ? st->peak_live : live;
  out[4] = st->capacity - bump + on_stack;
  out[5] = st->retire_tail - st->retire_head;
  out[6] = st->reclaimed;
  out[7] = st->retire_tail;
}

// Reclamation, phase 2: grid-stride copy of the freed ring entries and the
// parked spill slots onto the free stack.
__global__ void reclaim_copy_kernel(tree_state* st, const uint32_t* ring, const uint32_t* spill,
                                    uint32_t* free_stack) {
  const unsigned long long start = st->scratch[0], count = st->scratch[1], base = st->scratch[2],
                           spilled = st->scratch[3];
  const unsigned long long mask = st->retire_cap - 1;
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
       i < count + spilled; i += (unsigned long long)gridDim.x * blockDim.x) {
    free_stack[base + i] = i < count ? ring[(start + i) & mask] : spill[i - count];
  }
}

}  // namespace mvbt
