// Common definitions for the MI355X (gfx950) multiversion B-link tree:
// sentinels, memory-order primitives and the wave64 tile abstraction.
//
// Memory-order primitives restate the roles of the reference's
// cuda_memory<T>::load/store orders (include/memory_utils.hpp:21-197) for the
// AMDGPU memory model:
//   * "coherent" node loads are agent-scope relaxed atomic loads. On gfx950 the
//     backend emits them as `global_load ... sc1`, which does not hit in the
//     per-CU vector L1 (non-coherent across CUs), so a reader never consumes a
//     stale node from L1.
//   * node stores that may race with readers are agent-scope relaxed atomic
//     stores (`global_store ... sc1`).
//   * publication order between such stores (copy -> version word -> data ->
//     unlock) is enforced with order_release(): a workgroup-scope fence plus
//     `s_waitcnt vmcnt(0)`, i.e. every earlier store has completed at the
//     coherence point before the next one issues. An agent-scope release fence
//     would instead emit `buffer_wbl2 sc1` (write back the whole XCD L2), which
//     measured ~60 us per fence under load on MI355X; coherence here comes from
//     the sc1 accesses themselves, not from cache maintenance. (rocPRIM's
//     decoupled look-back uses the same scheme on gfx942/gfx950.)
//   * "weak" loads are plain loads, used only by readers when no writer runs
//     concurrently (a kernel boundary orders memory).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MVBT_DEV __device__ __forceinline__
#define MVBT_HD __host__ __device__ __forceinline__

namespace mvbt {

// Sentinels (reference: pair_type.hpp:21-64 default {MAX,MAX}; miss value
// versioned_node.hpp:267-271; reserved key 0, gpu_versioned_blink_tree.hpp:1460).
constexpr uint32_t kEmpty = 0xFFFFFFFFu;     // empty key / empty value / miss result
constexpr uint32_t kLatest = 0xFFFFFFFFu;    // snapshot id meaning "latest version"
constexpr uint32_t kNone = 0xFFFFFFFFu;      // no older version
constexpr uint32_t kLockBit = 1u << 31;      // metadata value word: lock
constexpr uint32_t kLeafBit = 1u << 30;      // metadata value word: leaf
constexpr uint32_t kAboveLeaves = 1u << 29;  // metadata value word: internal node whose children are leaves
constexpr uint32_t kPtrMask = (1u << 29) - 1;
constexpr uint32_t kNoSibling = kPtrMask;    // right-most node of a level
constexpr uint32_t kMaxNodes = kPtrMask;     // node indices must fit 29 bits

MVBT_HD bool valid_key(uint32_t k) { return k != 0u && k != kEmpty; }

// Error flags accumulated on the device (tree_state::error_flags).
enum : uint32_t {
  kErrPoolExhausted = 1u << 0,   // node pool out of slots
  kErrRetireOverflow = 1u << 1,  // retire ring full (copies leaked, not freed)
  kErrInvalidKey = 1u << 2,      // key 0 or 0xFFFFFFFF passed to insert/erase
  kErrVersionGap = 1u << 3,      // version chain ended above the snapshot
  kErrRangeOverflow = 1u << 4,   // a range query produced more than L pairs
  kErrCorrupt = 1u << 5,         // a pointer left the pool or a loop exceeded its bound
  kErrSnapshotId = 1u << 6,      // a snapshot id returned without waiting differs from the clock
  kErrStampOrder = 1u << 7,      // a fixed-stamp write met a leaf stamped after it (overlapping writers)
  kErrBadBatch = 1u << 8,        // a batched call was rejected: a batch writes one key twice, an
                                 // unknown op type, or offsets that disagree with the op count
};

// Loop bounds: every device loop terminates even on a corrupted structure.
constexpr uint32_t kMaxReadSteps = 1u << 20;   // node visits of one read
constexpr uint32_t kMaxWriteSteps = 1u << 20;  // iterations of one write (incl. lock retries)

// ---------------------------------------------------------------------------
// Memory-order primitives.
MVBT_DEV uint64_t ld_weak(const uint64_t* p) { return *p; }
MVBT_DEV uint64_t ld_coherent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MVBT_DEV uint32_t ld32_coherent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MVBT_DEV void st_coherent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MVBT_DEV void st32_coherent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MVBT_DEV void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
// Ordering-only release: all earlier memory operations of this wave complete
// before later ones issue (no cache maintenance).
MVBT_DEV void order_release() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));  // vmcnt(0)
}
MVBT_DEV void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

MVBT_DEV uint32_t atomic_or32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MVBT_DEV uint32_t atomic_and32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MVBT_DEV uint32_t atomic_add32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Compare-and-swap; returns the value found (== expected on success).
MVBT_DEV uint32_t atomic_cas32(uint32_t* p, uint32_t expected, uint32_t desired) {
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expected;
}
MVBT_DEV unsigned long long atomic_add64(unsigned long long* p, unsigned long long v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MVBT_DEV long long atomic_add_s64(long long* p, long long v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

MVBT_DEV uint64_t pack(uint32_t k, uint32_t v) { return (uint64_t)k | ((uint64_t)v << 32); }
MVBT_DEV uint32_t lo32(uint64_t x) { return (uint32_t)x; }
MVBT_DEV uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }

// ---------------------------------------------------------------------------
// Tile: W consecutive lanes of one wave64 cooperate on one operation. The
// reference uses cg::tiled_partition<B> on 32-wide warps; on CDNA a wave holds
// 64/W tiles that execute in lockstep, so tile code never spin-waits on another
// tile (see the op loops in mvbt_tree.hpp).
template <int W>
struct tile {
  static_assert(W == 16 || W == 32 || W == 64, "tile width must be 16, 32 or 64");
  static constexpr int width = W;
  static constexpr uint64_t kMask = (W == 64) ? ~0ull : ((1ull << W) - 1);
  uint32_t lane_;  // lane within the wave

  MVBT_DEV tile() : lane_(__lane_id()) {}
  MVBT_DEV uint32_t rank() const { return lane_ & (W - 1); }
  MVBT_DEV uint32_t base() const { return lane_ & (64 - W); }
  MVBT_DEV bool leader() const { return rank() == 0; }

  // Bitmask over the tile's lanes (bit i = tile lane i).
  MVBT_DEV uint64_t ballot(bool p) const {
    uint64_t m = __builtin_amdgcn_ballot_w64(p);
    if constexpr (W == 64) {
      return m;
    } else {
      return (m >> base()) & kMask;
    }
  }
  MVBT_DEV bool any(bool p) const { return ballot(p) != 0; }
  MVBT_DEV bool all(bool p) const { return ballot(p) == kMask; }

  MVBT_DEV uint32_t shfl(uint32_t v, uint32_t src) const {
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((base() + src) << 2), (int)v);
  }
  MVBT_DEV uint64_t shfl64(uint64_t v, uint32_t src) const {
    return pack(shfl(lo32(v), src), shfl(hi32(v), src));
  }
  // Value from lane rank-d (own value for rank < d).
  MVBT_DEV uint64_t shfl_up64(uint64_t v, uint32_t d) const {
    uint32_t r = rank();
    return shfl64(v, r >= d ? r - d : r);
  }
  // Value from lane rank+d (own value past the end).
  MVBT_DEV uint64_t shfl_down64(uint64_t v, uint32_t d) const {
    uint32_t r = rank();
    return shfl64(v, r + d < (uint32_t)W ? r + d : r);
  }
};

MVBT_DEV int highest_bit(uint64_t m) { return m ? 63 - __builtin_clzll(m) : -1; }
MVBT_DEV int lowest_bit(uint64_t m) { return m ? __builtin_ctzll(m) : -1; }

}  // namespace mvbt
