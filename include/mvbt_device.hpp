// mvbt_device.hpp - device-side cooperative API for user HIP kernels, with the
// reference's names and argument lists (owensgroup/MVGpuBTree
// gpu_versioned_blink_tree.hpp:304-559, gpu_blink_tree.hpp:268-390; apis.md
// "Device-side APIs").
//
// Host: mvbt_device_view(tree, &h) -> pass `h` by value to a kernel.
// Device: every lane of a tile of B lanes calls each operation together.
//
//   __global__ void k(mvbt_device_handle h, ...) {
//     GpuBTree::device::versioned_tree<16> tree(h);
//     GpuBTree::device::tile<16> tile;              // 4 tiles per wave64
//     auto allocator = tree.begin_writes(tile);     // allocation + retirement context
//     auto reclaimer = tree.reclaimer(allocator);   // the reference's DeviceReclaimer slot
//     tree.cooperative_insert(key, value, tile, allocator, reclaimer);
//     uint32_t s = tree.take_snapshot(tile);
//     uint32_t v = tree.cooperative_find(key, tile, allocator, s, /*concurrent=*/true);
//     tree.end_writes(allocator, tile);             // before the tile exits
//   }
//
// The reference passes a DeviceAllocator (bump or slab allocator) and a
// DeviceReclaimer (DEBR) built by the caller. Here one per-tile context holds
// the allocation magazine and the retirement buffer (begin_writes), so the
// reclaimer argument is a tag naming that context; code written against the
// reference keeps its call lines and replaces the two constructions. Reads
// accept the allocator argument and do not use it.
//
// Reads that run while other tiles or kernels write must pass concurrent =
// true (coherent loads; a snapshot read validates each leaf against its
// version and lock words, a read of the latest version reads each leaf under
// its lock). A snapshot taken inside a kernel is deterministic only for
// writes ordered after it (a later kernel, or writes issued after the
// snapshot's value has been observed). A tile never holds a lock between
// calls, so tiles of one wave may mix reads and writes freely.
#pragma once
#include "../mvgpubtree_amd/csrc/mvbt_tree.hpp"
#include "mvbt.h"

namespace GpuBTree {
namespace device {

template <int B>
using tile = mvbt::tile<B>;
// Per-tile allocation and retirement context (reference DeviceAllocator).
using allocator_context = mvbt::write_ctx;
// Names the retirement half of an allocator_context (reference DeviceReclaimer).
struct reclaimer_context {};

template <int B, bool VER>
struct tree_base {
  mvbt::tree_view<B, VER> v;
  using view_t = mvbt::tree_view<B, VER>;

  __host__ __device__ explicit tree_base(const mvbt_device_handle& h) {
    mvbt::tree_ptrs p;
    p.nodes = static_cast<uint64_t*>(h.nodes);
    p.st = static_cast<mvbt::tree_state*>(h.state);
    p.free_stack = static_cast<uint32_t*>(h.free_stack);
    p.spill = static_cast<uint32_t*>(h.spill);
    p.retire_ring = static_cast<uint32_t*>(h.retire_ring);
    p.retire_cap = h.retire_capacity;
    p.root_ = h.root;
    p.cap_ = h.capacity;
    v = mvbt::tree_view<B, VER>(p);
  }

  __device__ allocator_context begin_writes(const tile<B>& t) const { return v.begin_writes(t); }
  __device__ reclaimer_context reclaimer(allocator_context&) const { return reclaimer_context{}; }
  __device__ void end_writes(allocator_context& c, const tile<B>& t) const { v.finish(c, t); }

  // ---- this library's short forms -------------------------------------------
  // Value of `key` (MVBT_NOT_FOUND when absent) at `snapshot`.
  __device__ uint32_t cooperative_find(uint32_t key, const tile<B>& t, uint32_t snapshot = MVBT_LATEST,
                                       bool concurrent = true) const {
    return concurrent ? v.template find<true>(key, snapshot, t) : v.template find<false>(key, snapshot, t);
  }
  // [lower, upper): writes at most `capacity` pairs to `buffer`, returns the count.
  __device__ uint32_t concurrent_cooperative_range_query(uint32_t lower, uint32_t upper, const tile<B>& t,
                                                         uint32_t snapshot, mvbt_pair* buffer,
                                                         uint32_t capacity, bool concurrent = true) const {
    uint64_t* out = reinterpret_cast<uint64_t*>(buffer);
    return concurrent ? v.template range<true>(lower, upper, snapshot, out, capacity, t)
                      : v.template range<false>(lower, upper, snapshot, out, capacity, t);
  }

  // ---- the reference's argument lists -----------------------------------------
  // true when the key was new.
  __device__ bool cooperative_insert(uint32_t key, uint32_t value, const tile<B>& t,
                                     allocator_context& allocator) const {
    return v.insert(key, value, allocator, t) == view_t::kWriteNew;
  }
  // gpu_blink_tree.hpp:278-303 / gpu_versioned_blink_tree.hpp:347-353: latest version.
  __device__ uint32_t cooperative_find(uint32_t key, const tile<B>& t, allocator_context&,
                                       bool concurrent = false) const {
    return concurrent ? v.template find<true>(key, MVBT_LATEST, t) : v.template find<false>(key, MVBT_LATEST, t);
  }
  // gpu_blink_tree.hpp:358-390 / gpu_versioned_blink_tree.hpp:356-390: true when
  // the key was present. On a versioned tree the erase is versioned (snapshots
  // taken before it keep the key).
  __device__ bool cooperative_erase(uint32_t key, const tile<B>& t, allocator_context& allocator,
                                    bool concurrent = false) const {
    (void)concurrent;  // writers always use coherent accesses and locks
    return v.erase(key, allocator, t) == view_t::kWriteNew;
  }
};

// Reference gpu_versioned_btree device methods.
template <int B = 16>
struct versioned_tree : tree_base<B, true> {
  using base = tree_base<B, true>;
  using base::base;
  using base::cooperative_erase;
  using base::cooperative_find;
  using base::cooperative_insert;
  using base::concurrent_cooperative_range_query;

  // gpu_versioned_blink_tree.hpp:333-344: returns the id s; later writes are not visible at s.
  __device__ uint32_t take_snapshot(const tile<B>& t) const { return this->v.take_snapshot(t); }

  // :306-311, :314-320, :323-330. One insert algorithm here (a leaf is copied
  // once per snapshot epoch), so the in-place and out-of-place forms agree.
  __device__ bool cooperative_insert_in_place(uint32_t key, uint32_t value, const tile<B>& t,
                                              allocator_context& allocator) const {
    return this->cooperative_insert(key, value, t, allocator);
  }
  __device__ bool cooperative_insert_out_of_place(uint32_t key, uint32_t value, const tile<B>& t,
                                                  allocator_context& allocator, reclaimer_context&) const {
    return this->cooperative_insert(key, value, t, allocator);
  }
  __device__ bool cooperative_insert(uint32_t key, uint32_t value, const tile<B>& t, allocator_context& allocator,
                                     reclaimer_context&) const {
    return this->cooperative_insert(key, value, t, allocator);
  }
  // :439-474: value at `timestamp` (MVBT_LATEST: the latest version).
  __device__ uint32_t cooperative_find(uint32_t key, const tile<B>& t, allocator_context&, uint32_t timestamp,
                                       bool concurrent = false) const {
    return concurrent ? this->v.template find<true>(key, timestamp, t)
                      : this->v.template find<false>(key, timestamp, t);
  }
  // :479-501: the find at `timestamp` beside concurrent updates.
  __device__ uint32_t concurrent_cooperative_find(uint32_t key, const tile<B>& t, allocator_context&,
                                                  uint32_t timestamp) const {
    return this->v.template find<true>(key, timestamp, t);
  }
  // :394-436: versioned erase beside concurrent operations.
  __device__ bool cooperative_erase(uint32_t key, const tile<B>& t, allocator_context& allocator,
                                    reclaimer_context&) const {
    return this->v.erase(key, allocator, t) == base::view_t::kWriteNew;
  }
  // :505-559: [lower_bound, upper_bound) at `timestamp`; returns the count and,
  // with a buffer, writes every pair to it (unbounded, as in the reference:
  // the caller sizes the buffer; the overload with a capacity bounds it).
  __device__ uint32_t concurrent_cooperative_range_query(uint32_t lower_bound, uint32_t upper_bound,
                                                         const tile<B>& t, allocator_context&, uint32_t timestamp,
                                                         mvbt_pair* buffer = nullptr) const {
    return this->v.template range<true>(lower_bound, upper_bound, timestamp, reinterpret_cast<uint64_t*>(buffer),
                                        0xFFFFFFFFu, t);
  }
};

// Reference gpu_blink_tree device methods (latest version only).
template <int B = 16>
struct blink_tree : tree_base<B, false> {
  using base = tree_base<B, false>;
  using base::base;
  // gpu_blink_tree.hpp:307-355: [lower_bound, upper_bound); every pair is
  // written to the buffer (unbounded, as in the reference).
  __device__ uint32_t cooperative_range_query(uint32_t lower_bound, uint32_t upper_bound, const tile<B>& t,
                                              allocator_context&, mvbt_pair* buffer = nullptr,
                                              bool concurrent = false) const {
    uint64_t* out = reinterpret_cast<uint64_t*>(buffer);
    return concurrent ? this->v.template range<true>(lower_bound, upper_bound, MVBT_LATEST, out, 0xFFFFFFFFu, t)
                      : this->v.template range<false>(lower_bound, upper_bound, MVBT_LATEST, out, 0xFFFFFFFFu, t);
  }
};

}  // namespace device
}  // namespace GpuBTree
