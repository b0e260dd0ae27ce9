// mvbt_device.hpp - device-side cooperative API for user HIP kernels, with the
// reference's names (owensgroup/MVGpuBTree gpu_versioned_blink_tree.hpp:304-559,
// gpu_blink_tree.hpp:268-390; apis.md "Device-side APIs").
//
// Host: mvbt_device_view(tree, &h) -> pass `h` by value to a kernel.
// Device: every lane of a tile of B lanes calls each operation together.
//
//   __global__ void k(mvbt_device_handle h, ...) {
//     GpuBTree::device::versioned_tree<16> tree(h);
//     GpuBTree::device::tile<16> tile;            // 4 tiles per wave64
//     auto alloc = tree.begin_writes(tile);       // allocator + reclaimer context
//     tree.cooperative_insert(key, value, tile, alloc);
//     uint32_t s = tree.take_snapshot(tile);
//     uint32_t v = tree.cooperative_find(key, tile, s);
//     tree.end_writes(alloc, tile);               // before the tile exits
//   }
//
// Reads that run while other tiles or kernels write must pass concurrent = true
// (coherent loads + leaf version validation). A snapshot taken inside a kernel
// is deterministic only for writes ordered after it (a later kernel, or writes
// issued after the snapshot's value has been observed).
#pragma once
#include "../mvgpubtree_amd/csrc/mvbt_tree.hpp"
#include "mvbt.h"

namespace GpuBTree {
namespace device {

template <int B>
using tile = mvbt::tile<B>;
using allocator_context = mvbt::write_ctx;

template <int B, bool VER>
struct tree_base {
  mvbt::tree_view<B, VER> v;

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
  __device__ void end_writes(allocator_context& c, const tile<B>& t) const { v.finish(c, t); }

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
  // true when the key was new (insert) or present (erase).
  __device__ bool cooperative_insert(uint32_t key, uint32_t value, const tile<B>& t,
                                     allocator_context& c) const {
    return v.insert(key, value, c, t) == mvbt::tree_view<B, VER>::kWriteNew;
  }
  __device__ bool cooperative_erase(uint32_t key, const tile<B>& t, allocator_context& c) const {
    return v.erase(key, c, t) == mvbt::tree_view<B, VER>::kWriteNew;
  }
};

// Reference gpu_versioned_btree device methods.
template <int B = 16>
struct versioned_tree : tree_base<B, true> {
  using tree_base<B, true>::tree_base;
  // Takes a snapshot: returns its id; later writes are not visible at it.
  __device__ uint32_t take_snapshot(const tile<B>& t) const { return this->v.take_snapshot(t); }
};

// Reference gpu_blink_tree device methods (latest version only).
template <int B = 16>
struct blink_tree : tree_base<B, false> {
  using tree_base<B, false>::tree_base;
};

}  // namespace device
}  // namespace GpuBTree
