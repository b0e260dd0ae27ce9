// gpu_btree.h - header-only C++ host API over the C ABI in mvbt.h, keeping the
// reference's class and method names (owensgroup/MVGpuBTree include/gpu_btree.h,
// include/gpu_versioned_blink_tree.hpp:71-302, include/gpu_blink_tree.hpp:63-265)
// so that host code written against them compiles with hipStream_t in place of
// cudaStream_t. Errors throw GpuBTree::error (the reference terminates).
//
// Device-side cooperative operations (cooperative_find/insert/erase/
// range_query taking a tile and a snapshot id) cannot cross a C ABI; they are
// in mvgpubtree_amd/csrc/mvbt_device.hpp.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "mvbt.h"

namespace GpuBTree {

struct error : std::runtime_error {
  mvbt_status status;
  error(mvbt_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(mvbt_status s, const char* op) {
  if (s != MVBT_OK) throw error(s, std::string(op) + ": " + mvbt_last_error());
}

template <typename Key, typename Value>
struct pair_type {
  Key first;
  Value second;
};

namespace detail {
template <int B, uint32_t Kind>
class tree_base {
 public:
  static_assert(B == 16 || B == 32 || B == 64, "branching factor must be 16, 32 or 64");
  using size_type = uint32_t;
  using key_type = uint32_t;
  using value_type = uint32_t;
  using pair_type = GpuBTree::pair_type<uint32_t, uint32_t>;
  static constexpr int branching_factor = B;
  static constexpr key_type invalid_key = MVBT_NOT_FOUND;
  static constexpr value_type invalid_value = MVBT_NOT_FOUND;

  explicit tree_base(uint64_t pool_nodes = 0, uint64_t retire_capacity = 0, int device = -1) {
    mvbt_config c{};
    c.kind = Kind;
    c.node_width = B;
    c.pool_nodes = pool_nodes;
    c.retire_capacity = retire_capacity;
    c.device = device;
    check(mvbt_create(&c, &h_), "mvbt_create");
  }
  ~tree_base() { mvbt_destroy(h_); }
  tree_base(const tree_base&) = delete;
  tree_base& operator=(const tree_base&) = delete;

  void insert(const key_type* keys, const value_type* values, size_type num_keys,
              hipStream_t stream = 0) {
    check(mvbt_insert(h_, keys, values, num_keys, stream, 0), "insert");
  }
  void find(const key_type* keys, value_type* values, size_type num_keys, hipStream_t stream = 0,
            bool concurrent = false) const {
    check(mvbt_find(h_, keys, values, num_keys, MVBT_LATEST, stream, concurrent ? MVBT_CONCURRENT : 0),
          "find");
  }
  void erase(const key_type* keys, size_type num_keys, hipStream_t stream = 0, bool concurrent = false) {
    (void)concurrent;
    check(mvbt_erase(h_, keys, num_keys, stream, 0), "erase");
  }
  void range_query(const key_type* lower_bound, const key_type* upper_bound, pair_type* result,
                   size_type* counts, size_type average_range_length, size_type num_keys,
                   hipStream_t stream = 0, bool concurrent = false) {
    check(mvbt_range_query(h_, lower_bound, upper_bound, reinterpret_cast<mvbt_pair*>(result), counts,
                           average_range_length, num_keys, MVBT_LATEST, stream,
                           concurrent ? MVBT_CONCURRENT : 0),
          "range_query");
  }
  // Reference concurrent_insert_range / concurrent_find_erase: one launch per
  // call; on versioned trees the reads see the snapshot taken at the call.
  void concurrent_insert_range(const key_type* keys, const value_type* values, size_type num_insertion,
                               const key_type* lower_bound, const key_type* upper_bound, size_type num_ranges,
                               pair_type* result, size_type average_range_length, hipStream_t stream = 0,
                               size_type* counts = nullptr) {
    check(mvbt_concurrent_insert_range(h_, keys, values, num_insertion, lower_bound, upper_bound, num_ranges,
                                       reinterpret_cast<mvbt_pair*>(result), counts, average_range_length, stream),
          "concurrent_insert_range");
  }
  void concurrent_find_erase(const key_type* find_keys, value_type* find_results, size_type num_finds,
                             const key_type* erase_keys, size_type num_erasures, hipStream_t stream = 0) {
    check(mvbt_concurrent_find_erase(h_, find_keys, find_results, num_finds, erase_keys, num_erasures, stream),
          "concurrent_find_erase");
  }
  // Sorted bulk load of this new tree (fill 0 = default entries per node).
  void bulk_build(const key_type* keys, const value_type* values, uint64_t num_keys, uint32_t fill = 0,
                  hipStream_t stream = 0) {
    check(mvbt_bulk_build(h_, keys, values, num_keys, fill, stream), "bulk_build");
  }
  // Host structure check (every allocator); returns false on a broken invariant.
  bool validate_tree_structure(uint64_t* num_keys = nullptr, uint32_t* height = nullptr) {
    return mvbt_validate(h_, num_keys, height) == MVBT_OK;
  }
  mvbt_memstats memory_stats() {
    mvbt_memstats s{};
    check(mvbt_memory_stats(h_, &s), "memory_stats");
    return s;
  }
  double compute_memory_usage() {  // GiB of live nodes
    mvbt_memstats s = memory_stats();
    return double(s.live_nodes) * double(s.node_bytes) / double(1ull << 30);
  }
  uint32_t error_flags() {
    uint32_t f = 0;
    check(mvbt_error_flags(h_, &f), "error_flags");
    return f;
  }
  mvbt_tree* handle() const { return h_; }

 protected:
  mvbt_tree* h_ = nullptr;
};
}  // namespace detail

// Plain concurrent B-link tree (reference gpu_blink_tree<K,V,B,Alloc>).
template <typename Key = uint32_t, typename Value = uint32_t, int B = 16>
class gpu_blink_tree : public detail::tree_base<B, MVBT_PLAIN> {
  static_assert(sizeof(Key) == 4 && sizeof(Value) == 4, "keys and values are 32-bit");

 public:
  using base = detail::tree_base<B, MVBT_PLAIN>;
  using base::base;
  // Reference bulk-build constructor (gpu_blink_tree.hpp:85-136): sorted,
  // distinct keys only (the reference terminates otherwise; this throws).
  // pool_nodes 0 sizes the pool for the build plus as many inserts again.
  gpu_blink_tree(const Key* keys, const Value* values, typename base::size_type num_keys,
                 bool sorted_input = false, hipStream_t stream = 0, uint64_t pool_nodes = 0)
      : base(pool_nodes ? pool_nodes : std::max<uint64_t>(1ull << 22, uint64_t(num_keys) / 4 + (1ull << 20))) {
    if (!sorted_input) throw error(MVBT_ERR_INVALID_ARG, "gpu_blink_tree bulk build: only sorted input is supported");
    base::bulk_build(keys, values, num_keys, 0, stream);
  }
};

// Multiversion B-link tree (reference gpu_versioned_btree<K,V,B,Alloc>).
template <typename Key = uint32_t, typename Value = uint32_t, int B = 16>
class gpu_versioned_btree : public detail::tree_base<B, MVBT_VERSIONED> {
  static_assert(sizeof(Key) == 4 && sizeof(Value) == 4, "keys and values are 32-bit");
  using base = detail::tree_base<B, MVBT_VERSIONED>;

 public:
  using typename base::key_type;
  using typename base::pair_type;
  using typename base::size_type;
  using typename base::value_type;
  using base::base;
  using base::find;
  using base::range_query;

  // in_place is accepted for source compatibility; copies are made once per
  // snapshot epoch per leaf either way.
  void insert(const key_type* keys, const value_type* values, size_type num_keys,
              hipStream_t stream = 0, bool in_place = true) {
    (void)in_place;
    base::insert(keys, values, num_keys, stream);
  }
  void find(const key_type* keys, value_type* values, size_type num_keys, size_type timestamp,
            hipStream_t stream = 0, bool concurrent = false) const {
    check(mvbt_find(this->h_, keys, values, num_keys, timestamp, stream,
                    concurrent ? MVBT_CONCURRENT : 0),
          "find");
  }
  void range_query(const key_type* lower_bound, const key_type* upper_bound, pair_type* result,
                   size_type* counts, size_type average_range_length, size_type num_keys,
                   size_type timestamp, hipStream_t stream = 0, bool concurrent = false) {
    check(mvbt_range_query(this->h_, lower_bound, upper_bound, reinterpret_cast<mvbt_pair*>(result),
                           counts, average_range_length, num_keys, timestamp, stream,
                           concurrent ? MVBT_CONCURRENT : 0),
          "range_query");
  }
  // Returns the snapshot id (reference take_snapshot); stream-ordered, see mvbt_take_snapshot.
  size_type take_snapshot(hipStream_t stream = 0) {
    uint32_t id = 0;
    check(mvbt_take_snapshot(this->h_, stream, &id), "take_snapshot");
    return id;
  }
  void release_snapshot(size_type id) { check(mvbt_release_snapshot(this->h_, id), "release_snapshot"); }
  void reclaim(hipStream_t stream = 0) { check(mvbt_reclaim(this->h_, stream), "reclaim"); }
};

template <typename Key = uint32_t, typename Value = uint32_t, int B = 16>
using gpu_versioned_blink_tree = gpu_versioned_btree<Key, Value, B>;

}  // namespace GpuBTree
