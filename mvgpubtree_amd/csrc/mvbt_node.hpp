// Register-resident node view: one node of NW {key,value} pairs is held by a
// tile of NW lanes, lane i owning pair i (8 bytes), so a whole node moves in a
// single coalesced NW*8-byte access (128 B at NW = 16: one CDNA4 L2 line).
//
// Lane roles (reference layout versioned_node.hpp:583-593 / node.hpp:495-505,
// kept because its 8-byte lanes make every metadata word single-copy atomic):
//   lanes [0, kData)   sorted {key, value} in leaves, {pivot, child} in internal
//                      nodes; empty lanes hold {kEmpty, kEmpty} at the end
//   lane NW-2          (versioned only) version word {stamp, older-version index}
//   lane NW-1          metadata {high key, lock | leaf | level-1 | 29-bit right
//                      sibling}; level-1 marks an internal node whose children
//                      are leaves
//
// Search and update are ballot/shuffle based, as in the reference
// (find_next_lane versioned_node.hpp:214-236, get_key_value_from_node :243-271,
// insert :292-322, erase :324-346, get_in_range :349-363, do_split :89-142).
#pragma once
#include "mvbt_common.hpp"

namespace mvbt {

template <int NW, bool VER>
struct node {
  using tile_t = tile<NW>;
  static constexpr uint32_t kMetaLane = NW - 1;
  static constexpr uint32_t kVerLane = NW - 2;  // meaningful only when VER
  static constexpr uint32_t kData = VER ? NW - 2 : NW - 1;
  static constexpr uint32_t kLastData = kData - 1;
  static constexpr uint32_t kHalf = NW / 2;
  static constexpr uint64_t kEmptyPair = ~0ull;

  uint64_t p;  // this lane's pair: low word key, high word value

  MVBT_DEV static bool data_lane(uint32_t r) { return r < kData; }
  MVBT_DEV uint32_t key() const { return lo32(p); }
  MVBT_DEV uint32_t val() const { return hi32(p); }

  // ---- memory ------------------------------------------------------------
  MVBT_DEV void load_weak(const uint64_t* n, const tile_t& t) { p = ld_weak(n + t.rank()); }
  MVBT_DEV void load(const uint64_t* n, const tile_t& t) { p = ld_coherent(n + t.rank()); }
  MVBT_DEV void store(uint64_t* n, const tile_t& t) const { st_coherent(n + t.rank(), p); }
  // Store every lane except the metadata lane (whose lock bit is owned by
  // atomics) and, for versioned nodes, the version word.
  MVBT_DEV void store_body(uint64_t* n, const tile_t& t) const {
    uint32_t r = t.rank();
    if (r != kMetaLane && (!VER || r != kVerLane)) st_coherent(n + r, p);
  }
  // Copy with the lock bit cleared (reference store_unlocked_copy_at,
  // versioned_node.hpp:47-52).
  MVBT_DEV void store_unlocked_copy(uint64_t* n, const tile_t& t) const {
    uint64_t q = p;
    if (t.rank() == kMetaLane) q &= ~((uint64_t)kLockBit << 32);
    st_coherent(n + t.rank(), q);
  }

  // ---- metadata -----------------------------------------------------------
  MVBT_DEV uint64_t lane_pair(uint32_t r, const tile_t& t) const { return t.shfl64(p, r); }
  MVBT_DEV uint64_t meta(const tile_t& t) const { return lane_pair(kMetaLane, t); }
  MVBT_DEV uint64_t version(const tile_t& t) const { return lane_pair(kVerLane, t); }
  MVBT_DEV static uint32_t high_key(uint64_t meta) { return lo32(meta); }
  MVBT_DEV static uint32_t sibling(uint64_t meta) { return hi32(meta) & kPtrMask; }
  MVBT_DEV static bool is_leaf(uint64_t meta) { return (hi32(meta) & kLeafBit) != 0; }
  // Internal node whose children are leaves (level 1). A node's level never
  // changes except the root's, which clears the flag when it first splits.
  MVBT_DEV static bool above_leaves(uint64_t meta) { return (hi32(meta) & kAboveLeaves) != 0; }
  MVBT_DEV static bool is_locked(uint64_t meta) { return (hi32(meta) & kLockBit) != 0; }
  MVBT_DEV static uint32_t stamp(uint64_t ver) { return lo32(ver); }
  MVBT_DEV static uint32_t older(uint64_t ver) { return hi32(ver); }

  MVBT_DEV bool is_full(const tile_t& t) const {
    return t.ballot(t.rank() == kLastData && key() != kEmpty) != 0;
  }
  MVBT_DEV uint32_t count(const tile_t& t) const {
    return __builtin_popcountll(t.ballot(data_lane(t.rank()) && key() != kEmpty));
  }
  MVBT_DEV void set_lane(uint32_t r, uint64_t v, const tile_t& t) {
    if (t.rank() == r) p = v;
  }
  MVBT_DEV void set_lock_in_registers(bool locked, const tile_t& t) {
    if (t.rank() == kMetaLane) {
      uint32_t w = hi32(p);
      w = locked ? (w | kLockBit) : (w & ~kLockBit);
      p = pack(lo32(p), w);
    }
  }

  // ---- search -------------------------------------------------------------
  // Child for `key` in an internal node: value of the highest data lane whose
  // pivot <= key (lane 0 when none; the left-most node of every level keeps
  // pivot 0 in lane 0).
  MVBT_DEV uint32_t child_for(uint32_t key, const tile_t& t) const {
    uint32_t r = t.rank();
    bool ge = data_lane(r) && this->key() != kEmpty && key >= this->key();
    int l = highest_bit(t.ballot(ge));
    return t.shfl(val(), l < 0 ? 0u : (uint32_t)l);
  }
  // True if `idx` is a child pointer of this internal node.
  MVBT_DEV bool has_child(uint32_t idx, const tile_t& t) const {
    uint32_t r = t.rank();
    return t.ballot(data_lane(r) && key() != kEmpty && val() == idx) != 0;
  }
  // Leaf lookup: value or kEmpty.
  MVBT_DEV uint32_t lookup(uint32_t key, const tile_t& t) const {
    uint32_t r = t.rank();
    uint64_t m = t.ballot(data_lane(r) && this->key() == key);
    if (m == 0) return kEmpty;
    return t.shfl(val(), (uint32_t)lowest_bit(m));
  }
  MVBT_DEV bool contains(uint32_t key, const tile_t& t) const {
    return t.ballot(data_lane(t.rank()) && this->key() == key) != 0;
  }

  // ---- in-register updates -----------------------------------------------
  // Sorted insert; an existing key has its value replaced. Returns true when
  // the key was new. Caller guarantees the node is not full for a new key.
  MVBT_DEV bool insert(uint32_t key, uint32_t value, const tile_t& t) {
    uint32_t r = t.rank();
    uint64_t eq = t.ballot(data_lane(r) && this->key() == key);
    if (eq) {
      if (r == (uint32_t)lowest_bit(eq)) p = pack(key, value);
      return false;
    }
    uint64_t gt = t.ballot(data_lane(r) && this->key() != kEmpty && key > this->key());
    uint32_t pos = (uint32_t)(highest_bit(gt) + 1);
    uint64_t up = t.shfl_up64(p, 1);
    if (data_lane(r)) {
      if (r == pos) p = pack(key, value);
      else if (r > pos) p = up;
    }
    return true;
  }
  // Removes `key`; returns true if it was present.
  MVBT_DEV bool erase(uint32_t key, const tile_t& t) {
    uint32_t r = t.rank();
    uint64_t eq = t.ballot(data_lane(r) && this->key() == key);
    if (!eq) return false;
    uint32_t pos = (uint32_t)lowest_bit(eq);
    uint64_t down = t.shfl_down64(p, 1);
    if (data_lane(r) && r >= pos) p = (r == kLastData) ? kEmptyPair : down;
    return true;
  }

  // ---- range scan -----------------------------------------------------------
  // Pairs with lo <= key < hi are contiguous in a sorted node; they are written
  // to out[base + j] for j < cap - base. Returns how many matched.
  MVBT_DEV uint32_t collect(uint32_t lo, uint32_t hi, uint64_t* out, uint32_t base,
                            uint32_t cap, const tile_t& t) const {
    uint32_t r = t.rank();
    bool in = data_lane(r) && key() != kEmpty && key() >= lo && key() < hi;
    uint64_t m = t.ballot(in);
    if (!m) return 0;
    uint32_t pos = base + (r - (uint32_t)lowest_bit(m));
    if (in && out != nullptr && pos < cap) out[pos] = p;
    return (uint32_t)__builtin_popcountll(m);
  }

  // ---- split ----------------------------------------------------------------
  // Pivot = first key of the upper half.
  MVBT_DEV uint32_t split_pivot(const tile_t& t) const { return lo32(lane_pair(kHalf, t)); }

  // Upper half of this node as a new right sibling (data lanes shifted down
  // by kHalf); its metadata inherits this node's high key and sibling.
  MVBT_DEV node upper_half(uint64_t my_meta, uint64_t right_version, const tile_t& t) const {
    node y;
    uint32_t r = t.rank();
    uint64_t moved = t.shfl_down64(p, kHalf);
    if (data_lane(r)) {
      y.p = (r + kHalf < kData) ? moved : kEmptyPair;
    } else if (r == kMetaLane) {
      y.p = pack(high_key(my_meta), hi32(my_meta) & (kPtrMask | kLeafBit | kAboveLeaves));
    } else {
      y.p = right_version;  // version lane (VER only)
    }
    return y;
  }
  // Keep the lower half; metadata becomes {pivot, right} with flags kept.
  MVBT_DEV void keep_lower_half(uint32_t pivot, uint32_t right, const tile_t& t) {
    uint32_t r = t.rank();
    if (data_lane(r) && r >= kHalf) p = kEmptyPair;
    if (r == kMetaLane) p = pack(pivot, (hi32(p) & (kLockBit | kLeafBit | kAboveLeaves)) | right);
  }
};

}  // namespace mvbt
