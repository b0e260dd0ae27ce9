// Sorted bulk load of an empty tree (reference gpu_blink_tree sorted-input
// constructor, gpu_blink_tree.hpp:85-136, bulk_build_kernel
// btree_kernels.hpp:915-1025; available there only for the plain tree with the
// bump allocator, here for both tree kinds and every node width).
//
// The tree is laid out level by level with `fill` entries per node:
//   node 0                 the root (fixed, internal), over the top level
//   nodes 1 .. c0          leaves: leaf m holds keys [m*fill, (m+1)*fill)
//   then levels 1 .. top   internal nodes, node m of level k over children
//                          [m*fill, (m+1)*fill) of level k-1
// A node's first entry routes keys down from its low bound (pivot 0 for the
// left-most node of a level), its high key is the low bound of its right
// sibling, and the last node of a level has no sibling and high key kEmpty:
// the shape the insert path's splits produce, so later writes continue on it.
// Each tile writes one node with one coalesced store; no node depends on
// another, so the whole build is one grid-stride pass over the nodes.
#pragma once
#include "mvbt_tree.hpp"

namespace mvbt {

constexpr uint32_t kMaxBulkLevels = 34;  // fill >= 2 and n < 2^32 need at most 32

struct bulk_plan {
  unsigned long long n;      // keys
  unsigned long long total;  // nodes written, root included (= new bump frontier)
  uint32_t fill;             // entries per leaf and per internal node
  uint32_t levels;           // levels below the root; level 0 = leaves
  unsigned long long count[kMaxBulkLevels];  // nodes of each level
  unsigned long long base[kMaxBulkLevels];   // index of each level's first node
  unsigned long long span[kMaxBulkLevels];   // keys under one full node of each level
};

// Host: the level plan for n keys, or false when `root_cap` entries cannot
// hold the top level within kMaxBulkLevels levels.
inline bool make_bulk_plan(unsigned long long n, uint32_t fill, uint32_t root_cap, bulk_plan* p) {
  *p = bulk_plan{};
  p->n = n;
  p->fill = fill;
  unsigned long long c = (n + fill - 1) / fill, span = fill, base = 1;
  uint32_t lv = 0;
  for (;;) {
    if (lv >= kMaxBulkLevels) return false;
    p->count[lv] = c;
    p->base[lv] = base;
    p->span[lv] = span;
    base += c;
    ++lv;
    if (c <= root_cap) break;
    c = (c + fill - 1) / fill;
    span *= fill;
  }
  p->levels = lv;
  p->total = base;
  return true;
}

// Strictly ascending, no reserved key: *bad receives 1 otherwise.
__global__ void bulk_check_kernel(const uint32_t* keys, unsigned long long n, unsigned long long* bad) {
  bool b = false;
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x) {
    const uint32_t k = keys[i];
    b |= (k == 0u || k == kEmpty || (i + 1 < n && keys[i + 1] <= k));
  }
  if (__builtin_amdgcn_ballot_w64(b) != 0 && __lane_id() == 0) *bad = 1ull;
}

template <int NW, bool VER>
__global__ void __launch_bounds__(256) bulk_build_kernel(tree_view<NW, VER> tv, const uint32_t* keys,
                                                         const uint32_t* values, bulk_plan plan) {
  using node_t = node<NW, VER>;
  tile<NW> t;
  const uint32_t r = t.rank();
  const uint32_t fill = plan.fill;
  const unsigned long long tiles = (unsigned long long)gridDim.x * (blockDim.x / NW);
  const uint32_t stamp = VER ? ld32_coherent(&tv.st->clock) : 0u;
  for (unsigned long long id = blockIdx.x * (unsigned long long)(blockDim.x / NW) + threadIdx.x / NW;
       id < plan.total; id += tiles) {
    uint64_t p = node_t::kEmptyPair;
    if (id == 0) {
      // root: one entry per node of the top level
      const uint32_t top = plan.levels - 1;
      if (r < plan.count[top]) {
        const unsigned long long first = (unsigned long long)r * plan.span[top];
        p = pack(r == 0 ? 0u : keys[first], (uint32_t)(plan.base[top] + r));
      } else if (r == node_t::kMetaLane) {
        p = pack(kEmpty, kNoSibling | (plan.levels == 1 ? kAboveLeaves : 0u));
      } else if (VER && r == node_t::kVerLane) {
        p = pack(0u, kNone);
      }
    } else {
      uint32_t lv = 0;
      while (id >= plan.base[lv] + plan.count[lv]) ++lv;
      const unsigned long long m = id - plan.base[lv];
      const bool last = m + 1 == plan.count[lv];
      const uint32_t sib = last ? kNoSibling : (uint32_t)(id + 1);
      if (lv == 0) {
        const unsigned long long first = m * fill;
        const unsigned long long cnt = last ? plan.n - first : fill;
        if (r < cnt) {
          p = pack(keys[first + r], values[first + r]);
        } else if (r == node_t::kMetaLane) {
          p = pack(last ? kEmpty : keys[first + fill], sib | kLeafBit);
        } else if (VER && r == node_t::kVerLane) {
          p = pack(stamp, kNone);
        }
      } else {
        const unsigned long long below = plan.count[lv - 1], span = plan.span[lv - 1];
        const unsigned long long c0 = m * fill;
        const unsigned long long cnt = last ? below - c0 : fill;
        if (r < cnt) {
          const unsigned long long c = c0 + r;
          p = pack(c == 0 ? 0u : keys[c * span], (uint32_t)(plan.base[lv - 1] + c));
        } else if (r == node_t::kMetaLane) {
          p = pack(last ? kEmpty : keys[(c0 + fill) * span], sib | (lv == 1 ? kAboveLeaves : 0u));
        } else if (VER && r == node_t::kVerLane) {
          p = pack(0u, kNone);
        }
      }
    }
    st_coherent(tv.at((uint32_t)id) + r, p);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) tv.st->bump = plan.total;
}

}  // namespace mvbt
