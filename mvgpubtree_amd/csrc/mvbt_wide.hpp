// Read tiles narrower than the node: T lanes read an NW-pair node, P = NW / T
// consecutive pairs per lane (one 16-byte load per two pairs), so a wave
// holds 64 / T walks instead of 64 / NW. The reference ties its tile to the
// node (cg::tiled_partition<branching_factor>, btree_kernels.hpp:57,476,492);
// here the tile width of the read launches is a separate knob beside node
// width and waves per CU. Writers keep one pair per lane (node<NW, VER>):
// their in-register shifts and split halves are lane moves of whole pairs.
//
// The layout in memory is the one of mvbt_node.hpp; pair g of a node sits in
// lane g / P, slot g % P (P = 1, one pair per lane, serves the mixed calls'
// reader at 16-pair nodes). Two users: reads of launches in which no writer
// runs (snapshot reads of the versioned tree, plain-tree reads without
// MVBT_CONCURRENT), through the pass-based reader below, which follows
// tree_view::read_chunk step for step; and the routing of writers
// (descend_chunk_wide, coherent loads), whose hints are checked under the
// leaf lock like those of tree_view::descend_chunk.
#pragma once
#include "mvbt_tree.hpp"

namespace mvbt {

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

template <int NW, bool VER, int T>
struct wnode {
  static_assert(T == 16 || T == 32 || T == 64, "read tile width must be 16, 32 or 64");
  static_assert(NW % T == 0 && NW / T >= 1 && NW / T <= 4, "one, two or four pairs per lane");
  static constexpr int P = NW / T;
  using base_t = node<NW, VER>;
  using tile_t = tile<T>;
  static constexpr uint32_t kData = base_t::kData;
  static constexpr uint32_t kVerPair = NW - 2;

  uint64_t q[P];  // pairs rank * P .. rank * P + P - 1

  MVBT_DEV void load_weak(const uint64_t* n, const tile_t& t) {
    if constexpr (P == 1) {
      q[0] = n[t.rank()];
    } else {
      const u64x2* s = reinterpret_cast<const u64x2*>(n + t.rank() * P);
#pragma unroll
      for (int i = 0; i < P / 2; ++i) {
        const u64x2 v = s[i];
        q[2 * i] = v.x;
        q[2 * i + 1] = v.y;
      }
    }
  }
  // Coherent form for routing beside writers (8-byte agent-scope loads).
  MVBT_DEV void load_coherent(const uint64_t* n, const tile_t& t) {
#pragma unroll
    for (int i = 0; i < P; ++i) q[i] = ld_coherent(n + t.rank() * P + i);
  }
  MVBT_DEV static bool data_pair(uint32_t g) { return g < kData; }
  // metadata: pair NW - 1 (last lane, last slot); version word: pair NW - 2
  MVBT_DEV uint64_t meta(const tile_t& t) const { return t.shfl64(q[P - 1], T - 1); }
  MVBT_DEV uint64_t version(const tile_t& t) const { return t.shfl64(q[kVerPair % P], kVerPair / P); }

  // Child for `key`: value of the last data pair whose pivot <= key (pair 0
  // when none), as node::child_for.
  MVBT_DEV uint32_t child_for(uint32_t key, const tile_t& t) const {
    const uint32_t r = t.rank();
    bool any = false;
    uint32_t v = hi32(q[0]);
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const uint32_t k = lo32(q[s]);
      if (data_pair(r * P + s) && k != kEmpty && key >= k) {
        any = true;
        v = hi32(q[s]);
      }
    }
    const int l = highest_bit(t.ballot(any));
    return t.shfl(v, l < 0 ? 0u : (uint32_t)l);
  }
  // Leaf lookup: value or kEmpty.
  MVBT_DEV uint32_t lookup(uint32_t key, const tile_t& t) const {
    const uint32_t r = t.rank();
    bool f = false;
    uint32_t v = kEmpty;
#pragma unroll
    for (int s = 0; s < P; ++s) {
      if (data_pair(r * P + s) && lo32(q[s]) == key) {
        f = true;
        v = hi32(q[s]);
      }
    }
    const uint64_t m = t.ballot(f);
    return t.shfl(v, m ? (uint32_t)lowest_bit(m) : 0u);  // no match: every lane holds kEmpty
  }
  // Pairs with lo <= key < hi (a contiguous run of the sorted node) go to
  // out[base + j] for j < cap - base; returns how many matched.
  MVBT_DEV uint32_t collect(uint32_t lo, uint32_t hi, uint64_t* out, uint32_t base, uint32_t cap,
                            const tile_t& t) const {
    const uint32_t r = t.rank();
    uint32_t first = P, last = 0;  // this lane's first and last matching slot
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const uint32_t k = lo32(q[s]);
      if (data_pair(r * P + s) && k != kEmpty && k >= lo && k < hi) {
        if (first == (uint32_t)P) first = (uint32_t)s;
        last = (uint32_t)s;
      }
    }
    const bool mine = first < (uint32_t)P;
    const uint64_t m = t.ballot(mine);
    if (!m) return 0;
    const uint32_t lf = (uint32_t)lowest_bit(m), ll = (uint32_t)highest_bit(m);
    const uint32_t f = lf * P + t.shfl(first, lf);
    const uint32_t e = ll * P + t.shfl(last, ll);
    if (mine && out != nullptr) {
#pragma unroll
      for (int s = 0; s < P; ++s) {
        if ((uint32_t)s < first || (uint32_t)s > last) continue;
        const uint32_t pos = base + (r * P + (uint32_t)s - f);
        if (pos < cap) out[pos] = q[s];
      }
    }
    return e - f + 1;
  }
};

// Pass-based reads of a chunk of T keys with T-lane tiles over NW-pair
// nodes (no writer runs during the launch): tree_view::read_chunk with the
// node image spread over P pairs per lane. RANGE = false: find of `lo`,
// returns the value visible at `ts` or kEmpty; RANGE = true: range [lo, hi)
// (lo >= 1), pairs of lane j's query to out + j * cap (LANES: lane j's own
// `out`), returns the true count. LANES: `ts` per lane.
template <int T, bool RANGE, bool LANES, int RG, int NW, bool VER, class RC>
MVBT_DEV uint32_t read_chunk_wide(const tree_view<NW, VER>& tv, uint32_t lo, uint32_t hi, uint32_t ts,
                                  uint64_t* out, uint32_t cap, uint64_t todo, const tile<T>& t, const RC* rc) {
  using wn_t = wnode<NW, VER, T>;
  using node_t = node<NW, VER>;
  enum : uint32_t { kRoute = 0, kArrive = 1, kChain = 2 };
  const uint32_t r = t.rank();
  uint32_t cur = tv.root_, phase = kRoute, cnt = 0;
  uint32_t res = RANGE ? 0u : kEmpty;
  uint32_t err = 0;
  todo &= t.ballot(valid_key(lo) && (!RANGE || hi > lo));
  if (rc && ((todo >> r) & 1)) {
    uint32_t unused;
    cur = tv.route_lane(*rc, lo, unused);
  }
  for (int base = 0; base < T; base += RG) {
    uint64_t active = (todo >> base) & ((1ull << RG) - 1);
    for (uint32_t pass = 0; active != 0; ++pass) {
      if (pass >= kMaxReadSteps) {
        err |= kErrCorrupt;
        break;
      }
      wn_t w[RG];
#pragma unroll
      for (int j = 0; j < RG; ++j) {
        const uint32_t idx = t.shfl(cur, (uint32_t)(base + j));
        if ((active >> j) & 1) w[j].load_weak(tv.at(idx), t);
      }
#pragma unroll
      for (int j = 0; j < RG; ++j) {
        if (!((active >> j) & 1)) continue;
        const uint32_t lj = (uint32_t)(base + j);
        const wn_t& n = w[j];
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
            if (next == kNone) {
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
        if (!done && next >= tv.cap_) {
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
  if (err && t.leader()) atomic_or32(&tv.st->error_flags, err);
  return res;
}

// Pass-based snapshot reads beside writers, for the mixed calls, whose
// writes are all stamped after the call's snapshot `ts`: a write to a leaf
// visible at ts first stores a version copy, then publishes a new version
// word, and only then changes the data. A live leaf visible at ts is
// therefore used only once a later pass has read its version word unchanged:
// the image of the visiting pass is kept (spread over the tile like any
// image, one per walk of the group) and scanned when confirmed, while a
// changed word sends the walk down the leaf's version chain, whose copies
// are immutable once published. Loads are coherent. Lane j: key `lo` (find)
// or range [lo, hi) with pairs to its own `out` (at most cap), as
// read_chunk_wide with LANES; returns the value or kEmpty, or the count.
// `start`: lane j's first node (an internal node from the launch's route
// cache, which can only lie left of the key's path), or kNone for the root.
template <int T, bool RANGE, int RG, int NW>
MVBT_DEV uint32_t read_chunk_wide_conc(const tree_view<NW, true>& tv, uint32_t lo, uint32_t hi, uint32_t ts,
                                       uint64_t* out, uint32_t cap, uint64_t todo, const tile<T>& t,
                                       uint32_t start = kNone) {
  using wn_t = wnode<NW, true, T>;
  using node_t = node<NW, true>;
  enum : uint32_t { kRoute = 0, kArrive = 1, kCheck = 2, kChain = 3 };
  const uint32_t r = t.rank();
  uint32_t cur = start < tv.cap_ ? start : tv.root_, phase = kRoute, cnt = 0, seen_lo = 0, seen_hi = 0;
  uint32_t res = RANGE ? 0u : kEmpty;
  uint32_t err = 0;
  todo &= t.ballot(valid_key(lo) && (!RANGE || hi > lo));
  for (int base = 0; base < T; base += RG) {
    uint64_t active = (todo >> base) & ((1ull << RG) - 1);
    wn_t keep[RG];  // images waiting for their confirmation
    for (uint32_t pass = 0; active != 0; ++pass) {
      if (pass >= kMaxReadSteps) {
        err |= kErrCorrupt;
        break;
      }
      wn_t w[RG];
#pragma unroll
      for (int j = 0; j < RG; ++j) {
        const uint32_t idx = t.shfl(cur, (uint32_t)(base + j));
        if ((active >> j) & 1) w[j].load_coherent(tv.at(idx), t);
      }
#pragma unroll
      for (int j = 0; j < RG; ++j) {
        if (!((active >> j) & 1)) continue;
        const uint32_t lj = (uint32_t)(base + j);
        const uint32_t k = t.shfl(lo, lj);
        const uint32_t h = RANGE ? t.shfl(hi, lj) : 0u;
        uint64_t* oj = RANGE ? reinterpret_cast<uint64_t*>(t.shfl64(reinterpret_cast<uint64_t>(out), lj)) : nullptr;
        uint32_t ph = t.shfl(phase, lj), c = t.shfl(cnt, lj), next = t.shfl(cur, lj);
        const uint64_t ver = w[j].version(t);
        bool done = false;
        // scans an image visible at ts; continues to the sibling it names
        auto visit = [&](const wn_t& n) {
          const uint64_t m = n.meta(t);
          if constexpr (RANGE) {
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
        };
        bool fresh = true;  // w[j] is still to be looked at (a confirmed visit moves on)
        if (ph == kCheck) {
          if (ver == pack(t.shfl(seen_lo, lj), t.shfl(seen_hi, lj))) {
            visit(keep[j]);
            fresh = false;
          } else {
            ph = kArrive;  // written meanwhile: take this read as a fresh arrival
          }
        }
        if (fresh && ph == kRoute) {
          const uint64_t m = w[j].meta(t);
          if (k >= node_t::high_key(m)) next = node_t::sibling(m);
          else if (!node_t::is_leaf(m)) next = w[j].child_for(k, t);
          else ph = kArrive;
        }
        if (!fresh) {
        } else if (ph == kArrive) {
          if (node_t::stamp(ver) > ts) {
            next = node_t::older(ver);
            ph = kChain;
            if (next == kNone) {
              err |= kErrVersionGap;
              done = true;
            }
          } else {  // visible live leaf: read it again next pass before use
            keep[j] = w[j];
            ph = kCheck;
            if (r == lj) {
              seen_lo = lo32(ver);
              seen_hi = hi32(ver);
            }
          }
        } else if (ph == kChain) {
          if (node_t::stamp(ver) > ts) {
            next = node_t::older(ver);
            if (next == kNone) {
              err |= kErrVersionGap;
              done = true;
            }
          } else {
            visit(w[j]);  // a version copy: immutable once published
          }
        }
        if (!done && next >= tv.cap_) {
          err |= kErrCorrupt;
          done = true;
        }
        if (r == lj) {
          cnt = c;
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
  if (err && t.leader()) atomic_or32(&tv.st->error_flags, err);
  return res;
}

// Routing of a chunk with T-lane tiles over NW-pair nodes: the steps of
// tree_view::descend_chunk (one tree level per pass for RG keys of the tile,
// B-link right moves, `early` lanes stopping above the leaves with the leaf
// taken from their parent's entry), node images spread over P pairs per
// lane. CONC: coherent loads (writers of other tiles may run). On return
// lane j holds the leaf (or the node where its walk stopped) for key j and
// the node it was reached from (kNone after a right move), or kNone as leaf
// when the walk left the pool (kErrCorrupt is flagged). A writer's tile of
// NW lanes calls this through its NW / T sub-tiles, lane j of sub-tile s
// being lane s * T + j of the writer's tile, so each key's route lands in
// its own lane.
template <int T, bool CONC, int RG, int NW, bool VER>
MVBT_DEV void descend_chunk_wide(const tree_view<NW, VER>& tv, uint32_t key, uint64_t todo, uint32_t& leaf,
                                 uint32_t& parent, const tile<T>& t, uint64_t early, uint32_t start,
                                 uint32_t start_par) {
  using wn_t = wnode<NW, VER, T>;
  using node_t = node<NW, VER>;
  const uint32_t r = t.rank();
  uint32_t cur = tv.root_, par = kNone;
  if (start < tv.cap_) {
    cur = start;
    par = start_par;
  }
  bool corrupt = false;
  todo &= t.ballot(valid_key(key));
  for (int base = 0; base < T; base += RG) {
    uint64_t active = (todo >> base) & ((1ull << RG) - 1);
    for (uint32_t pass = 0; active != 0; ++pass) {
      if (pass >= kMaxReadSteps) {
        corrupt = true;
        if (r >= (uint32_t)base && r < (uint32_t)(base + RG) && ((active >> (r - base)) & 1)) cur = kNone;
        break;
      }
      wn_t w[RG];
#pragma unroll
      for (int j = 0; j < RG; ++j) {
        const uint32_t idx = t.shfl(cur, (uint32_t)(base + j));
        if ((active >> j) & 1) {
          if constexpr (CONC) w[j].load_coherent(tv.at(idx), t);
          else w[j].load_weak(tv.at(idx), t);
        }
      }
#pragma unroll
      for (int j = 0; j < RG; ++j) {
        if (!((active >> j) & 1)) continue;
        const uint32_t lj = (uint32_t)(base + j);
        const uint32_t idx = t.shfl(cur, lj);  // not yet advanced
        const wn_t& n = w[j];
        const uint32_t k = t.shfl(key, lj);
        const uint64_t m = n.meta(t);
        uint32_t next = kNone, npar = kNone;
        bool done = false, arrived = false;
        if (k >= node_t::high_key(m)) {
          next = node_t::sibling(m);
        } else if (node_t::is_leaf(m)) {
          done = true;
        } else {
          next = n.child_for(k, t);
          npar = idx;
          arrived = node_t::above_leaves(m) && ((early >> lj) & 1);
        }
        if (!done && next >= tv.cap_) {
          corrupt = true;
          done = true;
          if (r == lj) cur = kNone;
        } else if (!done && r == lj) {
          cur = next;
          par = npar;
        }
        if (done || arrived) active &= ~(1ull << j);
      }
    }
  }
  if (corrupt && t.leader()) atomic_or32(&tv.st->error_flags, kErrCorrupt);
  leaf = cur;
  parent = par;
}

}  // namespace mvbt
