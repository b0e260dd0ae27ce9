// C ABI implementation (include/mvbt.h): handle management, launch sizing,
// snapshot bookkeeping, epoch reclamation and host-side validation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mvbt.h"
#include "mvbt_batch.hpp"
#include "mvbt_bulk.hpp"
#include "mvbt_kernels.hpp"

using namespace mvbt;

namespace {

thread_local std::string g_last_error;

mvbt_status fail(mvbt_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

#define MVBT_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(MVBT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

mvbt_status hip_fail(hipError_t e) { return fail(MVBT_ERR_HIP, std::string("launch: ") + hipGetErrorString(e)); }

constexpr int kBlock = 256;

}  // namespace

struct mvbt_tree {
  mvbt_config cfg{};
  int device = 0;
  uint32_t nw = 16;
  bool versioned = false;
  int num_cus = 256;
  uint64_t* nodes = nullptr;
  tree_state* st = nullptr;
  uint32_t* free_stack = nullptr;
  uint32_t* spill = nullptr;
  uint32_t* retire_ring = nullptr;
  epoch_stat* series = nullptr;         // memory timeline (versioned trees), kSeriesCap entries
  snapshot_record* snap_rec = nullptr;  // pinned host slot (synchronous snapshots)
  std::mutex mu;
  std::map<uint32_t, uint32_t> live;  // live snapshot id -> refcount
  // Epoch edges of live snapshots: the retire-ring position when each was
  // taken, known on the host (slot < 0) or held in device slot snap_pos[slot]
  // (stream-ordered snapshots).
  struct bound {
    uint32_t id;
    uint64_t pos;
    int32_t slot;
  };
  std::vector<bound> bounds;
  // Stream-ordered snapshots. Every kernel that advances the device clock
  // (snapshot, mixed-call and batched setup) is ordered after the previous
  // one through clock_ev, so the clock advances in host call order and
  // `clock` (the host's copy of it) names each snapshot without a round trip.
  // Writes on another stream wait for clock_ev too: a write issued after
  // take_snapshot returns is stamped after the snapshot. The copy is exact
  // until a device view is handed out (user kernels may then take
  // snapshots); from then on take_snapshot reads the id back synchronously.
  static constexpr uint32_t kSnapSlots = 4096;
  unsigned long long* snap_pos = nullptr;  // device: kSnapSlots ring positions
  uint64_t snap_next = 0;                  // next slot (mod kSnapSlots)
  uint32_t clock = 0;
  hipEvent_t clock_ev = nullptr;
  hipStream_t clock_ev_stream = nullptr;
  bool clock_ev_set = false;
  // Free-list maintenance (spill merge, reclamation) must not overlap a
  // writer on another stream: writers wait for the last maintenance event,
  // and maintenance runs only when every other stream's last write is done.
  std::map<hipStream_t, hipEvent_t> write_ev;
  hipEvent_t maint_ev = nullptr;
  bool spill_maybe = false;
  // Launches with launch-fixed write stamps (batched form, mixed calls) run
  // alone among this handle's writers: they wait for every other stream's
  // last write, and writes issued later on other streams wait for them.
  hipEvent_t excl_ev = nullptr;
  hipStream_t excl_stream = nullptr;
  bool excl_set = false;
  // Batched form: device job descriptor, per-batch scratch, conflict table, graph.
  batch_job* job = nullptr;
  uint32_t* b_nwrites = nullptr;
  uint32_t* b_done = nullptr;
  int32_t* b_advance = nullptr;
  uint32_t b_capacity = 0;
  uint32_t* b_hash = nullptr;
  uint64_t b_hash_cap = 0;  // entries (a power of two)
  uint32_t* b_op_batch = nullptr;
  uint64_t b_op_cap = 0;    // entries
  // sorted form (worker mode 2): 8 arrays of b_op_cap entries and the sort's
  // segment counts and histograms, in one allocation
  uint32_t* b_sorted = nullptr;
  uint64_t b_sorted_cap = 0;  // entries per array
  uint32_t worker_waves = 16;
  // 0: one operation per tile; 1: chunks (batch_chunk_worker_kernel);
  // 2: sorted (writes grouped by leaf, then every read at its snapshot);
  // 3: by call size (the persistent worker below kAutoSortedOps operations,
  // the sorted form from there up). eff_mode: the mode of the current call.
  uint32_t worker_mode = 3;
  uint32_t eff_mode = 2;
  // launch graphs: [0] for worker modes 0 and 1, [k] for the sorted form in k windows
  static constexpr uint32_t kMaxWindows = 16;
  hipGraph_t graph[kMaxWindows + 1] = {};
  hipGraphExec_t graph_exec[kMaxWindows + 1] = {};
  hipGraphNode_t graph_setup[kMaxWindows + 1] = {};
  uint64_t window_ops = 1ull << 22;  // sorted form: operations per window (0: one window)
  // lanes per read tile of the launches in which no writer runs (finds and
  // ranges without MVBT_CONCURRENT, the sorted form's read passes); below the
  // node width each lane reads NW / read_tile pairs
  uint32_t read_tile = 0;  // 0: the node width
  bool external_writers = false;  // a device view was handed out
  bool fresh = true;              // no write, snapshot or device view yet (bulk build allowed)

  tree_ptrs ptrs() const {
    tree_ptrs v;
    v.nodes = nodes;
    v.st = st;
    v.free_stack = free_stack;
    v.spill = spill;
    v.retire_ring = retire_ring;
    v.retire_cap = cfg.retire_capacity;
    v.root_ = 0;
    v.cap_ = (uint32_t)cfg.pool_nodes;
    return v;
  }
  template <int NW, bool VER>
  tree_view<NW, VER> view() const {
    return tree_view<NW, VER>(ptrs());
  }
};

namespace {

// RAII device switch.
struct device_guard {
  int prev = -1;
  bool ok = true;
  explicit device_guard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
  }
  ~device_guard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

uint32_t grid_for(const mvbt_tree* t, uint64_t n, uint32_t nw, int waves_per_cu, int block = kBlock) {
  const uint64_t ops_per_block = (uint64_t)block;  // one op per lane per chunk
  uint64_t need = (n + ops_per_block - 1) / ops_per_block;
  uint64_t cap = (uint64_t)t->num_cus * (uint64_t)waves_per_cu * 64 / (uint64_t)block;
  (void)nw;
  if (need < 1) need = 1;
  return (uint32_t)std::min<uint64_t>(need, std::max<uint64_t>(cap, 1));
}

template <int NW, bool VER>
struct impl {
  static mvbt_status init(mvbt_tree* t) {
    hipLaunchKernelGGL((init_kernel<NW, VER>), dim3(1), dim3(64), 0, 0, t->template view<NW, VER>());
    MVBT_HIP(hipGetLastError());
    return MVBT_OK;
  }
  // Read launches with tiles narrower than the node (no writer runs).
  template <int T>
  static bool find_wide(mvbt_tree* t, const uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t ts, hipStream_t s,
                        uint32_t g) {
    if constexpr (T < NW && NW / T <= 4) {
      if (t->read_tile == (uint32_t)T) {
        hipLaunchKernelGGL((find_wide_kernel<NW, VER, T>), dim3(g), dim3(kFindBlock), 0, s,
                           t->template view<NW, VER>(), keys, vals, n, ts);
        return true;
      }
    }
    return false;
  }
  template <int T>
  static bool range_wide(mvbt_tree* t, const uint32_t* lo, const uint32_t* hi, uint64_t* out, uint32_t* counts,
                         uint32_t L, uint64_t n, uint32_t ts, hipStream_t s, uint32_t g) {
    if constexpr (T < NW && NW / T <= 4) {
      if (t->read_tile == (uint32_t)T) {
        hipLaunchKernelGGL((range_wide_kernel<NW, VER, T>), dim3(g), dim3(kBlock), 0, s,
                           t->template view<NW, VER>(), lo, hi, out, counts, L, n, ts);
        return true;
      }
    }
    return false;
  }
  static mvbt_status find(mvbt_tree* t, const uint32_t* keys, uint32_t* vals, uint64_t n,
                          uint32_t ts, hipStream_t s, bool conc) {
    uint32_t g = grid_for(t, n, NW, 32, kFindBlock);
    if (!conc && (find_wide<16>(t, keys, vals, n, ts, s, g) || find_wide<32>(t, keys, vals, n, ts, s, g))) {
      MVBT_HIP(hipGetLastError());
      return MVBT_OK;
    }
    if (conc)
      hipLaunchKernelGGL((find_kernel<NW, VER, true>), dim3(g), dim3(kFindBlock), 0, s,
                         t->template view<NW, VER>(), keys, vals, n, ts);
    else
      hipLaunchKernelGGL((find_kernel<NW, VER, false>), dim3(g), dim3(kFindBlock), 0, s,
                         t->template view<NW, VER>(), keys, vals, n, ts);
    MVBT_HIP(hipGetLastError());
    return MVBT_OK;
  }
  static mvbt_status range(mvbt_tree* t, const uint32_t* lo, const uint32_t* hi, uint64_t* out,
                           uint32_t* counts, uint32_t L, uint64_t n, uint32_t ts, hipStream_t s,
                           bool conc) {
    uint32_t g = grid_for(t, n, NW, 32);
    if (!conc && (range_wide<16>(t, lo, hi, out, counts, L, n, ts, s, g) ||
                  range_wide<32>(t, lo, hi, out, counts, L, n, ts, s, g))) {
      MVBT_HIP(hipGetLastError());
      return MVBT_OK;
    }
    if (conc)
      hipLaunchKernelGGL((range_kernel<NW, VER, true>), dim3(g), dim3(kBlock), 0, s,
                         t->template view<NW, VER>(), lo, hi, out, counts, L, n, ts);
    else
      hipLaunchKernelGGL((range_kernel<NW, VER, false>), dim3(g), dim3(kBlock), 0, s,
                         t->template view<NW, VER>(), lo, hi, out, counts, L, n, ts);
    MVBT_HIP(hipGetLastError());
    return MVBT_OK;
  }
  template <int OP>
  static mvbt_status write(mvbt_tree* t, const uint32_t* keys, const uint32_t* vals, uint64_t n,
                           hipStream_t s) {
    uint32_t g = grid_for(t, n, NW, 32);
    hipLaunchKernelGGL((write_kernel<NW, VER, OP>), dim3(g), dim3(kBlock), 0, s,
                       t->template view<NW, VER>(), keys, vals, n);
    MVBT_HIP(hipGetLastError());
    return MVBT_OK;
  }
  static mvbt_status bulk(mvbt_tree* t, const uint32_t* keys, const uint32_t* vals, const bulk_plan& plan,
                          hipStream_t s) {
    const uint64_t tiles_per_block = kBlock / NW;
    const uint64_t blocks = (plan.total + tiles_per_block - 1) / tiles_per_block;
    const uint64_t cap = (uint64_t)t->num_cus * 32 * 64 / kBlock;
    hipLaunchKernelGGL((bulk_build_kernel<NW, VER>), dim3((uint32_t)std::min(blocks, cap)), dim3(kBlock), 0,
                       s, t->template view<NW, VER>(), keys, vals, plan);
    MVBT_HIP(hipGetLastError());
    return MVBT_OK;
  }
  static constexpr uint32_t data_lanes() { return node<NW, VER>::kData; }
  template <bool IR>
  static mvbt_status mixed(mvbt_tree* t, const uint32_t* wk, const uint32_t* wv, uint64_t nw, const uint32_t* rlo,
                           const uint32_t* rhi, uint64_t nr, uint32_t* fout, uint64_t* pairs, uint32_t* counts,
                           uint32_t L, hipStream_t s) {
    if (VER) hipLaunchKernelGGL(mixed_setup_kernel, dim3(1), dim3(64), 0, s, t->st);
    uint32_t g = grid_for(t, nw + nr, NW, 32);
    hipLaunchKernelGGL((mixed_kernel<NW, VER, IR>), dim3(g), dim3(kBlock), 0, s, t->template view<NW, VER>(), wk, wv,
                       nw, rlo, rhi, nr, fout, pairs, counts, L);
    MVBT_HIP(hipGetLastError());
    return MVBT_OK;
  }
};

template <int NW>
void* worker_fn() {
  return reinterpret_cast<void*>(&batch_worker_kernel<NW>);
}
// The sorted form's width-specific kernels: route, write, find, range.
struct sorted_fns {
  void* route;
  void* write;
  void* find;
  void* range;
};
template <int NW>
sorted_fns sorted_fn(uint32_t read_tile) {
  // routing: 16-lane tiles at 32- and 64-pair nodes, whatever the read tile
  sorted_fns f{reinterpret_cast<void*>(&batch_route_kernel<NW, NW >= 32 ? 16 : NW>),
               reinterpret_cast<void*>(&batch_sorted_write_kernel<NW>),
               reinterpret_cast<void*>(&batch_sorted_find_kernel<NW>),
               reinterpret_cast<void*>(&batch_sorted_range_kernel<NW>)};
  if constexpr (NW >= 32) {
    if (read_tile == 16) {
      f.find = reinterpret_cast<void*>(&batch_sorted_find_kernel<NW, 16>);
      f.range = reinterpret_cast<void*>(&batch_sorted_range_kernel<NW, 16>);
    }
  }
  if constexpr (NW == 64) {
    if (read_tile == 32) {
      f.find = reinterpret_cast<void*>(&batch_sorted_find_kernel<NW, 32>);
      f.range = reinterpret_cast<void*>(&batch_sorted_range_kernel<NW, 32>);
    }
  }
  return f;
}
template <int NW>
void* chunk_worker_fn() {
  return reinterpret_cast<void*>(&batch_chunk_worker_kernel<NW>);
}

struct dispatch_table {
  mvbt_status (*init)(mvbt_tree*);
  mvbt_status (*find)(mvbt_tree*, const uint32_t*, uint32_t*, uint64_t, uint32_t, hipStream_t, bool);
  mvbt_status (*range)(mvbt_tree*, const uint32_t*, const uint32_t*, uint64_t*, uint32_t*, uint32_t,
                       uint64_t, uint32_t, hipStream_t, bool);
  mvbt_status (*insert)(mvbt_tree*, const uint32_t*, const uint32_t*, uint64_t, hipStream_t);
  mvbt_status (*erase)(mvbt_tree*, const uint32_t*, const uint32_t*, uint64_t, hipStream_t);
  void* (*worker)();
  void* (*chunk_worker)();
  sorted_fns (*sorted)(uint32_t read_tile);
  mvbt_status (*bulk)(mvbt_tree*, const uint32_t*, const uint32_t*, const bulk_plan&, hipStream_t);
  uint32_t data_lanes;
  using mixed_fn = mvbt_status (*)(mvbt_tree*, const uint32_t*, const uint32_t*, uint64_t, const uint32_t*,
                                   const uint32_t*, uint64_t, uint32_t*, uint64_t*, uint32_t*, uint32_t, hipStream_t);
  mixed_fn insert_range;
  mixed_fn find_erase;
};

template <int NW, bool VER>
dispatch_table make_table() {
  using I = impl<NW, VER>;
  return dispatch_table{&I::init, &I::find, &I::range, &I::template write<0>, &I::template write<1>,
                        VER ? &worker_fn<NW> : nullptr, VER ? &chunk_worker_fn<NW> : nullptr,
                        VER ? &sorted_fn<NW> : nullptr, &I::bulk,
                        I::data_lanes(),
                        &I::template mixed<true>, &I::template mixed<false>};
}

const dispatch_table* table_for(uint32_t nw, bool versioned) {
  static const dispatch_table t16p = make_table<16, false>();
  static const dispatch_table t16v = make_table<16, true>();
  static const dispatch_table t32p = make_table<32, false>();
  static const dispatch_table t32v = make_table<32, true>();
  static const dispatch_table t64p = make_table<64, false>();
  static const dispatch_table t64v = make_table<64, true>();
  switch (nw) {
    case 16: return versioned ? &t16v : &t16p;
    case 32: return versioned ? &t32v : &t32p;
    case 64: return versioned ? &t64v : &t64p;
    default: return nullptr;
  }
}

bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

// Launches the two reclamation kernels: frees ring entries [head, end_pos)
// (end_pos = 0 frees none) and returns parked spill slots to the free stack.
// A reclamation (record) also appends an epoch to the memory timeline; a
// spill merge alone does not.
mvbt_status launch_maintenance(mvbt_tree* t, hipStream_t s, unsigned long long end_pos,
                               const unsigned long long* end_from = nullptr, bool record = false) {
  hipLaunchKernelGGL(reclaim_prep_kernel, dim3(1), dim3(64), 0, s, t->st, end_pos, end_from,
                     record ? t->series : nullptr);
  MVBT_HIP(hipGetLastError());
  hipLaunchKernelGGL(reclaim_copy_kernel, dim3(t->num_cus * 4), dim3(kBlock), 0, s, t->st,
                     t->retire_ring, t->spill, t->free_stack);
  MVBT_HIP(hipGetLastError());
  if (!t->maint_ev) MVBT_HIP(hipEventCreateWithFlags(&t->maint_ev, hipEventDisableTiming));
  MVBT_HIP(hipEventRecord(t->maint_ev, s));
  t->spill_maybe = false;
  return MVBT_OK;
}

bool other_writers_busy(mvbt_tree* t, hipStream_t s) {
  for (auto& kv : t->write_ev)
    if (kv.first != s && hipEventQuery(kv.second) == hipErrorNotReady) return true;
  return false;
}

// Before a write kernel on stream s (caller holds t->mu).
// Orders work on stream s after the last clock-advancing kernel when that ran
// on another stream.
mvbt_status after_clock(mvbt_tree* t, hipStream_t s) {
  if (t->clock_ev_set && t->clock_ev_stream != s) MVBT_HIP(hipStreamWaitEvent(s, t->clock_ev, 0));
  return MVBT_OK;
}
// Records that a clock-advancing kernel was just launched on stream s.
mvbt_status clock_advanced(mvbt_tree* t, hipStream_t s, uint32_t by) {
  if (!t->clock_ev) MVBT_HIP(hipEventCreateWithFlags(&t->clock_ev, hipEventDisableTiming));
  MVBT_HIP(hipEventRecord(t->clock_ev, s));
  t->clock_ev_stream = s;
  t->clock_ev_set = true;
  t->clock += by;
  return MVBT_OK;
}

mvbt_status before_write(mvbt_tree* t, hipStream_t s) {
  t->fresh = false;
  if (mvbt_status rc = after_clock(t, s)) return rc;
  if (t->excl_set && t->excl_stream != s) MVBT_HIP(hipStreamWaitEvent(s, t->excl_ev, 0));
  if (t->maint_ev) MVBT_HIP(hipStreamWaitEvent(s, t->maint_ev, 0));
  if ((t->spill_maybe || t->external_writers) && !other_writers_busy(t, s))
    return launch_maintenance(t, s, 0);
  return MVBT_OK;
}

// Before a launch with launch-fixed write stamps on stream s: it must not
// overlap another stream's writes (a leaf stamped by them after its own
// stamps would break the stamp order of version chains), so stream s waits
// for every other stream's last write.
mvbt_status isolate(mvbt_tree* t, hipStream_t s) {
  for (auto& kv : t->write_ev)
    if (kv.first != s) MVBT_HIP(hipStreamWaitEvent(s, kv.second, 0));
  return MVBT_OK;
}
// After such a launch: writes issued later on other streams wait for it.
mvbt_status isolated_done(mvbt_tree* t, hipStream_t s) {
  if (!t->excl_ev) MVBT_HIP(hipEventCreateWithFlags(&t->excl_ev, hipEventDisableTiming));
  MVBT_HIP(hipEventRecord(t->excl_ev, s));
  t->excl_stream = s;
  t->excl_set = true;
  return MVBT_OK;
}
// Snapshot ids a call may still reserve: ids run up to 0xFFFFFFFE
// (0xFFFFFFFF names the latest version). The host copy of the clock is exact
// until a device view is handed out; then the device clock is read.
mvbt_status reserve_ids(mvbt_tree* t, uint64_t count) {
  uint64_t clock = t->clock;
  if (t->external_writers) {
    uint32_t c = 0;
    MVBT_HIP(hipDeviceSynchronize());
    MVBT_HIP(hipMemcpy(&c, &t->st->clock, sizeof(c), hipMemcpyDeviceToHost));
    clock = c;
  }
  if (clock + count > (uint64_t)MVBT_LATEST)
    return fail(MVBT_ERR_UNSUPPORTED, "snapshot ids exhausted: the call needs " + std::to_string(count) +
                                          " ids, the clock is at " + std::to_string(clock));
  return MVBT_OK;
}

// After a write kernel on stream s (caller holds t->mu).
mvbt_status after_write(mvbt_tree* t, hipStream_t s) {
  auto it = t->write_ev.find(s);
  if (it == t->write_ev.end()) {
    hipEvent_t e;
    MVBT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    it = t->write_ev.emplace(s, e).first;
  }
  MVBT_HIP(hipEventRecord(it->second, s));
  t->spill_maybe = true;
  return MVBT_OK;
}

mvbt_status check_tree(mvbt_tree* t) {
  if (!t) return fail(MVBT_ERR_INVALID_ARG, "null tree handle");
  return MVBT_OK;
}

struct host_state_view {
  tree_state s;
};

mvbt_status read_state(mvbt_tree* t, tree_state* out) {
  MVBT_HIP(hipDeviceSynchronize());
  MVBT_HIP(hipMemcpy(out, t->st, sizeof(tree_state), hipMemcpyDeviceToHost));
  return MVBT_OK;
}

// Host copy of the node pool plus a walk over the latest tree.
struct host_walk {
  uint32_t nw = 16;
  bool versioned = false;
  std::vector<uint64_t> pool;
  uint64_t nodes = 0;
  uint32_t data_lanes() const { return versioned ? nw - 2 : nw - 1; }
  const uint64_t* at(uint32_t i) const { return pool.data() + (size_t)i * nw; }
  uint64_t meta(uint32_t i) const { return at(i)[nw - 1]; }
  static uint32_t key(uint64_t p) { return (uint32_t)p; }
  static uint32_t val(uint64_t p) { return (uint32_t)(p >> 32); }
};

mvbt_status copy_pool(mvbt_tree* t, host_walk& w) {
  tree_state s;
  mvbt_status rc = read_state(t, &s);
  if (rc != MVBT_OK) return rc;
  w.nw = t->nw;
  w.versioned = t->versioned;
  w.nodes = std::min<uint64_t>(s.bump, s.capacity);
  try {
    w.pool.resize(w.nodes * t->nw);
  } catch (const std::bad_alloc&) {
    return fail(MVBT_ERR_OUT_OF_MEMORY, "host copy of the node pool failed to allocate");
  }
  MVBT_HIP(hipMemcpy(w.pool.data(), t->nodes, w.nodes * t->nw * sizeof(uint64_t),
                     hipMemcpyDeviceToHost));
  return MVBT_OK;
}

// Walks the latest tree level by level. Checks invariants (reference
// validate_tree_structure, gpu_versioned_blink_tree.hpp:1098-1259) and
// appends leaf pairs to `pairs` when non-null.
mvbt_status walk_tree(const host_walk& w, std::vector<mvbt_pair>* pairs, uint64_t* num_keys,
                      uint32_t* height) {
  auto bad = [](const std::string& m) { return fail(MVBT_ERR_DEVICE_FLAGS, "validate: " + m); };
  const uint32_t nw = w.nw, nd = w.data_lanes();
  std::vector<uint32_t> level{0u};  // root
  uint32_t h = 0;
  uint64_t keys = 0;
  bool parent_above = false;  // the level above flagged its children as leaves
  while (true) {
    ++h;
    if (h > 64) return bad("tree deeper than 64 levels");
    // Follow the sibling chain from the left-most node and compare with `level`.
    std::vector<uint32_t> next;
    bool leaf_level = false, above_level = false;
    uint32_t prev_high = 0;
    for (size_t li = 0; li < level.size(); ++li) {
      uint32_t idx = level[li];
      if (idx >= w.nodes) return bad("node index out of range: " + std::to_string(idx));
      uint64_t m = w.meta(idx);
      uint32_t flags = (uint32_t)(m >> 32);
      if (flags & kLockBit) return bad("node " + std::to_string(idx) + " is locked");
      bool leaf = (flags & kLeafBit) != 0;
      bool above = (flags & kAboveLeaves) != 0;
      if (li == 0) {
        leaf_level = leaf;
        above_level = above;
        if (h > 1 && leaf != parent_above) return bad("level-1 flag does not match the level below");
      } else if (leaf != leaf_level) {
        return bad("mixed leaf/internal level");
      } else if (above != above_level) {
        return bad("mixed level-1 flags in one level");
      }
      if (leaf && above) return bad("leaf " + std::to_string(idx) + " flagged as level 1");
      uint32_t high = (uint32_t)m;
      uint32_t sib = flags & kPtrMask;
      if (li + 1 < level.size()) {
        if (sib != level[li + 1]) return bad("sibling chain does not match child order");
      } else if (sib != kNoSibling || high != kEmpty) {
        return bad("last node of a level has a sibling or a finite high key");
      }
      uint32_t last = 0;
      bool seen = false, ended = false;
      for (uint32_t l = 0; l < nd; ++l) {
        uint64_t p = w.at(idx)[l];
        uint32_t k = host_walk::key(p);
        if (k == kEmpty) {
          ended = true;
          continue;
        }
        if (ended) return bad("hole in node " + std::to_string(idx));
        if (seen && k <= last) return bad("unsorted node " + std::to_string(idx));
        if (k >= high) return bad("key >= high key in node " + std::to_string(idx));
        if (li > 0 && k < prev_high) return bad("key below left neighbour's high key");
        seen = true;
        last = k;
        if (leaf) {
          ++keys;
          if (pairs) pairs->push_back(mvbt_pair{k, host_walk::val(p)});
        } else {
          next.push_back(host_walk::val(p));
        }
      }
      if (!leaf && !seen) return bad("empty internal node " + std::to_string(idx));
      prev_high = high;
    }
    if (leaf_level) break;
    parent_above = above_level;
    level.swap(next);
  }
  if (num_keys) *num_keys = keys;
  if (height) *height = h;
  (void)nw;
  return MVBT_OK;
}

// ---- batched form helpers ----------------------------------------------------

// Conflict-table entries for a call of n operations: a power of two at least
// twice n (and at least 1024).
constexpr uint64_t conflict_entries(uint64_t n) {
  uint64_t e = 1024;
  while (e < 2 * n) e *= 2;
  return e;
}

// Entries of the sorted form's allocation before its histograms.
constexpr uint64_t kSortedArrays = 14;  // 8 working arrays + 6 gathered write fields
constexpr uint64_t kSegWords = (uint64_t)kOpClasses * kSortSegments;  // the split's segment counts
constexpr uint64_t sorted_scratch_words(uint64_t cap) {
  return kSortedArrays * cap + kSegWords + (uint64_t)kRadix * kSortSegments + kRadix;
}

// ---- allocation-free batched form (mvbt_run_batches_noalloc) ----------------
// Limits of one call, and the static scratch they need: the job descriptor,
// three per-batch counters, the conflict table, each op's batch, and the
// sorted form's arrays (about 37 MB per device).
constexpr uint64_t kNoAllocMaxOps = 1ull << 19;
constexpr uint32_t kNoAllocMaxBatches = 1u << 14;
constexpr uint64_t kNoAllocHash = conflict_entries(kNoAllocMaxOps);
constexpr uint64_t kNoAllocJobWords = (sizeof(batch_job) + 255) / 256 * 64;
constexpr uint64_t kNoAllocWords = kNoAllocJobWords + 3ull * kNoAllocMaxBatches + kNoAllocHash + kNoAllocMaxOps +
                                   sorted_scratch_words(kNoAllocMaxOps);
constexpr int kMaxDevices = 64;

}  // namespace

// Scratch of the allocation-free entry point. Each device the library runs on
// holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns on that device
// (g_scratch_ptr), never through this name.
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

namespace {

std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices];  // per device: the address of `scratch` there

// Looks up the device address of `scratch` on `device` (which must be
// current) once per device.
mvbt_status bind_noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(MVBT_ERR_UNSUPPORTED, "device ordinal above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return MVBT_OK;
  void* p = nullptr;
  MVBT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(MVBT_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch_ptr[device] = static_cast<uint32_t*>(p);
  return MVBT_OK;
}

uint32_t* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

void destroy_graph(mvbt_tree* t);

mvbt_status ensure_batch_scratch(mvbt_tree* t, uint32_t nb, uint64_t nops) {
  if (!t->job) {
    MVBT_HIP(hipMalloc(&t->job, sizeof(batch_job)));
    MVBT_HIP(hipMemset(t->job, 0, sizeof(batch_job)));
  }
  const uint64_t want_hash = conflict_entries(nops);
  const bool sorted = t->eff_mode == 2;
  if (nb <= t->b_capacity && want_hash <= t->b_hash_cap && nops <= t->b_op_cap &&
      (!sorted || nops <= t->b_sorted_cap))
    return MVBT_OK;
  MVBT_HIP(hipDeviceSynchronize());
  if (sorted && nops > t->b_sorted_cap) {
    uint64_t cap = 1024;
    while (cap < nops) cap *= 2;
    if (t->b_sorted) (void)hipFree(t->b_sorted);
    t->b_sorted = nullptr;
    t->b_sorted_cap = 0;
    MVBT_HIP(hipMalloc(&t->b_sorted, sorted_scratch_words(cap) * sizeof(uint32_t)));
    t->b_sorted_cap = cap;
    uint32_t* a = t->b_sorted;
    // w_rel, w_par, sort_k[0..1], sort_v[0..1], rd_find, rd_range, seg, hist, gat[0..5]
    uint32_t* sp[16];
    for (uint64_t k = 0; k < 8; ++k) sp[k] = a + k * cap;
    sp[8] = a + kSortedArrays * cap;
    sp[9] = sp[8] + kSegWords;
    for (uint64_t k = 0; k < 6; ++k) sp[10 + k] = a + (8 + k) * cap;
    static_assert(offsetof(batch_job, gat) == offsetof(batch_job, w_rel) + 10 * sizeof(void*), "sorted scratch layout");
    MVBT_HIP(hipMemcpy(&t->job->w_rel, sp, sizeof(sp), hipMemcpyHostToDevice));
    destroy_graph(t);  // the sort kernels of the graph take these pointers
  }
  if (nb <= t->b_capacity && want_hash <= t->b_hash_cap && nops <= t->b_op_cap) return MVBT_OK;
  if (nb > t->b_capacity) {
    uint32_t cap = 1024;
    while (cap < nb) cap *= 2;
    if (t->b_nwrites) (void)hipFree(t->b_nwrites);
    if (t->b_done) (void)hipFree(t->b_done);
    if (t->b_advance) (void)hipFree(t->b_advance);
    t->b_nwrites = nullptr;
    t->b_done = nullptr;
    t->b_advance = nullptr;
    t->b_capacity = 0;
    MVBT_HIP(hipMalloc(&t->b_nwrites, cap * sizeof(uint32_t)));
    MVBT_HIP(hipMalloc(&t->b_done, cap * sizeof(uint32_t)));
    MVBT_HIP(hipMalloc(&t->b_advance, cap * sizeof(int32_t)));
    t->b_capacity = cap;
  }
  if (want_hash > t->b_hash_cap) {
    if (t->b_hash) (void)hipFree(t->b_hash);
    t->b_hash = nullptr;
    t->b_hash_cap = 0;
    MVBT_HIP(hipMalloc(&t->b_hash, want_hash * sizeof(uint32_t)));
    t->b_hash_cap = want_hash;
  }
  if (nops > t->b_op_cap) {
    uint64_t cap = 1024;
    while (cap < nops) cap *= 2;
    if (t->b_op_batch) (void)hipFree(t->b_op_batch);
    t->b_op_batch = nullptr;
    t->b_op_cap = 0;
    MVBT_HIP(hipMalloc(&t->b_op_batch, cap * sizeof(uint32_t)));
    t->b_op_cap = cap;
  }
  // The scratch pointers live in the device descriptor; per-call parameters do not.
  uint32_t* ptrs[5] = {t->b_nwrites, t->b_done, reinterpret_cast<uint32_t*>(t->b_advance), t->b_hash,
                       t->b_op_batch};
  static_assert(offsetof(batch_job, op_batch) == offsetof(batch_job, nwrites) + 4 * sizeof(void*), "scratch layout");
  MVBT_HIP(hipMemcpy(&t->job->nwrites, ptrs, sizeof(ptrs), hipMemcpyHostToDevice));
  return MVBT_OK;
}

uint32_t worker_blocks(const mvbt_tree* t) {
  uint64_t b = (uint64_t)t->num_cus * t->worker_waves * 64 / kBlock;
  return (uint32_t)std::max<uint64_t>(b, 1);
}

uint32_t worker_blocks(const mvbt_tree* t);

// The sorted form's launch sequence after setup/clear/count: enqueued on a
// stream, or added to a graph, through `add(fn, grid, block, args)`.
// Radix passes cover the bits of the largest sort key (the pool capacity).
uint32_t sort_passes(const mvbt_tree* t) {
  const uint32_t cap = (uint32_t)t->cfg.pool_nodes;
  const uint32_t bits = 32 - (uint32_t)__builtin_clz(cap | 1u);
  return (bits + 7) / 8;
}

// A sorted-form call may free version copies between its windows when no
// snapshot outside it is live (none taken through this API, no device view
// that could hold one) and the caller did not opt out (MVBT_NO_RECLAIM).
bool may_reclaim_in_call(const mvbt_tree* t, uint32_t flags) {
  return !(flags & MVBT_NO_RECLAIM) && t->live.empty() && !t->external_writers;
}

// Window k of the sorted form: begin, split, route, sort, copy planning,
// write, find, range, end (reclamation when allowed, spill merge).
template <class Add>
mvbt_status sorted_window(mvbt_tree* t, Add& add, batch_job* job, uint32_t* a, uint64_t cap, uint32_t k) {
  uint32_t* sk[2] = {a + 2 * cap, a + 3 * cap};
  {
    void* args[] = {&job, &k};
    if (hipError_t e = add(reinterpret_cast<void*>(&batch_window_begin_kernel), 1, 64, args)) return hip_fail(e);
  }
  uint32_t* sv[2] = {a + 4 * cap, a + 5 * cap};
  uint32_t* seg = a + kSortedArrays * cap;
  uint32_t* hist = seg + kSegWords;
  uint32_t* totals = hist + (uint64_t)kRadix * kSortSegments;
  uint32_t* n_writes = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(job) + offsetof(batch_job, n_writes));
  uint32_t segs = kSortSegments;
  const sorted_fns f = table_for(t->nw, true)->sorted(t->read_tile);
  tree_ptrs tp = t->ptrs();
  const uint32_t grid = (uint32_t)t->num_cus * 8;
  {
    void* args[] = {&job};
    if (hipError_t e = add(reinterpret_cast<void*>(&batch_split_count_kernel), kSortSegments, kSortBlock, args))
      return hip_fail(e);
  }
  {
    // one row per operation class: totals land in n_writes, n_finds, n_ranges
    static_assert(offsetof(batch_job, n_ranges) == offsetof(batch_job, n_writes) + 2 * sizeof(uint32_t), "counts");
    void* args[] = {&seg, &segs, &n_writes};
    if (hipError_t e = add(reinterpret_cast<void*>(&radix_scan_kernel), kOpClasses, kSortBlock, args))
      return hip_fail(e);
  }
  {
    void* args[] = {&job};
    if (hipError_t e = add(reinterpret_cast<void*>(&batch_split_kernel), kSortSegments, kSortBlock, args))
      return hip_fail(e);
  }
  {
    void* args[] = {&tp, &job};
    if (hipError_t e = add(f.route, grid, kBlock, args)) return hip_fail(e);
  }
  const uint32_t passes = sort_passes(t);
  uint32_t shifts[4] = {0, 8, 16, 24};
  for (uint32_t p = 0; p < passes; ++p) {
    uint32_t* kin = sk[p & 1];
    uint32_t* vin = sv[p & 1];
    uint32_t* kout = sk[(p + 1) & 1];
    uint32_t* vout = sv[(p + 1) & 1];
    void* hargs[] = {&kin, &n_writes, &shifts[p], &hist};
    if (hipError_t e = add(reinterpret_cast<void*>(&radix_hist_kernel), kSortSegments, kSortBlock, hargs))
      return hip_fail(e);
    void* cargs[] = {&hist, &segs, &totals};
    if (hipError_t e = add(reinterpret_cast<void*>(&radix_scan_kernel), kRadix, kSortBlock, cargs))
      return hip_fail(e);
    void* sargs[] = {&kin, &vin, &kout, &vout, &n_writes, &shifts[p], &hist, &totals};
    if (hipError_t e = add(reinterpret_cast<void*>(&radix_scatter_kernel), kSortSegments, kSortBlock, sargs))
      return hip_fail(e);
  }
  uint32_t which = passes & 1;  // the buffer the last pass wrote
  {
    // copy planning: number the (leaf, batch) groups, reserve a slot for each
    uint32_t* n_epochs = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(job) + offsetof(batch_job, n_epochs));
    void* args[] = {&job, &which};
    if (hipError_t e = add(reinterpret_cast<void*>(&batch_epoch_count_kernel), kSortSegments, kSortBlock, args))
      return hip_fail(e);
    void* sargs[] = {&seg, &segs, &n_epochs};
    if (hipError_t e = add(reinterpret_cast<void*>(&radix_scan_kernel), 1, kSortBlock, sargs)) return hip_fail(e);
    if (hipError_t e = add(reinterpret_cast<void*>(&batch_epoch_assign_kernel), kSortSegments, kSortBlock, args))
      return hip_fail(e);
    void* rargs[] = {&job, &tp};
    if (hipError_t e = add(reinterpret_cast<void*>(&batch_copy_reserve_kernel), 1, 64, rargs)) return hip_fail(e);
    void* cargs[] = {&job, &tp, &which};
    if (hipError_t e = add(reinterpret_cast<void*>(&batch_copy_slots_kernel), grid, kBlock, cargs))
      return hip_fail(e);
    if (hipError_t e = add(reinterpret_cast<void*>(&batch_gather_kernel), grid, kBlock, args)) return hip_fail(e);
  }
  {
    // the write pass's occupancy is the tunable one (worker waves per CU)
    void* args[] = {&tp, &job, &which};
    if (hipError_t e = add(f.write, worker_blocks(t), kBlock, args)) return hip_fail(e);
  }
  {
    void* args[] = {&tp, &job};
    if (hipError_t e = add(f.find, (uint32_t)t->num_cus * 2, kFindBlock, args)) return hip_fail(e);
  }
  {
    void* args[] = {&tp, &job};
    if (hipError_t e = add(f.range, grid, kBlock, args)) return hip_fail(e);
  }
  {
    tree_state* st = t->st;
    epoch_stat* series = t->series;
    void* eargs[] = {&job, &st, &series};
    if (hipError_t e = add(reinterpret_cast<void*>(&batch_window_end_kernel), 1, 64, eargs)) return hip_fail(e);
    uint32_t* ring = t->retire_ring;
    uint32_t* spill = t->spill;
    uint32_t* fs = t->free_stack;
    void* cargs[] = {&st, &ring, &spill, &fs};
    if (hipError_t e = add(reinterpret_cast<void*>(&reclaim_copy_kernel), (uint32_t)t->num_cus * 4, kBlock, cargs))
      return hip_fail(e);
  }
  return MVBT_OK;
}

template <class Add>
mvbt_status sorted_sequence(mvbt_tree* t, Add&& add, batch_job* job, uint32_t* a, uint64_t cap, uint32_t windows) {
  for (uint32_t k = 0; k < windows; ++k)
    if (mvbt_status rc = sorted_window(t, add, job, a, cap, k)) return rc;
  return MVBT_OK;
}

void* worker_kernel(const mvbt_tree* t) {
  const dispatch_table* tb = table_for(t->nw, true);
  return t->eff_mode == 1 ? tb->chunk_worker() : tb->worker();
}

void destroy_graph(mvbt_tree* t) {
  for (uint32_t g = 0; g <= mvbt_tree::kMaxWindows; ++g) {
    if (t->graph_exec[g]) (void)hipGraphExecDestroy(t->graph_exec[g]);
    if (t->graph[g]) (void)hipGraphDestroy(t->graph[g]);
    t->graph_exec[g] = nullptr;
    t->graph[g] = nullptr;
    t->graph_setup[g] = nullptr;
  }
}

// Windows of a sorted-form call: about window_ops operations each, at most
// one per batch and kMaxWindows.
uint32_t windows_for(const mvbt_tree* t, uint64_t num_ops, uint32_t num_batches) {
  if (t->eff_mode != 2 || t->window_ops == 0) return 1;
  const uint64_t k = (num_ops + t->window_ops - 1) / t->window_ops;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({k, num_batches, mvbt_tree::kMaxWindows}));
}
uint32_t graph_index(const mvbt_tree* t, uint32_t windows) { return t->eff_mode == 2 ? windows : 0; }

// Below this many operations a call in mode 3 runs on the one-operation-per-
// tile worker: the sorted form's 25-odd launches per window then cost more
// than the frontier waits they save (served steps of one or two requests).
constexpr uint64_t kAutoSortedOps = 16384;
uint32_t effective_mode(const mvbt_tree* t, uint64_t num_ops) {
  if (t->worker_mode != 3) return t->worker_mode;
  return num_ops < kAutoSortedOps ? 0u : 2u;
}

// Explicit graph: setup -> count -> mask -> clear -> conflict -> (frontier -> worker | the sorted
// form's windows) -> finish. Every node but setup reads its inputs from the
// device descriptor, so a call only updates the setup node's parameters.
mvbt_status build_graph(mvbt_tree* t, const batch_params& bp) {
  const uint32_t gi = graph_index(t, bp.windows);
  MVBT_HIP(hipGraphCreate(&t->graph[gi], 0));
  batch_job* job = t->job;
  tree_state* st = t->st;
  batch_params p = bp;
  void* setup_args[] = {&job, &st, &p};
  void* job_args[] = {&job};
  hipKernelNodeParams kp{};
  hipGraphNode_t prev = nullptr, node = nullptr;
  auto add = [&](void* fn, unsigned grid, unsigned block, void** args) -> hipError_t {
    kp.func = fn;
    kp.gridDim = dim3(grid);
    kp.blockDim = dim3(block);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    hipError_t e = hipGraphAddKernelNode(&node, t->graph[gi], prev ? &prev : nullptr, prev ? 1 : 0, &kp);
    prev = node;
    return e;
  };
  void* count_args[] = {&job, &st};
  MVBT_HIP(add(reinterpret_cast<void*>(&batch_setup_kernel), 1, 64, setup_args));
  t->graph_setup[gi] = node;
  MVBT_HIP(add(reinterpret_cast<void*>(&batch_count_kernel), t->num_cus * 4, kBlock, count_args));
  MVBT_HIP(add(reinterpret_cast<void*>(&batch_mask_kernel), 1, 64, job_args));
  MVBT_HIP(add(reinterpret_cast<void*>(&batch_clear_kernel), t->num_cus * 4, kBlock, job_args));
  MVBT_HIP(add(reinterpret_cast<void*>(&batch_conflict_kernel), t->num_cus * 8, kBlock, count_args));
  if (t->eff_mode == 2) {
    if (mvbt_status rc = sorted_sequence(t, add, t->job, t->b_sorted, t->b_sorted_cap, bp.windows)) return rc;
  } else {
    MVBT_HIP(add(reinterpret_cast<void*>(&batch_frontier_kernel), t->num_cus, kBlock, job_args));
    tree_ptrs tp = t->ptrs();
    void* worker_args[] = {&tp, &job};
    MVBT_HIP(add(worker_kernel(t), worker_blocks(t), kBlock, worker_args));
  }
  MVBT_HIP(add(reinterpret_cast<void*>(&batch_finish_kernel), 1, 64, job_args));
  MVBT_HIP(hipGraphInstantiate(&t->graph_exec[gi], t->graph[gi], nullptr, nullptr, 0));
  return MVBT_OK;
}

}  // namespace

extern "C" {

const char* mvbt_last_error(void) { return g_last_error.c_str(); }

mvbt_status mvbt_create(const mvbt_config* config, mvbt_tree** out) {
  if (!config || !out) return fail(MVBT_ERR_INVALID_ARG, "null config or out pointer");
  *out = nullptr;
  mvbt_config cfg = *config;
  if (cfg.kind != MVBT_PLAIN && cfg.kind != MVBT_VERSIONED)
    return fail(MVBT_ERR_INVALID_ARG, "unknown tree kind");
  if (cfg.node_width == 0) cfg.node_width = 16;
  if (!table_for(cfg.node_width, true)) return fail(MVBT_ERR_INVALID_ARG, "node_width must be 16, 32 or 64");
  if (cfg.pool_nodes == 0) cfg.pool_nodes = 1ull << 22;
  if (cfg.pool_nodes < 4 || cfg.pool_nodes >= kMaxNodes)
    return fail(MVBT_ERR_INVALID_ARG, "pool_nodes must be in [4, 2^29-1)");
  if (cfg.retire_capacity != 0 && !is_pow2(cfg.retire_capacity))
    return fail(MVBT_ERR_INVALID_ARG, "retire_capacity must be a power of two");
  // Every retired copy awaiting reclamation occupies a pool slot, so a ring
  // of at least pool_nodes entries never fills (a full ring would leak).
  while (cfg.retire_capacity < cfg.pool_nodes) cfg.retire_capacity = cfg.retire_capacity ? cfg.retire_capacity * 2 : 1;

  int dev = cfg.device;
  if (dev < 0) MVBT_HIP(hipGetDevice(&dev));
  device_guard g(dev);
  if (!g.ok) return fail(MVBT_ERR_HIP, "hipSetDevice failed");

  std::unique_ptr<mvbt_tree> t(new (std::nothrow) mvbt_tree());
  if (!t) return fail(MVBT_ERR_OUT_OF_MEMORY, "host allocation failed");
  t->cfg = cfg;
  t->device = dev;
  t->nw = cfg.node_width;
  t->versioned = cfg.kind == MVBT_VERSIONED;
  hipDeviceProp_t prop;
  MVBT_HIP(hipGetDeviceProperties(&prop, dev));
  t->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  // The allocation-free entry point's scratch address on this device (looked
  // up again by that entry point if this fails).
  if (t->versioned && bind_noalloc_scratch(dev) != MVBT_OK) g_last_error.clear();

  auto cleanup = [&]() {
    if (t->nodes) (void)hipFree(t->nodes);
    if (t->st) (void)hipFree(t->st);
    if (t->free_stack) (void)hipFree(t->free_stack);
    if (t->spill) (void)hipFree(t->spill);
    if (t->retire_ring) (void)hipFree(t->retire_ring);
    if (t->snap_pos) (void)hipFree(t->snap_pos);
    if (t->snap_rec) (void)hipHostFree(t->snap_rec);
    if (t->series) (void)hipFree(t->series);
  };
  const size_t node_bytes = (size_t)t->nw * sizeof(uint64_t);
  if (hipMalloc(&t->nodes, cfg.pool_nodes * node_bytes) != hipSuccess ||
      hipMalloc(&t->st, sizeof(tree_state)) != hipSuccess ||
      hipMalloc(&t->free_stack, cfg.pool_nodes * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&t->spill, cfg.pool_nodes * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&t->retire_ring, cfg.retire_capacity * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&t->snap_pos, mvbt_tree::kSnapSlots * sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc(&t->snap_rec, sizeof(snapshot_record), hipHostMallocDefault) != hipSuccess ||
      (t->versioned && hipMalloc(&t->series, kSeriesCap * sizeof(epoch_stat)) != hipSuccess)) {
    (void)hipGetLastError();
    cleanup();
    return fail(MVBT_ERR_OUT_OF_MEMORY, "device allocation failed");
  }
  tree_state s{};
  s.root = 0;
  s.clock = 0;
  s.bump = 2;  // root + first leaf
  s.capacity = cfg.pool_nodes;
  s.free_top = 0;
  s.retire_cap = cfg.retire_capacity;
  s.peak_live = 2;
  if (hipMemcpy(t->st, &s, sizeof(s), hipMemcpyHostToDevice) != hipSuccess) {
    cleanup();
    return fail(MVBT_ERR_HIP, "state upload failed");
  }
  mvbt_status rc = table_for(t->nw, t->versioned)->init(t.get());
  if (rc == MVBT_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(MVBT_ERR_HIP, "init kernel failed");
  // the ordering events every writer uses, made here once rather than on first use
  if (rc == MVBT_OK && (hipEventCreateWithFlags(&t->clock_ev, hipEventDisableTiming) != hipSuccess ||
                        hipEventCreateWithFlags(&t->maint_ev, hipEventDisableTiming) != hipSuccess ||
                        hipEventCreateWithFlags(&t->excl_ev, hipEventDisableTiming) != hipSuccess))
    rc = fail(MVBT_ERR_HIP, "event creation failed");
  if (rc != MVBT_OK) {
    if (t->clock_ev) (void)hipEventDestroy(t->clock_ev);
    if (t->maint_ev) (void)hipEventDestroy(t->maint_ev);
    if (t->excl_ev) (void)hipEventDestroy(t->excl_ev);
    cleanup();
    return rc;
  }
  *out = t.release();
  return MVBT_OK;
}

mvbt_status mvbt_destroy(mvbt_tree* t) {
  if (!t) return MVBT_OK;
  device_guard g(t->device);
  (void)hipDeviceSynchronize();
  (void)hipFree(t->nodes);
  (void)hipFree(t->st);
  (void)hipFree(t->free_stack);
  (void)hipFree(t->spill);
  (void)hipFree(t->retire_ring);
  if (t->series) (void)hipFree(t->series);
  (void)hipFree(t->snap_pos);
  (void)hipHostFree(t->snap_rec);
  for (auto& kv : t->write_ev) (void)hipEventDestroy(kv.second);
  if (t->maint_ev) (void)hipEventDestroy(t->maint_ev);
  if (t->clock_ev) (void)hipEventDestroy(t->clock_ev);
  destroy_graph(t);
  if (t->job) (void)hipFree(t->job);
  if (t->b_nwrites) (void)hipFree(t->b_nwrites);
  if (t->b_done) (void)hipFree(t->b_done);
  if (t->b_advance) (void)hipFree(t->b_advance);
  if (t->b_hash) (void)hipFree(t->b_hash);
  if (t->b_op_batch) (void)hipFree(t->b_op_batch);
  if (t->b_sorted) (void)hipFree(t->b_sorted);
  if (t->excl_ev) (void)hipEventDestroy(t->excl_ev);
  delete t;
  return MVBT_OK;
}

// First stage of a large insert into an empty tree (see mvbt_insert).
constexpr uint64_t kFirstStage = 1ull << 14;

mvbt_status mvbt_insert(mvbt_tree* t, const uint32_t* keys, const uint32_t* values, uint64_t n,
                        void* stream, uint32_t flags) {
  (void)flags;
  if (mvbt_status rc = check_tree(t)) return rc;
  if (n == 0) return MVBT_OK;
  if (!keys || !values) return fail(MVBT_ERR_INVALID_ARG, "null keys or values");
  device_guard g(t->device);
  std::lock_guard<std::mutex> lk(t->mu);
  hipStream_t s = (hipStream_t)stream;
  const bool was_fresh = t->fresh;
  if (mvbt_status rc = before_write(t, s)) return rc;
  const dispatch_table* tb = table_for(t->nw, t->versioned);
  if (was_fresh && n > kFirstStage) {
    // An empty tree takes a large batch in growing stages (kFirstStage keys,
    // then four times as many as are in by then): in one launch the early
    // keys all split the same few leaves and update the same parents, one
    // lock holder at a time. The stages run in order on the stream, within
    // the same epoch (no snapshot between them).
    uint64_t done = 0, stage = kFirstStage;
    while (done < n) {
      const uint64_t m = std::min(stage, n - done);
      if (mvbt_status rc = tb->insert(t, keys + done, values + done, m, s)) return rc;
      done += m;
      stage = 4 * done;
    }
  } else if (mvbt_status rc = tb->insert(t, keys, values, n, s)) {
    return rc;
  }
  return after_write(t, s);
}

mvbt_status mvbt_erase(mvbt_tree* t, const uint32_t* keys, uint64_t n, void* stream, uint32_t flags) {
  (void)flags;
  if (mvbt_status rc = check_tree(t)) return rc;
  if (n == 0) return MVBT_OK;
  if (!keys) return fail(MVBT_ERR_INVALID_ARG, "null keys");
  device_guard g(t->device);
  std::lock_guard<std::mutex> lk(t->mu);
  hipStream_t s = (hipStream_t)stream;
  if (mvbt_status rc = before_write(t, s)) return rc;
  if (mvbt_status rc = table_for(t->nw, t->versioned)->erase(t, keys, nullptr, n, s)) return rc;
  return after_write(t, s);
}

mvbt_status mvbt_find(mvbt_tree* t, const uint32_t* keys, uint32_t* values, uint64_t n,
                      uint32_t snapshot, void* stream, uint32_t flags) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (n == 0) return MVBT_OK;
  if (!keys || !values) return fail(MVBT_ERR_INVALID_ARG, "null keys or values");
  if (!t->versioned && snapshot != MVBT_LATEST)
    return fail(MVBT_ERR_UNSUPPORTED, "plain trees have no snapshots");
  device_guard g(t->device);
  return table_for(t->nw, t->versioned)
      ->find(t, keys, values, n, snapshot, (hipStream_t)stream, (flags & MVBT_CONCURRENT) != 0);
}

mvbt_status mvbt_range_query(mvbt_tree* t, const uint32_t* lower, const uint32_t* upper,
                             mvbt_pair* out, uint32_t* counts, uint32_t L, uint64_t n,
                             uint32_t snapshot, void* stream, uint32_t flags) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (n == 0) return MVBT_OK;
  if (!lower || !upper) return fail(MVBT_ERR_INVALID_ARG, "null bounds");
  if (out && L == 0) return fail(MVBT_ERR_INVALID_ARG, "L must be > 0 when out is given");
  if (!t->versioned && snapshot != MVBT_LATEST)
    return fail(MVBT_ERR_UNSUPPORTED, "plain trees have no snapshots");
  device_guard g(t->device);
  return table_for(t->nw, t->versioned)
      ->range(t, lower, upper, reinterpret_cast<uint64_t*>(out), counts, L, n, snapshot,
              (hipStream_t)stream, (flags & MVBT_CONCURRENT) != 0);
}

mvbt_status mvbt_take_snapshot(mvbt_tree* t, void* stream, uint32_t* snapshot_id) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!snapshot_id) return fail(MVBT_ERR_INVALID_ARG, "null snapshot_id");
  if (!t->versioned) return fail(MVBT_ERR_UNSUPPORTED, "plain trees have no snapshots");
  device_guard g(t->device);
  std::lock_guard<std::mutex> lk(t->mu);
  hipStream_t s = (hipStream_t)stream;
  t->fresh = false;
  if (mvbt_status rc = after_clock(t, s)) return rc;
  // Stream-ordered: the id is the host's copy of the clock; the kernel
  // records the epoch edge in a device slot (one not held by a live bound).
  const int32_t slot = (int32_t)(t->snap_next % mvbt_tree::kSnapSlots);
  bool slot_free = true;
  for (const auto& b : t->bounds) slot_free = slot_free && b.slot != slot;
  if (!t->external_writers && slot_free) {
    const uint32_t id = t->clock;
    if (id == MVBT_LATEST) return fail(MVBT_ERR_UNSUPPORTED, "snapshot ids exhausted");
    // a reclamation launched earlier (on any stream) may still have to read
    // the edge a released snapshot left in this slot
    if (t->maint_ev) MVBT_HIP(hipStreamWaitEvent(s, t->maint_ev, 0));
    hipLaunchKernelGGL(snapshot_kernel, dim3(1), dim3(64), 0, s, t->st, nullptr, t->snap_pos + slot, id);
    MVBT_HIP(hipGetLastError());
    if (mvbt_status rc = clock_advanced(t, s, 1)) return rc;
    t->snap_next += 1;
    t->live[id] += 1;
    t->bounds.push_back({id, 0, slot});
    *snapshot_id = id;
    return MVBT_OK;
  }
  // Synchronous: the id is read back from the device.
  hipLaunchKernelGGL(snapshot_kernel, dim3(1), dim3(64), 0, s, t->st, t->snap_rec, nullptr, kNone);
  MVBT_HIP(hipGetLastError());
  MVBT_HIP(hipStreamSynchronize(s));
  uint32_t id = t->snap_rec->id;
  if (id == MVBT_LATEST) return fail(MVBT_ERR_UNSUPPORTED, "snapshot ids exhausted");
  if (mvbt_status rc = clock_advanced(t, s, 0)) return rc;
  t->clock = id + 1;
  t->live[id] += 1;
  t->bounds.push_back({id, t->snap_rec->ring_pos, -1});
  *snapshot_id = id;
  return MVBT_OK;
}

mvbt_status mvbt_release_snapshot(mvbt_tree* t, uint32_t id) {
  if (mvbt_status rc = check_tree(t)) return rc;
  std::lock_guard<std::mutex> lk(t->mu);
  auto it = t->live.find(id);
  if (it == t->live.end()) return fail(MVBT_ERR_INVALID_ARG, "snapshot is not live");
  if (--it->second == 0) t->live.erase(it);
  return MVBT_OK;
}

mvbt_status mvbt_reclaim(mvbt_tree* t, void* stream) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!t->versioned) return MVBT_OK;
  device_guard g(t->device);
  std::lock_guard<std::mutex> lk(t->mu);
  // Copies retired before the edge of the oldest live snapshot are garbage;
  // with no live snapshot, every retired copy is.
  unsigned long long end = ~0ull;
  const unsigned long long* end_from = nullptr;  // the edge, when it is held on the device
  if (!t->live.empty()) {
    uint32_t oldest = t->live.begin()->first;
    end = 0;
    for (const auto& b : t->bounds)
      if (b.id == oldest) {
        if (b.slot >= 0) end_from = t->snap_pos + b.slot;
        else end = b.pos;
      }
  }
  // Forget edges no longer needed.
  std::vector<mvbt_tree::bound> keep;
  for (const auto& b : t->bounds)
    if (t->live.count(b.id)) keep.push_back(b);
  t->bounds.swap(keep);
  hipStream_t s = (hipStream_t)stream;
  // Writers on other streams are ordered after this by maint_ev; writers
  // already running on them must have finished (documented precondition).
  for (auto& kv : t->write_ev)
    if (kv.first != s) MVBT_HIP(hipStreamWaitEvent(s, kv.second, 0));
  // a device-held edge is written by a snapshot kernel, possibly on another stream
  if (mvbt_status rc = after_clock(t, s)) return rc;
  return launch_maintenance(t, s, end, end_from, true);
}

mvbt_status mvbt_memory_stats(mvbt_tree* t, mvbt_memstats* out) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!out) return fail(MVBT_ERR_INVALID_ARG, "null out");
  device_guard g(t->device);
  tree_state s;
  if (mvbt_status rc = read_state(t, &s)) return rc;
  uint64_t on_stack = (s.free_top > 0 ? (uint64_t)s.free_top : 0) + s.spill_count;
  uint64_t bump = std::min<uint64_t>(s.bump, s.capacity);
  out->node_bytes = (uint64_t)t->nw * 8;
  out->capacity_nodes = s.capacity;
  out->live_nodes = bump - on_stack;
  out->peak_live_nodes = std::max<uint64_t>(s.peak_live, out->live_nodes);
  out->free_nodes = s.capacity - bump + on_stack;
  out->retired_pending = s.retire_tail - s.retire_head;
  out->reclaimed_nodes = s.reclaimed;
  out->retired_total = s.retire_tail;
  return MVBT_OK;
}

mvbt_status mvbt_memory_series(mvbt_tree* t, mvbt_epoch_stat* out, uint32_t capacity, uint32_t* count) {
  static_assert(sizeof(mvbt_epoch_stat) == sizeof(epoch_stat), "epoch stat layout");
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!count || (capacity && !out)) return fail(MVBT_ERR_INVALID_ARG, "null out or count");
  *count = 0;
  if (!t->versioned) return MVBT_OK;
  device_guard g(t->device);
  tree_state s;
  if (mvbt_status rc = read_state(t, &s)) return rc;  // synchronises the device
  const uint64_t have = std::min<uint64_t>(s.series_count, kSeriesCap);
  const uint64_t n = std::min<uint64_t>(have, capacity);
  std::vector<epoch_stat> all(kSeriesCap);
  MVBT_HIP(hipMemcpy(all.data(), t->series, kSeriesCap * sizeof(epoch_stat), hipMemcpyDeviceToHost));
  for (uint64_t i = 0; i < n; ++i) {  // the latest n, oldest first
    const uint64_t e = s.series_count - n + i;
    std::memcpy(out + i, &all[e % kSeriesCap], sizeof(epoch_stat));
  }
  *count = (uint32_t)n;
  return MVBT_OK;
}

mvbt_status mvbt_error_flags(mvbt_tree* t, uint32_t* flags) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!flags) return fail(MVBT_ERR_INVALID_ARG, "null flags");
  device_guard g(t->device);
  tree_state s;
  if (mvbt_status rc = read_state(t, &s)) return rc;
  *flags = s.error_flags;
  return MVBT_OK;
}

mvbt_status mvbt_clear_error_flags(mvbt_tree* t) {
  if (mvbt_status rc = check_tree(t)) return rc;
  device_guard g(t->device);
  MVBT_HIP(hipDeviceSynchronize());
  uint32_t zero = 0;
  MVBT_HIP(hipMemcpy(&t->st->error_flags, &zero, sizeof(zero), hipMemcpyHostToDevice));
  return MVBT_OK;
}

mvbt_status mvbt_validate(mvbt_tree* t, uint64_t* num_keys, uint32_t* height) {
  if (mvbt_status rc = check_tree(t)) return rc;
  device_guard g(t->device);
  host_walk w;
  if (mvbt_status rc = copy_pool(t, w)) return rc;
  return walk_tree(w, nullptr, num_keys, height);
}

mvbt_status mvbt_export(mvbt_tree* t, mvbt_pair* out, uint64_t capacity, uint64_t* count) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!count) return fail(MVBT_ERR_INVALID_ARG, "null count");
  device_guard g(t->device);
  host_walk w;
  if (mvbt_status rc = copy_pool(t, w)) return rc;
  std::vector<mvbt_pair> pairs;
  if (mvbt_status rc = walk_tree(w, &pairs, nullptr, nullptr)) return rc;
  *count = pairs.size();
  if (out) std::memcpy(out, pairs.data(), std::min<uint64_t>(capacity, pairs.size()) * sizeof(mvbt_pair));
  return MVBT_OK;
}

mvbt_status mvbt_export_dot(mvbt_tree* t, char* buf, uint64_t capacity, uint64_t* length, uint32_t max_nodes) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!length) return fail(MVBT_ERR_INVALID_ARG, "null length");
  device_guard g(t->device);
  host_walk w;
  if (mvbt_status rc = copy_pool(t, w)) return rc;
  if (mvbt_status rc = walk_tree(w, nullptr, nullptr, nullptr)) return rc;  // structure must be sound
  const uint32_t nd = w.data_lanes(), nw = w.nw;
  std::string dot = "digraph mvbt {\n  node [shape=record, height=.05, fontsize=9];\n";
  auto pair_text = [](uint64_t p) {
    return std::to_string(host_walk::key(p)) + ":" + std::to_string(host_walk::val(p));
  };
  auto emit = [&](uint32_t idx, bool copy) {
    const uint64_t* n = w.at(idx);
    const uint64_t m = n[nw - 1];
    const uint32_t flags = (uint32_t)(m >> 32);
    const bool leaf = (flags & kLeafBit) != 0;
    std::string label = "<h>" + std::to_string(idx) + (copy ? " (copy)" : leaf ? " leaf" : "");
    if (w.versioned && leaf) label += " @" + std::to_string((uint32_t)n[nw - 2]);
    for (uint32_t l = 0; l < nd; ++l)
      if (host_walk::key(n[l]) != kEmpty) label += "|<p" + std::to_string(l) + ">" + pair_text(n[l]);
    label += "|hi " + ((uint32_t)m == kEmpty ? std::string("inf") : std::to_string((uint32_t)m));
    dot += "  n" + std::to_string(idx) + " [label=\"" + label + "\"" + (copy ? ", style=dashed" : "") + "];\n";
  };
  // breadth-first over the latest tree, then each leaf's version chain
  std::vector<uint32_t> level{0u};
  uint64_t emitted = 0;
  const uint64_t limit = max_nodes ? max_nodes : ~0ull;
  while (!level.empty() && emitted < limit) {
    std::vector<uint32_t> next;
    for (uint32_t idx : level) {
      if (emitted++ >= limit) break;
      emit(idx, false);
      const uint64_t* n = w.at(idx);
      const uint64_t m = n[nw - 1];
      const uint32_t flags = (uint32_t)(m >> 32);
      const uint32_t sib = flags & kPtrMask;
      if (sib != kNoSibling) dot += "  n" + std::to_string(idx) + " -> n" + std::to_string(sib) + " [style=dotted, constraint=false];\n";
      if (!(flags & kLeafBit)) {
        for (uint32_t l = 0; l < nd; ++l)
          if (host_walk::key(n[l]) != kEmpty) {
            const uint32_t c = host_walk::val(n[l]);
            dot += "  n" + std::to_string(idx) + ":p" + std::to_string(l) + " -> n" + std::to_string(c) + ":h;\n";
            next.push_back(c);
          }
      } else if (w.versioned) {
        uint32_t older = (uint32_t)(n[nw - 2] >> 32), from = idx;
        for (uint32_t hop = 0; older != kNone && older < w.nodes && hop < 64 && emitted < limit; ++hop, ++emitted) {
          emit(older, true);
          dot += "  n" + std::to_string(from) + " -> n" + std::to_string(older) + " [style=dashed, label=\"older\"];\n";
          from = older;
          older = (uint32_t)(w.at(older)[nw - 2] >> 32);
        }
      }
    }
    level.swap(next);
  }
  dot += "}\n";
  *length = dot.size();
  if (buf && capacity) {
    const uint64_t k = std::min<uint64_t>(capacity - 1, dot.size());
    std::memcpy(buf, dot.data(), k);
    buf[k] = '\0';
  }
  return MVBT_OK;
}

mvbt_status mvbt_run_batches(mvbt_tree* t, const mvbt_op* ops, uint64_t num_ops,
                             const uint64_t* batch_offsets, uint32_t num_batches, uint32_t* results,
                             mvbt_pair* pairs, uint64_t pair_slots, uint32_t L, uint32_t* snapshot_base,
                             void* stream, uint32_t flags) {
  static_assert(sizeof(mvbt_op) == sizeof(op_record), "op layout");
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!t->versioned) return fail(MVBT_ERR_UNSUPPORTED, "the batched form needs a versioned tree");
  if (num_ops == 0 || num_batches == 0) return MVBT_OK;
  if (!ops || !batch_offsets || !results) return fail(MVBT_ERR_INVALID_ARG, "null ops, offsets or results");
  if (num_batches > (1u << 30)) return fail(MVBT_ERR_INVALID_ARG, "too many batches");
  if (num_ops >= (1ull << 31)) return fail(MVBT_ERR_INVALID_ARG, "num_ops must be below 2^31");
  if (pairs && L == 0) return fail(MVBT_ERR_INVALID_ARG, "L must be > 0 when pairs is given");
  device_guard g(t->device);
  std::lock_guard<std::mutex> lk(t->mu);
  t->eff_mode = effective_mode(t, num_ops);
  if (mvbt_status rc = reserve_ids(t, num_batches)) return rc;
  if (mvbt_status rc = ensure_batch_scratch(t, num_batches, num_ops)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (mvbt_status rc = before_write(t, s)) return rc;
  if (mvbt_status rc = isolate(t, s)) return rc;
  batch_params p{};
  p.ops = reinterpret_cast<const op_record*>(ops);
  p.nops = num_ops;
  p.offsets = reinterpret_cast<const unsigned long long*>(batch_offsets);
  p.nb = num_batches;
  p.L = L;
  p.results = results;
  p.pairs = reinterpret_cast<uint64_t*>(pairs);
  p.pair_slots = pairs ? pair_slots : 0;
  p.base_out = snapshot_base;
  p.hash_mask = (uint32_t)(conflict_entries(num_ops) - 1);
  p.windows = windows_for(t, num_ops, num_batches);
  p.reclaim = may_reclaim_in_call(t, flags) ? 1u : 0u;
  if (flags & MVBT_USE_GRAPH) {
    const uint32_t gi = graph_index(t, p.windows);
    if (!t->graph_exec[gi]) {
      if (mvbt_status rc = build_graph(t, p)) return rc;
    }
    batch_job* job = t->job;
    tree_state* st = t->st;
    void* args[] = {&job, &st, &p};
    hipKernelNodeParams kp{};
    kp.func = reinterpret_cast<void*>(&batch_setup_kernel);
    kp.gridDim = dim3(1);
    kp.blockDim = dim3(64);
    kp.kernelParams = args;
    MVBT_HIP(hipGraphExecKernelNodeSetParams(t->graph_exec[gi], t->graph_setup[gi], &kp));
    MVBT_HIP(hipGraphLaunch(t->graph_exec[gi], s));
  } else {
    hipLaunchKernelGGL(batch_setup_kernel, dim3(1), dim3(64), 0, s, t->job, t->st, p);
    hipLaunchKernelGGL(batch_count_kernel, dim3(t->num_cus * 4), dim3(kBlock), 0, s, t->job, t->st);
    hipLaunchKernelGGL(batch_mask_kernel, dim3(1), dim3(64), 0, s, t->job);
    hipLaunchKernelGGL(batch_clear_kernel, dim3(t->num_cus * 4), dim3(kBlock), 0, s, t->job);
    hipLaunchKernelGGL(batch_conflict_kernel, dim3(t->num_cus * 8), dim3(kBlock), 0, s, t->job, t->st);
    if (t->eff_mode == 2) {
      auto launch = [s](void* fn, unsigned grid, unsigned block, void** args) -> hipError_t {
        return hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, s);
      };
      if (mvbt_status rc = sorted_sequence(t, launch, t->job, t->b_sorted, t->b_sorted_cap, p.windows)) return rc;
    } else {
      hipLaunchKernelGGL(batch_frontier_kernel, dim3(t->num_cus), dim3(kBlock), 0, s, t->job);
      tree_ptrs tp = t->ptrs();
      batch_job* job = t->job;
      void* wargs[] = {&tp, &job};
      MVBT_HIP(hipLaunchKernel(worker_kernel(t), dim3(worker_blocks(t)), dim3(kBlock),
                               wargs, 0, s));
    }
    hipLaunchKernelGGL(batch_finish_kernel, dim3(1), dim3(64), 0, s, t->job);
  }
  MVBT_HIP(hipGetLastError());
  if (mvbt_status rc = clock_advanced(t, s, num_batches)) return rc;
  if (mvbt_status rc = isolated_done(t, s)) return rc;
  return after_write(t, s);
}

mvbt_status mvbt_reserve_batches(mvbt_tree* t, uint64_t num_ops, uint32_t num_batches) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!t->versioned) return fail(MVBT_ERR_UNSUPPORTED, "the batched form needs a versioned tree");
  if (num_batches > (1u << 30)) return fail(MVBT_ERR_INVALID_ARG, "too many batches");
  if (num_ops >= (1ull << 31)) return fail(MVBT_ERR_INVALID_ARG, "num_ops must be below 2^31");
  device_guard g(t->device);
  std::lock_guard<std::mutex> lk(t->mu);
  // every mode calls of up to num_ops operations may run in (both, in mode 3)
  std::vector<uint32_t> modes{effective_mode(t, num_ops)};
  if (t->worker_mode == 3 && num_ops >= kAutoSortedOps) modes.push_back(0);
  for (uint32_t m : modes) {
    t->eff_mode = m;
    if (mvbt_status rc = ensure_batch_scratch(t, std::max<uint32_t>(num_batches, 1), std::max<uint64_t>(num_ops, 1)))
      return rc;
    const uint32_t kmax = windows_for(t, num_ops, std::max<uint32_t>(num_batches, 1));
    for (uint32_t k = 1; k <= kmax; ++k) {
      batch_params p{};
      p.hash_mask = 1023;
      p.windows = k;
      if (!t->graph_exec[graph_index(t, k)])
        if (mvbt_status rc = build_graph(t, p)) return rc;
    }
  }
  MVBT_HIP(hipDeviceSynchronize());
  return MVBT_OK;
}

mvbt_status mvbt_run_batches_noalloc(mvbt_tree* t, const mvbt_op* ops, uint64_t num_ops,
                                     const uint64_t* batch_offsets, uint32_t num_batches, uint32_t* results,
                                     mvbt_pair* pairs, uint64_t pair_slots, uint32_t L, uint32_t* snapshot_base,
                                     void* stream) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!t->versioned) return fail(MVBT_ERR_UNSUPPORTED, "the batched form needs a versioned tree");
  if (num_ops == 0 || num_batches == 0) return MVBT_OK;
  if (!ops || !batch_offsets || !results) return fail(MVBT_ERR_INVALID_ARG, "null ops, offsets or results");
  if (num_ops > kNoAllocMaxOps || num_batches > kNoAllocMaxBatches)
    return fail(MVBT_ERR_INVALID_ARG, "mvbt_run_batches_noalloc takes at most " + std::to_string(kNoAllocMaxOps) +
                                          " operations and " + std::to_string(kNoAllocMaxBatches) +
                                          " batches per call (got " + std::to_string(num_ops) + " and " +
                                          std::to_string(num_batches) + ")");
  if (pairs && L == 0) return fail(MVBT_ERR_INVALID_ARG, "L must be > 0 when pairs is given");
  device_guard g(t->device);
  std::lock_guard<std::mutex> lk(t->mu);
  if (mvbt_status rc = bind_noalloc_scratch(t->device)) return rc;
  uint32_t* base = noalloc_scratch(t->device);
  if (mvbt_status rc = reserve_ids(t, num_batches)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (mvbt_status rc = before_write(t, s)) return rc;
  if (mvbt_status rc = isolate(t, s)) return rc;
  // layout: job | nwrites, done, advance (per batch) | conflict table | op batch | sorted arrays
  batch_job* job = reinterpret_cast<batch_job*>(base);
  scratch_layout l{};
  uint32_t* w = base + kNoAllocJobWords;
  l.nwrites = w;
  l.done = w + kNoAllocMaxBatches;
  l.advance = reinterpret_cast<int32_t*>(w + 2 * kNoAllocMaxBatches);
  l.hash = w + 3 * kNoAllocMaxBatches;
  l.op_batch = l.hash + kNoAllocHash;
  uint32_t* sorted = l.op_batch + kNoAllocMaxOps;
  for (uint64_t k = 0; k < 8; ++k) l.sorted[k] = sorted + k * kNoAllocMaxOps;
  l.sorted[8] = sorted + kSortedArrays * kNoAllocMaxOps;
  l.sorted[9] = l.sorted[8] + kSegWords;
  for (uint64_t k = 0; k < 6; ++k) l.sorted[10 + k] = sorted + (8 + k) * kNoAllocMaxOps;
  batch_params p{};
  p.ops = reinterpret_cast<const op_record*>(ops);
  p.nops = num_ops;
  p.offsets = reinterpret_cast<const unsigned long long*>(batch_offsets);
  p.nb = num_batches;
  p.L = L;
  p.results = results;
  p.pairs = reinterpret_cast<uint64_t*>(pairs);
  p.pair_slots = pairs ? pair_slots : 0;
  p.base_out = snapshot_base;
  p.hash_mask = (uint32_t)(conflict_entries(num_ops) - 1);
  p.windows = 1;
  if (t->window_ops) p.windows = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(
      {(num_ops + t->window_ops - 1) / t->window_ops, num_batches, mvbt_tree::kMaxWindows}));
  p.reclaim = may_reclaim_in_call(t, 0) ? 1u : 0u;
  hipLaunchKernelGGL(batch_bind_kernel, dim3(1), dim3(64), 0, s, job, l);
  hipLaunchKernelGGL(batch_setup_kernel, dim3(1), dim3(64), 0, s, job, t->st, p);
  hipLaunchKernelGGL(batch_count_kernel, dim3(t->num_cus * 4), dim3(kBlock), 0, s, job, t->st);
  hipLaunchKernelGGL(batch_mask_kernel, dim3(1), dim3(64), 0, s, job);
  hipLaunchKernelGGL(batch_clear_kernel, dim3(t->num_cus * 4), dim3(kBlock), 0, s, job);
  hipLaunchKernelGGL(batch_conflict_kernel, dim3(t->num_cus * 8), dim3(kBlock), 0, s, job, t->st);
  auto launch = [s](void* fn, unsigned grid, unsigned block, void** args) -> hipError_t {
    return hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, s);
  };
  if (mvbt_status rc = sorted_sequence(t, launch, job, sorted, kNoAllocMaxOps, p.windows)) return rc;
  hipLaunchKernelGGL(batch_finish_kernel, dim3(1), dim3(64), 0, s, job);
  MVBT_HIP(hipGetLastError());
  if (mvbt_status rc = clock_advanced(t, s, num_batches)) return rc;
  if (mvbt_status rc = isolated_done(t, s)) return rc;
  return after_write(t, s);
}

mvbt_status mvbt_memory_stats_async(mvbt_tree* t, mvbt_memstats* out, void* stream) {
  static_assert(sizeof(mvbt_memstats) == 8 * sizeof(unsigned long long), "memstats layout");
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!out) return fail(MVBT_ERR_INVALID_ARG, "null out");
  device_guard g(t->device);
  hipLaunchKernelGGL(stats_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, t->st,
                     (unsigned long long)t->nw * 8, reinterpret_cast<unsigned long long*>(out));
  MVBT_HIP(hipGetLastError());
  return MVBT_OK;
}

mvbt_status mvbt_set_worker_waves(mvbt_tree* t, uint32_t waves_per_cu) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (waves_per_cu < 1 || waves_per_cu > 64) return fail(MVBT_ERR_INVALID_ARG, "waves_per_cu in [1, 64]");
  std::lock_guard<std::mutex> lk(t->mu);
  if (t->worker_waves != waves_per_cu) {
    device_guard g(t->device);
    (void)hipDeviceSynchronize();
    destroy_graph(t);
    t->worker_waves = waves_per_cu;
  }
  return MVBT_OK;
}

mvbt_status mvbt_set_worker_mode(mvbt_tree* t, uint32_t mode) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (mode > 3)
    return fail(MVBT_ERR_INVALID_ARG,
                "worker mode is 0 (one operation per tile), 1 (chunks), 2 (sorted) or 3 (by call size)");
  std::lock_guard<std::mutex> lk(t->mu);
  if (t->worker_mode != mode) {
    device_guard g(t->device);
    (void)hipDeviceSynchronize();
    destroy_graph(t);
    t->worker_mode = mode;
  }
  return MVBT_OK;
}

mvbt_status mvbt_set_read_tile(mvbt_tree* t, uint32_t lanes) {
  if (mvbt_status rc = check_tree(t)) return rc;
  const uint32_t nw = t->nw;
  if (lanes != 0 && (lanes > nw || (lanes != 16 && lanes != 32 && lanes != 64) || nw / lanes > 4))
    return fail(MVBT_ERR_INVALID_ARG, "read tile: 0 (the node width) or 16, 32, 64 lanes, at most the node width");
  const uint32_t v = lanes == nw ? 0u : lanes;
  std::lock_guard<std::mutex> lk(t->mu);
  if (t->read_tile != v) {
    device_guard g(t->device);
    (void)hipDeviceSynchronize();
    destroy_graph(t);  // the sorted form's graphs hold the read kernels
    t->read_tile = v;
  }
  return MVBT_OK;
}

mvbt_status mvbt_set_batch_window(mvbt_tree* t, uint64_t ops_per_window) {
  if (mvbt_status rc = check_tree(t)) return rc;
  std::lock_guard<std::mutex> lk(t->mu);
  t->window_ops = ops_per_window;
  return MVBT_OK;
}

mvbt_status mvbt_device_view(mvbt_tree* t, mvbt_device_handle* out) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!out) return fail(MVBT_ERR_INVALID_ARG, "null out");
  std::lock_guard<std::mutex> lk(t->mu);
  tree_ptrs p = t->ptrs();
  out->nodes = p.nodes;
  out->state = p.st;
  out->free_stack = p.free_stack;
  out->spill = p.spill;
  out->retire_ring = p.retire_ring;
  out->retire_capacity = p.retire_cap;
  out->root = p.root_;
  out->capacity = p.cap_;
  out->node_width = t->nw;
  out->kind = t->versioned ? MVBT_VERSIONED : MVBT_PLAIN;
  t->external_writers = true;
  t->fresh = false;
  return MVBT_OK;
}

mvbt_status mvbt_concurrent_insert_range(mvbt_tree* t, const uint32_t* keys, const uint32_t* values,
                                         uint64_t num_insertions, const uint32_t* lower, const uint32_t* upper,
                                         uint64_t num_ranges, mvbt_pair* result, uint32_t* counts, uint32_t L,
                                         void* stream) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (num_insertions + num_ranges == 0) return MVBT_OK;
  if (num_insertions && (!keys || !values)) return fail(MVBT_ERR_INVALID_ARG, "null keys or values");
  if (num_ranges && (!lower || !upper)) return fail(MVBT_ERR_INVALID_ARG, "null range bounds");
  if (num_ranges && result && L == 0) return fail(MVBT_ERR_INVALID_ARG, "L must be > 0 when result is given");
  device_guard g(t->device);
  std::lock_guard<std::mutex> lk(t->mu);
  if (t->versioned)
    if (mvbt_status rc = reserve_ids(t, 1)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (mvbt_status rc = before_write(t, s)) return rc;
  if (mvbt_status rc = isolate(t, s)) return rc;
  if (mvbt_status rc = table_for(t->nw, t->versioned)
                           ->insert_range(t, keys, values, num_insertions, lower, upper, num_ranges, nullptr,
                                          reinterpret_cast<uint64_t*>(result), counts, L, s))
    return rc;
  if (t->versioned)  // the call reserved one snapshot id
    if (mvbt_status rc = clock_advanced(t, s, 1)) return rc;
  if (mvbt_status rc = isolated_done(t, s)) return rc;
  return after_write(t, s);
}

mvbt_status mvbt_concurrent_find_erase(mvbt_tree* t, const uint32_t* find_keys, uint32_t* find_results,
                                       uint64_t num_finds, const uint32_t* erase_keys, uint64_t num_erasures,
                                       void* stream) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (num_finds + num_erasures == 0) return MVBT_OK;
  if (num_finds && (!find_keys || !find_results)) return fail(MVBT_ERR_INVALID_ARG, "null find keys or results");
  if (num_erasures && !erase_keys) return fail(MVBT_ERR_INVALID_ARG, "null erase keys");
  device_guard g(t->device);
  std::lock_guard<std::mutex> lk(t->mu);
  if (t->versioned)
    if (mvbt_status rc = reserve_ids(t, 1)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (mvbt_status rc = before_write(t, s)) return rc;
  if (mvbt_status rc = isolate(t, s)) return rc;
  if (mvbt_status rc = table_for(t->nw, t->versioned)
                           ->find_erase(t, erase_keys, nullptr, num_erasures, find_keys, nullptr, num_finds,
                                        find_results, nullptr, nullptr, 0, s))
    return rc;
  if (t->versioned)  // the call reserved one snapshot id
    if (mvbt_status rc = clock_advanced(t, s, 1)) return rc;
  if (mvbt_status rc = isolated_done(t, s)) return rc;
  return after_write(t, s);
}

mvbt_status mvbt_bulk_build(mvbt_tree* t, const uint32_t* keys, const uint32_t* values, uint64_t n,
                            uint32_t fill, void* stream) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (n == 0) return MVBT_OK;
  if (!keys || !values) return fail(MVBT_ERR_INVALID_ARG, "null keys or values");
  const dispatch_table* tb = table_for(t->nw, t->versioned);
  if (fill == 0) fill = (2 * tb->data_lanes + 2) / 3;
  if (fill < 2 || fill > tb->data_lanes)
    return fail(MVBT_ERR_INVALID_ARG, "fill must be in [2, " + std::to_string(tb->data_lanes) + "]");
  device_guard g(t->device);
  std::lock_guard<std::mutex> lk(t->mu);
  if (!t->fresh)
    return fail(MVBT_ERR_INVALID_ARG, "bulk build needs a new tree (no write, snapshot or device view yet)");
  MVBT_HIP(hipDeviceSynchronize());  // nothing else may touch the hand-off word meanwhile
  bulk_plan plan;
  if (!make_bulk_plan(n, fill, tb->data_lanes, &plan))
    return fail(MVBT_ERR_INVALID_ARG, "too many levels for this fill");
  if (plan.total > t->cfg.pool_nodes)
    return fail(MVBT_ERR_INVALID_ARG, "node pool too small: the build needs " + std::to_string(plan.total) +
                                          " nodes, the pool has " + std::to_string(t->cfg.pool_nodes));
  hipStream_t s = (hipStream_t)stream;
  // Validate first, so a rejected input leaves the tree untouched and new.
  unsigned long long* bad = &t->st->scratch[0];
  MVBT_HIP(hipMemsetAsync(bad, 0, sizeof(*bad), s));
  const uint64_t blocks = std::min<uint64_t>((n + kBlock - 1) / kBlock, (uint64_t)t->num_cus * 8);
  hipLaunchKernelGGL(bulk_check_kernel, dim3((uint32_t)blocks), dim3(kBlock), 0, s, keys,
                     (unsigned long long)n, bad);
  MVBT_HIP(hipGetLastError());
  unsigned long long h_bad = 0;
  MVBT_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(h_bad), hipMemcpyDeviceToHost, s));
  MVBT_HIP(hipStreamSynchronize(s));
  if (h_bad)
    return fail(MVBT_ERR_INVALID_ARG,
                "bulk build keys must be strictly ascending and exclude the reserved keys 0 and 0xFFFFFFFF");
  if (mvbt_status rc = before_write(t, s)) return rc;
  if (mvbt_status rc = tb->bulk(t, keys, values, plan, s)) return rc;
  if (mvbt_status rc = after_write(t, s)) return rc;
  MVBT_HIP(hipStreamSynchronize(s));
  return MVBT_OK;
}

mvbt_status mvbt_node_width(mvbt_tree* t, uint32_t* nw) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!nw) return fail(MVBT_ERR_INVALID_ARG, "null out");
  *nw = t->nw;
  return MVBT_OK;
}

mvbt_status mvbt_kind_of(mvbt_tree* t, uint32_t* kind) {
  if (mvbt_status rc = check_tree(t)) return rc;
  if (!kind) return fail(MVBT_ERR_INVALID_ARG, "null out");
  *kind = t->versioned ? MVBT_VERSIONED : MVBT_PLAIN;
  return MVBT_OK;
}

}  // extern "C"
