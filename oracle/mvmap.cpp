// ORACLE - test infrastructure only. Nothing in the library links or calls
// this file; tests, the smoke check and the benchmark's CPU-baseline leg use it
// as the checker.
//
// CPU restatement of the observable semantics the reference validates against
// (owensgroup/MVGpuBTree, paths under include/ and test/):
//   * find: value of the key, or 0xFFFFFFFF when absent
//     (validation.hpp:75-104, versioned_node.hpp:267-271);
//   * range [lo, hi): ascending pairs from std::set::lower_bound(lo) while
//     key < hi, packed at slot i*L, with a count per query
//     (validation.hpp:184-264, unittest_versioning.cu:315-334);
//   * insert of an existing key replaces its value (versioned_node.hpp:302-304);
//     erase of an absent key is a no-op (versioned_node.hpp:324-346);
//   * snapshots: take_snapshot returns the counter and increments it
//     (gpu_versioned_blink_tree.hpp:297-302, :333-344); a read at s observes
//     exactly the writes made while the counter was <= s (writes stamped with
//     the counter value, gpu_versioned_blink_tree.hpp:836,1075; walk
//     newest->oldest until stamp <= s, :1578-1603); s = 0xFFFFFFFF reads the
//     latest state (:347-353).
// Keys 0 and 0xFFFFFFFF are reserved (tombstone and empty sentinel,
// gpu_versioned_blink_tree.hpp:1460): writes of them are ignored, reads of
// them miss.
//
// State: std::map<key, history>, where history is the ordered list of
// (stamp, present, value) records of that key.
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kLatest = 0xFFFFFFFFu;

struct record {
  uint32_t stamp;
  uint32_t value;
  bool present;
};

struct mvmap {
  std::map<uint32_t, std::vector<record>> keys;
  uint32_t clock = 0;

  static bool valid(uint32_t k) { return k != 0 && k != kEmpty; }

  void put(uint32_t k, uint32_t v, bool present) {
    auto& h = keys[k];
    if (!h.empty() && h.back().stamp == clock) {
      h.back().value = v;
      h.back().present = present;
    } else {
      h.push_back(record{clock, v, present});
    }
  }
  // Record visible at ts, or nullptr.
  static const record* at(const std::vector<record>& h, uint32_t ts) {
    if (ts == kLatest) return h.empty() ? nullptr : &h.back();
    auto it = std::upper_bound(h.begin(), h.end(), ts,
                               [](uint32_t t, const record& r) { return t < r.stamp; });
    if (it == h.begin()) return nullptr;
    return &*(it - 1);
  }
  uint32_t find(uint32_t k, uint32_t ts) const {
    if (!valid(k)) return kEmpty;
    auto it = keys.find(k);
    if (it == keys.end()) return kEmpty;
    const record* r = at(it->second, ts);
    return (r && r->present) ? r->value : kEmpty;
  }
  uint32_t range(uint32_t lo, uint32_t hi, uint32_t ts, uint64_t* out, uint32_t cap) const {
    uint32_t count = 0;
    if (lo == 0) lo = 1;
    for (auto it = keys.lower_bound(lo); it != keys.end() && it->first < hi; ++it) {
      const record* r = at(it->second, ts);
      if (!r || !r->present) continue;
      if (out && count < cap) out[count] = (uint64_t)it->first | ((uint64_t)r->value << 32);
      ++count;
    }
    return count;
  }
};

}  // namespace

extern "C" {

void* oracle_create() { return new mvmap(); }
void oracle_destroy(void* h) { delete static_cast<mvmap*>(h); }

// Bulk build from keys already in ascending order (hinted insertion, O(1)
// amortised per key); stamps them with the current clock.
void oracle_build_sorted(void* h, const uint32_t* keys, const uint32_t* values, uint64_t n) {
  mvmap* m = static_cast<mvmap*>(h);
  auto hint = m->keys.end();
  for (uint64_t i = 0; i < n; ++i) {
    if (!mvmap::valid(keys[i])) continue;
    hint = m->keys.emplace_hint(hint, keys[i], std::vector<record>{record{m->clock, values[i], true}});
    ++hint;
  }
}

void oracle_insert(void* h, const uint32_t* keys, const uint32_t* values, uint64_t n) {
  mvmap* m = static_cast<mvmap*>(h);
  for (uint64_t i = 0; i < n; ++i)
    if (mvmap::valid(keys[i])) m->put(keys[i], values[i], true);
}

// Returns how many keys were present (and are now erased).
uint64_t oracle_erase(void* h, const uint32_t* keys, uint64_t n) {
  mvmap* m = static_cast<mvmap*>(h);
  uint64_t erased = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (m->find(keys[i], kLatest) == kEmpty) continue;
    m->put(keys[i], 0, false);
    ++erased;
  }
  return erased;
}

uint32_t oracle_take_snapshot(void* h) { return static_cast<mvmap*>(h)->clock++; }

void oracle_find(void* h, const uint32_t* keys, uint32_t* out, uint64_t n, uint32_t ts) {
  const mvmap* m = static_cast<const mvmap*>(h);
  for (uint64_t i = 0; i < n; ++i) out[i] = m->find(keys[i], ts);
}

// out may be null (count only); pairs of query i start at out + i*L.
void oracle_range(void* h, const uint32_t* lo, const uint32_t* hi, uint64_t* out, uint32_t* counts,
                  uint32_t L, uint64_t n, uint32_t ts) {
  const mvmap* m = static_cast<const mvmap*>(h);
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t c = m->range(lo[i], hi[i], ts, out ? out + i * (uint64_t)L : nullptr, L);
    if (counts) counts[i] = c;
  }
}

// Number of keys present at ts.
uint64_t oracle_size(void* h, uint32_t ts) {
  const mvmap* m = static_cast<const mvmap*>(h);
  uint64_t c = 0;
  for (const auto& kv : m->keys) {
    const record* r = mvmap::at(kv.second, ts);
    if (r && r->present) ++c;
  }
  return c;
}

// Pairs present at ts in key order; returns the count (writes at most cap).
uint64_t oracle_export(void* h, uint64_t* out, uint64_t cap, uint32_t ts) {
  const mvmap* m = static_cast<const mvmap*>(h);
  uint64_t c = 0;
  for (const auto& kv : m->keys) {
    const record* r = mvmap::at(kv.second, ts);
    if (!r || !r->present) continue;
    if (out && c < cap) out[c] = (uint64_t)kv.first | ((uint64_t)r->value << 32);
    ++c;
  }
  return c;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batch replay: the observable semantics of a stream of batches of mixed
// operations when every batch is issued through the reference's per-batch
// host API with a snapshot taken before it (take_snapshot,
// gpu_versioned_blink_tree.hpp:297-302; find/range_query at that snapshot,
// :154-164 and :196-216; then insert :106-142 and erase :166-173). Reads of
// batch b therefore see exactly the writes of batches < b. The batched form
// of the library claims this result for one persistent launch.
//
// ops: n rows {type, key, arg, slot}; type 0 find(key), 1 insert(key, arg),
// 2 erase(key), 3 range [key, arg) with pairs at pairs + slot * L.
// offsets: nb + 1 absolute op indices; batch b = ops[offsets[b], offsets[b+1]).
// results[i] (absolute index): find -> value or 0xFFFFFFFF; range -> true
// count (at most L pairs written); insert -> 1 new key, 0 value replaced;
// erase -> 1 erased, 0 absent; a write of a reserved key -> 0xFFFFFFFF (no
// effect). A batch's writes must name distinct keys: otherwise the order of
// two writes of one key inside a batch would decide the result.
namespace {

enum : uint32_t { kFind = 0, kInsert = 1, kErase = 2, kRange = 3 };

// 0 when the call is well formed; otherwise 1 + the offending op index
// (a repeated write key inside one batch, or an unknown op type) as a
// negative number.
int64_t check_batches(const uint32_t* ops, uint64_t n, const uint64_t* offsets, uint32_t nb) {
  for (uint32_t b = 0; b < nb; ++b) {
    if (offsets[b] > offsets[b + 1] || offsets[b + 1] > n) return -1;
    std::unordered_set<uint32_t> written;
    for (uint64_t i = offsets[b]; i < offsets[b + 1]; ++i) {
      const uint32_t* o = ops + 4 * i;
      if (o[0] > kRange) return -(int64_t)(i + 1);
      if ((o[0] == kInsert || o[0] == kErase) && mvmap::valid(o[1]) && !written.insert(o[1]).second)
        return -(int64_t)(i + 1);
    }
  }
  return 0;
}

}  // namespace

extern "C" {

// Exact replay on the multiversion map (its state advances: one snapshot per
// batch). Returns 0, or the negative code of check_batches (nothing applied).
int64_t oracle_replay_batches(void* h, const uint32_t* ops, uint64_t n, const uint64_t* offsets, uint32_t nb,
                              uint32_t* results, uint64_t* pairs, uint64_t pair_slots, uint32_t L) {
  mvmap* m = static_cast<mvmap*>(h);
  if (int64_t e = check_batches(ops, n, offsets, nb)) return e;
  for (uint32_t b = 0; b < nb; ++b) {
    const uint32_t s = m->clock++;  // the batch's snapshot
    for (uint64_t i = offsets[b]; i < offsets[b + 1]; ++i) {
      const uint32_t* o = ops + 4 * i;
      if (o[0] == kFind) {
        results[i] = m->find(o[1], s);
      } else if (o[0] == kRange) {
        uint64_t* out = (pairs && o[3] < pair_slots) ? pairs + (uint64_t)o[3] * L : nullptr;
        results[i] = m->range(o[1], o[2], s, out, L);
      }
    }
    for (uint64_t i = offsets[b]; i < offsets[b + 1]; ++i) {
      const uint32_t* o = ops + 4 * i;
      if (o[0] != kInsert && o[0] != kErase) continue;
      if (!mvmap::valid(o[1])) {
        results[i] = kEmpty;
        continue;
      }
      auto it = m->keys.find(o[1]);
      const record* r = it == m->keys.end() ? nullptr : mvmap::at(it->second, kLatest);
      const bool present = r && r->present;
      if (o[0] == kInsert) {
        results[i] = present ? 0u : 1u;
        m->put(o[1], o[2], true);
      } else {
        results[i] = present ? 1u : 0u;
        if (present) m->put(o[1], 0, false);
      }
    }
  }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Write-once replay for full-size streams (tens of millions to 500M initial
// keys): the same results as oracle_replay_batches when no key is written
// more than once in the whole stream, which holds for the benchmark streams
// (inserts of fresh keys, erases of distinct initial keys). Each written key
// then has one event: its state before the event (from the sorted initial
// pairs) and after it. A read of batch b sees a key's post-event state iff
// the event's batch is < b. Answers do not depend on each other, so any
// subset of ops is answered independently, on several threads.
namespace {

struct wo_event {
  uint32_t key, batch;
  uint32_t init_val, post_val;
  bool init_present, post_present;
};

struct wo_ctx {
  const uint32_t* k0;
  const uint32_t* v0;
  uint64_t n0;
  const uint32_t* ops;
  uint64_t n;
  std::vector<uint64_t> offsets;
  std::vector<wo_event> ev;  // sorted by key

  uint32_t batch_of(uint64_t i) const {
    return (uint32_t)(std::upper_bound(offsets.begin(), offsets.end(), i) - offsets.begin()) - 1;
  }
  const wo_event* event(uint32_t k) const {
    auto it = std::lower_bound(ev.begin(), ev.end(), k, [](const wo_event& e, uint32_t x) { return e.key < x; });
    return (it != ev.end() && it->key == k) ? &*it : nullptr;
  }
  bool initial(uint32_t k, uint32_t* v) const {
    const uint32_t* p = std::lower_bound(k0, k0 + n0, k);
    if (p == k0 + n0 || *p != k) return false;
    *v = v0[p - k0];
    return true;
  }
  // State of event key at the reads of batch b.
  static bool at(const wo_event& e, uint32_t b, uint32_t* v) {
    const bool post = b > e.batch;
    *v = post ? e.post_val : e.init_val;
    return post ? e.post_present : e.init_present;
  }
  uint32_t answer(uint64_t i, uint64_t* out, uint32_t L) const {
    const uint32_t* o = ops + 4 * i;
    const uint32_t b = batch_of(i);
    if (o[0] == kFind) {
      if (!mvmap::valid(o[1])) return kEmpty;
      uint32_t v;
      if (const wo_event* e = event(o[1])) return at(*e, b, &v) ? v : kEmpty;
      return initial(o[1], &v) ? v : kEmpty;
    }
    if (o[0] == kRange) {
      const uint32_t lo = o[1] == 0 ? 1u : o[1], hi = o[2];
      uint32_t count = 0;
      if (hi <= lo) return 0;
      const uint32_t* p = std::lower_bound(k0, k0 + n0, lo);
      auto q = std::lower_bound(ev.begin(), ev.end(), lo, [](const wo_event& e, uint32_t x) { return e.key < x; });
      for (;;) {
        const uint32_t kp = p < k0 + n0 ? *p : kEmpty;
        const uint32_t kq = q != ev.end() ? q->key : kEmpty;
        const uint32_t k = std::min(kp, kq);
        if (k >= hi || k == kEmpty) break;
        uint32_t v = 0;
        bool present;
        if (kq == k) {
          present = at(*q, b, &v);
          ++q;
          if (kp == k) ++p;
        } else {
          present = true;
          v = v0[p - k0];
          ++p;
        }
        if (!present) continue;
        if (out && count < L) out[count] = (uint64_t)k | ((uint64_t)v << 32);
        ++count;
      }
      return count;
    }
    if (!mvmap::valid(o[1])) return kEmpty;
    const wo_event* e = event(o[1]);  // every valid write has its event
    const bool present = e->init_present;
    return o[0] == kInsert ? (present ? 0u : 1u) : (present ? 1u : 0u);
  }
};

}  // namespace

extern "C" {

// Builds the context; the arrays must outlive it. *status: 0 OK; 1 initial
// keys not strictly ascending or reserved; 2 a key written twice in the
// stream (op index in *bad_op); 3 an unknown op type or bad offsets.
void* oracle_wo_create(const uint32_t* k0, const uint32_t* v0, uint64_t n0, const uint32_t* ops, uint64_t n,
                       const uint64_t* offsets, uint32_t nb, int32_t* status, uint64_t* bad_op) {
  *status = 0;
  *bad_op = 0;
  for (uint64_t i = 0; i < n0; ++i)
    if (!mvmap::valid(k0[i]) || (i && k0[i] <= k0[i - 1])) {
      *status = 1;
      *bad_op = i;
      return nullptr;
    }
  wo_ctx* c = new wo_ctx();
  c->k0 = k0;
  c->v0 = v0;
  c->n0 = n0;
  c->ops = ops;
  c->n = n;
  c->offsets.assign(offsets, offsets + nb + 1);
  struct w {
    uint32_t key;
    uint64_t op;
    uint32_t batch;
  };
  std::vector<w> writes;
  for (uint32_t b = 0; b < nb; ++b) {
    if (offsets[b] > offsets[b + 1] || offsets[b + 1] > n) {
      *status = 3;
      delete c;
      return nullptr;
    }
    for (uint64_t i = offsets[b]; i < offsets[b + 1]; ++i) {
      const uint32_t* o = ops + 4 * i;
      if (o[0] > kRange) {
        *status = 3;
        *bad_op = i;
        delete c;
        return nullptr;
      }
      if ((o[0] == kInsert || o[0] == kErase) && mvmap::valid(o[1])) writes.push_back(w{o[1], i, b});
    }
  }
  std::sort(writes.begin(), writes.end(), [](const w& a, const w& b) { return a.key < b.key || (a.key == b.key && a.op < b.op); });
  c->ev.reserve(writes.size());
  for (size_t j = 0; j < writes.size(); ++j) {
    if (j && writes[j].key == writes[j - 1].key) {
      *status = 2;
      *bad_op = writes[j].op;
      delete c;
      return nullptr;
    }
    const uint32_t* o = ops + 4 * writes[j].op;
    wo_event e{};
    e.key = o[1];
    e.batch = writes[j].batch;
    e.init_present = c->initial(o[1], &e.init_val);
    if (!e.init_present) e.init_val = kEmpty;
    if (o[0] == kInsert) {
      e.post_present = true;
      e.post_val = o[2];
    } else {
      e.post_present = false;
      e.post_val = kEmpty;
    }
    c->ev.push_back(e);
  }
  return c;
}

void oracle_wo_destroy(void* c) { delete static_cast<wo_ctx*>(c); }

// Answers ops op_idx[0..m): results[j], and for a range its pairs at
// pairs + j * L (pairs may be null). Runs on `threads` host threads.
void oracle_wo_answer(void* cp, const uint64_t* op_idx, uint64_t m, uint32_t* results, uint64_t* pairs, uint32_t L,
                      uint32_t threads) {
  const wo_ctx* c = static_cast<const wo_ctx*>(cp);
  if (threads < 1) threads = 1;
  auto work = [&](uint64_t a, uint64_t z) {
    for (uint64_t j = a; j < z; ++j) results[j] = c->answer(op_idx[j], pairs ? pairs + j * L : nullptr, L);
  };
  if (threads == 1 || m < 4096) {
    work(0, m);
    return;
  }
  std::vector<std::thread> th;
  for (uint32_t t = 0; t < threads; ++t) th.emplace_back(work, m * t / threads, m * (t + 1) / threads);
  for (auto& x : th) x.join();
}

}  // extern "C"
