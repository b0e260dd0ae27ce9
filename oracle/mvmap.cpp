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
