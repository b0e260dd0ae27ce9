// C++ host program over include/gpu_btree.h: the reference unit tests'
// closed-form cases (test/unittest_btree.cu:138-251,
// test/unittest_versioning.cu:145-376) run through the C++ API, with I/O in
// pinned host memory passed straight to the kernels as those tests do.
// keys[i] = 2(i+1), value = key, absent keys = key + 1, ranges [k, k + 2L)
// hold exactly the L keys k, k+2, ..., k+2(L-1) (fewer at the tail).
// Prints "OK" and exits 0 when every check passes.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpu_btree.h"

namespace {

int g_errors = 0;
#define EXPECT(cond, ...)                  \
  do {                                     \
    if (!(cond)) {                         \
      if (g_errors < 10) {                 \
        std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
        std::printf(__VA_ARGS__);          \
        std::printf("\n");                 \
      }                                    \
      ++g_errors;                          \
    }                                      \
  } while (0)

template <typename T>
T* pinned(size_t n) {
  void* p = nullptr;
  if (hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault) != hipSuccess) std::abort();
  return static_cast<T*>(p);
}

void shuffle(uint32_t* a, uint32_t n, uint64_t seed) {
  for (uint32_t i = n; i > 1; --i) {
    seed = seed * 6364136223846793005ull + 1442695040888963407ull;
    uint32_t j = (uint32_t)((seed >> 33) % i);
    uint32_t t = a[i - 1];
    a[i - 1] = a[j];
    a[j] = t;
  }
}

template <typename Tree>
void plain_cases(uint32_t n) {
  Tree tree(1u << 16);
  uint32_t* keys = pinned<uint32_t>(n);
  uint32_t* vals = pinned<uint32_t>(n);
  uint32_t* absent = pinned<uint32_t>(n);
  uint32_t* out = pinned<uint32_t>(n);
  for (uint32_t i = 0; i < n; ++i) keys[i] = 2 * (i + 1);
  shuffle(keys, n, 7);
  for (uint32_t i = 0; i < n; ++i) {
    vals[i] = keys[i];
    absent[i] = keys[i] + 1;
  }
  tree.insert(keys, vals, n);
  (void)hipDeviceSynchronize();
  uint64_t count = 0;
  EXPECT(tree.validate_tree_structure(&count), "structure after insert");
  EXPECT(count == n, "key count %llu != %u", (unsigned long long)count, n);
  tree.find(keys, out, n);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == keys[i], "FindExist key %u got %u", keys[i], out[i]);
  tree.find(absent, out, n);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == MVBT_NOT_FOUND, "FindNotExist key %u got %u", absent[i], out[i]);
  tree.erase(absent, n);  // EraseNone
  tree.find(keys, out, n);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == keys[i], "EraseNone key %u got %u", keys[i], out[i]);
  tree.erase(keys, n);  // EraseAll
  tree.find(keys, out, n);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == MVBT_NOT_FOUND, "EraseAll key %u got %u", keys[i], out[i]);
  tree.insert(keys, vals, n);  // EraseAllInsertAll
  tree.find(keys, out, n);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == keys[i], "Reinsert key %u got %u", keys[i], out[i]);
  EXPECT(tree.validate_tree_structure(), "structure at end");
  EXPECT(tree.error_flags() == 0, "device error flags 0x%x", tree.error_flags());
  (void)hipHostFree(keys);
  (void)hipHostFree(vals);
  (void)hipHostFree(absent);
  (void)hipHostFree(out);
}

void versioned_cases(uint32_t n, uint32_t L) {
  GpuBTree::gpu_versioned_btree<uint32_t, uint32_t, 16> tree(1u << 16);
  using pair_t = GpuBTree::gpu_versioned_btree<uint32_t, uint32_t, 16>::pair_type;
  uint32_t* keys = pinned<uint32_t>(n);
  uint32_t* vals = pinned<uint32_t>(n);
  uint32_t* out = pinned<uint32_t>(n);
  uint32_t* upper = pinned<uint32_t>(n);
  uint32_t* counts = pinned<uint32_t>(n);
  pair_t* pairs = pinned<pair_t>((size_t)n * L);
  for (uint32_t i = 0; i < n; ++i) keys[i] = 2 * (i + 1);
  shuffle(keys, n, 11);
  for (uint32_t i = 0; i < n; ++i) vals[i] = keys[i];
  const uint32_t half = n / 2;
  tree.insert(keys, vals, half);
  uint32_t s0 = tree.take_snapshot();
  tree.insert(keys + half, vals + half, n - half);
  uint32_t s1 = tree.take_snapshot();
  EXPECT(s1 == s0 + 1, "snapshot ids %u %u", s0, s1);
  // FindExist / FindNotExist at the snapshot taken after the first half.
  tree.find(keys, out, n, s0);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t want = i < half ? keys[i] : MVBT_NOT_FOUND;
    EXPECT(out[i] == want, "find@s0 key %u got %u want %u", keys[i], out[i], want);
  }
  tree.find(keys, out, n, s1);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == keys[i], "find@s1 key %u got %u", keys[i], out[i]);
  // RangeQueryTest: [k, k+2L) at s1 holds k, k+2, ..., up to the last key 2n.
  for (uint32_t i = 0; i < n; ++i) upper[i] = keys[i] + 2 * L;
  tree.range_query(keys, upper, pairs, counts, L, n, s1);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t k = keys[i];
    uint32_t want = (2 * n - k) / 2 + 1 < L ? (2 * n - k) / 2 + 1 : L;
    EXPECT(counts[i] == want, "range count at %u: %u want %u", k, counts[i], want);
    for (uint32_t j = 0; j < want && j < counts[i]; ++j) {
      pair_t p = pairs[(size_t)i * L + j];
      EXPECT(p.first == k + 2 * j && p.second == k + 2 * j, "range pair %u of [%u): {%u,%u}", j, k,
             p.first, p.second);
    }
  }
  // Versioned erase: gone at latest, still visible at s1.
  tree.erase(keys, half);
  uint32_t s2 = tree.take_snapshot();
  tree.find(keys, out, n, s2);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t want = i < half ? MVBT_NOT_FOUND : keys[i];
    EXPECT(out[i] == want, "find@s2 key %u got %u want %u", keys[i], out[i], want);
  }
  tree.find(keys, out, n, s1);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == keys[i], "find@s1 after erase key %u got %u", keys[i], out[i]);
  // Reclamation keeps what live snapshots read.
  tree.release_snapshot(s0);
  tree.reclaim();
  tree.find(keys, out, n, s1);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == keys[i], "find@s1 after reclaim key %u got %u", keys[i], out[i]);
  EXPECT(tree.validate_tree_structure(), "versioned structure");
  EXPECT(tree.error_flags() == 0, "device error flags 0x%x", tree.error_flags());
  (void)hipHostFree(keys);
  (void)hipHostFree(vals);
  (void)hipHostFree(out);
  (void)hipHostFree(upper);
  (void)hipHostFree(counts);
  (void)hipHostFree(pairs);
}

// Reference bulk-build constructor: sorted keys 2, 4, ..., 2n; then finds of
// every key and every odd (absent) key, and a few inserts on the loaded tree.
void bulk_cases(uint32_t n) {
  uint32_t* keys = pinned<uint32_t>(n);
  uint32_t* vals = pinned<uint32_t>(n);
  uint32_t* absent = pinned<uint32_t>(n);
  uint32_t* out = pinned<uint32_t>(n);
  for (uint32_t i = 0; i < n; ++i) {
    keys[i] = 2 * (i + 1);
    vals[i] = keys[i] ^ 0x9E3779B9u;
    absent[i] = keys[i] + 1;
  }
  GpuBTree::gpu_blink_tree<uint32_t, uint32_t, 16> tree(keys, vals, n, /*sorted_input=*/true);
  uint64_t count = 0;
  EXPECT(tree.validate_tree_structure(&count), "structure after bulk build");
  EXPECT(count == n, "bulk key count %llu != %u", (unsigned long long)count, n);
  tree.find(keys, out, n);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == vals[i], "bulk FindExist key %u got %u", keys[i], out[i]);
  tree.find(absent, out, n);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == MVBT_NOT_FOUND, "bulk FindNotExist key %u got %u", absent[i], out[i]);
  tree.insert(absent, absent, n);
  tree.find(absent, out, n);
  (void)hipDeviceSynchronize();
  for (uint32_t i = 0; i < n; ++i) EXPECT(out[i] == absent[i], "insert after bulk key %u got %u", absent[i], out[i]);
  EXPECT(tree.validate_tree_structure(&count) && count == 2ull * n, "structure after inserts");
  bool threw = false;
  try {
    GpuBTree::gpu_blink_tree<uint32_t, uint32_t, 16> unsorted(keys, vals, n, /*sorted_input=*/false);
  } catch (const GpuBTree::error&) {
    threw = true;
  }
  EXPECT(threw, "unsorted bulk build must throw");
  (void)hipHostFree(keys);
  (void)hipHostFree(vals);
  (void)hipHostFree(absent);
  (void)hipHostFree(out);
}

}  // namespace

int main(int argc, char** argv) {
  uint32_t n = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 1024;
  uint32_t L = argc > 2 ? (uint32_t)std::strtoul(argv[2], nullptr, 10) : 2;
  try {
    plain_cases<GpuBTree::gpu_blink_tree<uint32_t, uint32_t, 16>>(n);
    plain_cases<GpuBTree::gpu_versioned_btree<uint32_t, uint32_t, 16>>(n);
    versioned_cases(n, L);
    bulk_cases(n);
  } catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 2;
  }
  if (g_errors) {
    std::printf("FAILED %d checks\n", g_errors);
    return 1;
  }
  std::printf("OK n=%u L=%u\n", n, L);
  return 0;
}
