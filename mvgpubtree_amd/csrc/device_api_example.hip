// User kernels on the device-side cooperative API (include/mvbt_device.hpp),
// written with the reference's argument lists: the three insert forms from a
// user kernel, take_snapshot(tile) inside a kernel, then a kernel where some
// tiles insert while others read that snapshot (cooperative_find with a
// timestamp, concurrent_cooperative_range_query), then both erase forms beside
// concurrent_cooperative_find at the snapshot and latest-version finds; and
// the plain tree's insert, concurrent erase, find and range query. The
// expected values follow from closed forms (keys 2(i+1), value = key, as in
// the reference unit tests test/unittest_versioning.cu:82-95). Prints "OK".
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mvbt.h"
#include "mvbt_device.hpp"

#define HIP_OK(x)                                                          \
  do {                                                                     \
    hipError_t e = (x);                                                    \
    if (e != hipSuccess) {                                                 \
      std::printf("FAIL %s: %s\n", #x, hipGetErrorString(e));              \
      std::exit(2);                                                        \
    }                                                                      \
  } while (0)
#define MVBT_OK_(x)                                                        \
  do {                                                                     \
    if ((x) != MVBT_OK) {                                                  \
      std::printf("FAIL %s: %s\n", #x, mvbt_last_error());                 \
      std::exit(2);                                                        \
    }                                                                      \
  } while (0)

constexpr int B = 16;
using tree_t = GpuBTree::device::versioned_tree<B>;
using plain_t = GpuBTree::device::blink_tree<B>;
using tile_t = GpuBTree::device::tile<B>;

__device__ uint32_t tile_id() { return (blockIdx.x * blockDim.x + threadIdx.x) / B; }
__device__ uint32_t tile_count() { return gridDim.x * blockDim.x / B; }

// The reference's insert forms (gpu_versioned_blink_tree.hpp:306-330), one per
// key in turn.
__global__ void insert_kernel(mvbt_device_handle h, const uint32_t* keys, uint32_t n) {
  tree_t tree(h);
  tile_t t;
  auto allocator = tree.begin_writes(t);
  auto reclaimer = tree.reclaimer(allocator);
  for (uint32_t i = tile_id(); i < n; i += tile_count()) {
    if (i % 3 == 0) tree.cooperative_insert(keys[i], keys[i], t, allocator, reclaimer);
    else if (i % 3 == 1) tree.cooperative_insert_in_place(keys[i], keys[i], t, allocator);
    else tree.cooperative_insert_out_of_place(keys[i], keys[i], t, allocator, reclaimer);
  }
  tree.end_writes(allocator, t);
}

__global__ void snapshot_kernel(mvbt_device_handle h, uint32_t* out) {
  tree_t tree(h);
  tile_t t;
  if (threadIdx.x < B) {
    uint32_t s = tree.take_snapshot(t);
    if (t.leader()) *out = s;
  }
}

// Ops [0, n): find keys[i] at snapshot; [n, 2n): range [keys[i], keys[i]+2L) at
// snapshot into pairs + i*L (the reference's unbounded form: such a range
// holds at most L keys); [2n, 3n - half): insert keys[half + j].
__global__ void mixed_kernel(mvbt_device_handle h, const uint32_t* keys, uint32_t n, uint32_t half,
                             const uint32_t* snap, uint32_t* found, mvbt_pair* pairs,
                             uint32_t* counts, uint32_t L) {
  tree_t tree(h);
  tile_t t;
  const uint32_t s = *snap;
  auto allocator = tree.begin_writes(t);
  auto reclaimer = tree.reclaimer(allocator);
  const uint32_t total = 3 * n;
  for (uint32_t i = tile_id(); i < total; i += tile_count()) {
    if (i < n) {
      uint32_t v = tree.cooperative_find(keys[i], t, allocator, s, /*concurrent=*/true);
      if (t.leader()) found[i] = v;
    } else if (i < 2 * n) {
      uint32_t q = i - n;
      uint32_t c = tree.concurrent_cooperative_range_query(keys[q], keys[q] + 2 * L, t, allocator, s,
                                                           pairs + (size_t)q * L);
      if (t.leader()) counts[q] = c;
    } else if (i - 2 * n < n - half) {
      uint32_t k = keys[half + (i - 2 * n)];
      tree.cooperative_insert(k, k, t, allocator, reclaimer);
    }
  }
  tree.end_writes(allocator, t);
}

// After the mixed kernel: keys[i] for i % 4 == 0 are erased (both erase
// forms) while other tiles read: concurrent_cooperative_find at the
// snapshot, and the latest version with cooperative_find(.., concurrent).
// Latest reads of keys that are never erased must hit; snapshot reads must
// equal the snapshot. erased[i] gets the erase result.
__global__ void erase_kernel(mvbt_device_handle h, const uint32_t* keys, uint32_t n, const uint32_t* snap,
                             uint32_t* at_snap, uint32_t* latest, uint32_t* erased) {
  tree_t tree(h);
  tile_t t;
  const uint32_t s = *snap;
  auto allocator = tree.begin_writes(t);
  auto reclaimer = tree.reclaimer(allocator);
  for (uint32_t i = tile_id(); i < 2 * n; i += tile_count()) {
    if (i < n) {
      const uint32_t k = keys[i];
      if (i % 4 == 0) {
        const bool ok = (i % 8 == 0) ? tree.cooperative_erase(k, t, allocator, reclaimer)
                                     : tree.cooperative_erase(k, t, allocator, /*concurrent=*/true);
        if (t.leader()) erased[i] = ok ? 1u : 0u;
      } else {
        const uint32_t v = tree.cooperative_find(k, t, allocator, /*concurrent=*/true);
        if (t.leader()) latest[i] = v;
      }
    } else {
      const uint32_t q = i - n;
      const uint32_t v = tree.concurrent_cooperative_find(keys[q], t, allocator, s);
      if (t.leader()) at_snap[q] = v;
    }
  }
  tree.end_writes(allocator, t);
}

// Plain tree (gpu_blink_tree.hpp:269-390): inserts, then erases of every
// fourth key beside concurrent finds and range queries of the others.
__global__ void plain_insert_kernel(mvbt_device_handle h, const uint32_t* keys, uint32_t n) {
  plain_t tree(h);
  tile_t t;
  auto allocator = tree.begin_writes(t);
  for (uint32_t i = tile_id(); i < n; i += tile_count()) tree.cooperative_insert(keys[i], keys[i], t, allocator);
  tree.end_writes(allocator, t);
}
__global__ void plain_mixed_kernel(mvbt_device_handle h, const uint32_t* keys, uint32_t n, uint32_t* found,
                                   uint32_t* erased, mvbt_pair* pairs, uint32_t* counts) {
  plain_t tree(h);
  tile_t t;
  auto allocator = tree.begin_writes(t);
  for (uint32_t i = tile_id(); i < 2 * n; i += tile_count()) {
    if (i < n) {
      const uint32_t k = keys[i];
      if (i % 4 == 0) {
        const bool ok = tree.cooperative_erase(k, t, allocator, /*concurrent=*/true);
        if (t.leader()) erased[i] = ok ? 1u : 0u;
      } else {
        const uint32_t v = tree.cooperative_find(k, t, allocator, /*concurrent=*/true);
        if (t.leader()) found[i] = v;
      }
    } else {
      // [k + 1, k + 2): holds no key (keys are even), so the count is 0
      const uint32_t q = i - n;
      const uint32_t c = tree.cooperative_range_query(keys[q] + 1, keys[q] + 2, t, allocator, pairs + q, true);
      if (t.leader()) counts[q] = c;
    }
  }
  tree.end_writes(allocator, t);
}

int main(int argc, char** argv) {
  const uint32_t n = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 100000;
  const uint32_t L = 8, half = n / 2;
  std::vector<uint32_t> keys(n);
  for (uint32_t i = 0; i < n; ++i) keys[i] = 2 * (i + 1);
  // first half and second half are each shuffled (insertion order), the split
  // into halves stays by key rank so the snapshot holds keys 2..2*half
  uint64_t seed = 12345;
  auto shuffle = [&](uint32_t lo, uint32_t hi) {
    for (uint32_t i = hi; i > lo + 1; --i) {
      seed = seed * 6364136223846793005ull + 1442695040888963407ull;
      uint32_t j = lo + (uint32_t)((seed >> 33) % (i - lo));
      std::swap(keys[i - 1], keys[j]);
    }
  };
  shuffle(0, half);
  shuffle(half, n);

  mvbt_config cfg{};
  cfg.kind = MVBT_VERSIONED;
  cfg.node_width = B;
  cfg.pool_nodes = 1u << 20;
  cfg.device = -1;
  mvbt_tree* tree = nullptr;
  MVBT_OK_(mvbt_create(&cfg, &tree));
  mvbt_device_handle h;
  MVBT_OK_(mvbt_device_view(tree, &h));

  uint32_t *d_keys, *d_snap, *d_found, *d_counts;
  mvbt_pair* d_pairs;
  HIP_OK(hipMalloc(&d_keys, n * 4));
  HIP_OK(hipMalloc(&d_snap, 4));
  HIP_OK(hipMalloc(&d_found, n * 4));
  HIP_OK(hipMalloc(&d_counts, n * 4));
  HIP_OK(hipMalloc(&d_pairs, (size_t)n * L * sizeof(mvbt_pair)));
  HIP_OK(hipMemcpy(d_keys, keys.data(), n * 4, hipMemcpyHostToDevice));

  hipLaunchKernelGGL(insert_kernel, dim3(512), dim3(256), 0, 0, h, d_keys, half);
  hipLaunchKernelGGL(snapshot_kernel, dim3(1), dim3(64), 0, 0, h, d_snap);
  hipLaunchKernelGGL(mixed_kernel, dim3(1024), dim3(256), 0, 0, h, d_keys, n, half, d_snap, d_found,
                     d_pairs, d_counts, L);
  HIP_OK(hipGetLastError());
  HIP_OK(hipDeviceSynchronize());

  std::vector<uint32_t> found(n), counts(n), latest(n);
  std::vector<mvbt_pair> pairs((size_t)n * L);
  uint32_t snap = 0;
  HIP_OK(hipMemcpy(&snap, d_snap, 4, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(found.data(), d_found, n * 4, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(counts.data(), d_counts, n * 4, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(pairs.data(), d_pairs, pairs.size() * sizeof(mvbt_pair), hipMemcpyDeviceToHost));
  MVBT_OK_(mvbt_find(tree, d_keys, d_found, n, MVBT_LATEST, nullptr, 0));
  HIP_OK(hipMemcpy(latest.data(), d_found, n * 4, hipMemcpyDeviceToHost));

  int errors = 0;
  auto fail = [&](const char* what, uint32_t i, uint32_t got, uint32_t want) {
    if (errors++ < 10) std::printf("FAIL %s at %u: got %u want %u\n", what, i, got, want);
  };
  const uint32_t max_key = 2 * half;  // the snapshot holds keys 2..2*half
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t k = keys[i];
    uint32_t want = k <= max_key ? k : MVBT_NOT_FOUND;
    if (found[i] != want) fail("find@snapshot", i, found[i], want);
    if (latest[i] != k) fail("find@latest", i, latest[i], k);
    uint32_t wc = k > max_key ? 0 : std::min<uint32_t>(L, (max_key - k) / 2 + 1);
    if (counts[i] != wc) fail("range count", i, counts[i], wc);
    for (uint32_t j = 0; j < wc && j < counts[i]; ++j) {
      mvbt_pair p = pairs[(size_t)i * L + j];
      if (p.first != k + 2 * j || p.second != k + 2 * j) fail("range pair", i, p.first, k + 2 * j);
    }
  }
  uint64_t nk = 0;
  uint32_t flags = 0;
  MVBT_OK_(mvbt_validate(tree, &nk, nullptr));
  MVBT_OK_(mvbt_error_flags(tree, &flags));
  if (nk != n) fail("key count", 0, (uint32_t)nk, n);
  if (flags) fail("device flags", 0, flags, 0);

  // erases beside snapshot and latest readers
  uint32_t *d_at_snap, *d_latest, *d_erased;
  HIP_OK(hipMalloc(&d_at_snap, n * 4));
  HIP_OK(hipMalloc(&d_latest, n * 4));
  HIP_OK(hipMalloc(&d_erased, n * 4));
  hipLaunchKernelGGL(erase_kernel, dim3(1024), dim3(256), 0, 0, h, d_keys, n, d_snap, d_at_snap, d_latest, d_erased);
  HIP_OK(hipGetLastError());
  HIP_OK(hipDeviceSynchronize());
  std::vector<uint32_t> at_snap(n), lat(n), erased(n);
  HIP_OK(hipMemcpy(at_snap.data(), d_at_snap, n * 4, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(lat.data(), d_latest, n * 4, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(erased.data(), d_erased, n * 4, hipMemcpyDeviceToHost));
  MVBT_OK_(mvbt_find(tree, d_keys, d_found, n, MVBT_LATEST, nullptr, 0));
  HIP_OK(hipMemcpy(latest.data(), d_found, n * 4, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t k = keys[i];
    const uint32_t want_snap = k <= max_key ? k : MVBT_NOT_FOUND;
    if (at_snap[i] != want_snap) fail("concurrent find@snapshot", i, at_snap[i], want_snap);
    if (i % 4 == 0) {
      if (erased[i] != 1) fail("erase result", i, erased[i], 1);
      if (latest[i] != MVBT_NOT_FOUND) fail("erased key still found", i, latest[i], MVBT_NOT_FOUND);
    } else {
      if (lat[i] != k) fail("concurrent find@latest", i, lat[i], k);
      if (latest[i] != k) fail("find@latest after erases", i, latest[i], k);
    }
  }
  MVBT_OK_(mvbt_validate(tree, &nk, nullptr));
  const uint32_t n_erased = (n + 3) / 4;
  if (nk != n - n_erased) fail("key count after erases", 0, (uint32_t)nk, n - n_erased);
  MVBT_OK_(mvbt_error_flags(tree, &flags));
  if (flags) fail("device flags", 1, flags, 0);
  MVBT_OK_(mvbt_destroy(tree));

  // plain tree
  cfg.kind = MVBT_PLAIN;
  MVBT_OK_(mvbt_create(&cfg, &tree));
  MVBT_OK_(mvbt_device_view(tree, &h));
  hipLaunchKernelGGL(plain_insert_kernel, dim3(512), dim3(256), 0, 0, h, d_keys, n);
  hipLaunchKernelGGL(plain_mixed_kernel, dim3(1024), dim3(256), 0, 0, h, d_keys, n, d_found, d_erased, d_pairs,
                     d_counts);
  HIP_OK(hipGetLastError());
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipMemcpy(lat.data(), d_found, n * 4, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(erased.data(), d_erased, n * 4, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(counts.data(), d_counts, n * 4, hipMemcpyDeviceToHost));
  MVBT_OK_(mvbt_find(tree, d_keys, d_found, n, MVBT_LATEST, nullptr, 0));
  HIP_OK(hipMemcpy(latest.data(), d_found, n * 4, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t k = keys[i];
    if (counts[i] != 0) fail("plain range count", i, counts[i], 0);
    if (i % 4 == 0) {
      if (erased[i] != 1) fail("plain erase result", i, erased[i], 1);
      if (latest[i] != MVBT_NOT_FOUND) fail("plain erased key found", i, latest[i], MVBT_NOT_FOUND);
    } else {
      if (lat[i] != k) fail("plain concurrent find", i, lat[i], k);
      if (latest[i] != k) fail("plain find after erases", i, latest[i], k);
    }
  }
  MVBT_OK_(mvbt_validate(tree, &nk, nullptr));
  if (nk != n - n_erased) fail("plain key count", 0, (uint32_t)nk, n - n_erased);
  MVBT_OK_(mvbt_error_flags(tree, &flags));
  if (flags) fail("plain device flags", 0, flags, 0);
  MVBT_OK_(mvbt_destroy(tree));
  if (errors) {
    std::printf("FAILED %d checks\n", errors);
    return 1;
  }
  std::printf("OK n=%u snapshot=%u\n", n, snap);
  return 0;
}
