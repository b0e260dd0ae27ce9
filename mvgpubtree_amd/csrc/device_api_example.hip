// User kernels on the device-side cooperative API (include/mvbt_device.hpp):
// cooperative_insert from a user kernel, take_snapshot(tile) inside a kernel,
// then one kernel where some tiles insert while other tiles read that snapshot
// with cooperative_find and concurrent_cooperative_range_query. The expected
// values follow from closed forms (keys 2(i+1), value = key, as in the
// reference unit tests test/unittest_versioning.cu:82-95). Prints "OK".
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
using tile_t = GpuBTree::device::tile<B>;

__device__ uint32_t tile_id() { return (blockIdx.x * blockDim.x + threadIdx.x) / B; }
__device__ uint32_t tile_count() { return gridDim.x * blockDim.x / B; }

__global__ void insert_kernel(mvbt_device_handle h, const uint32_t* keys, uint32_t n) {
  tree_t tree(h);
  tile_t t;
  auto ctx = tree.begin_writes(t);
  for (uint32_t i = tile_id(); i < n; i += tile_count()) tree.cooperative_insert(keys[i], keys[i], t, ctx);
  tree.end_writes(ctx, t);
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
// snapshot into pairs + i*L; [2n, 2n + late): insert keys[half + j].
__global__ void mixed_kernel(mvbt_device_handle h, const uint32_t* keys, uint32_t n, uint32_t half,
                             const uint32_t* snap, uint32_t* found, mvbt_pair* pairs,
                             uint32_t* counts, uint32_t L) {
  tree_t tree(h);
  tile_t t;
  const uint32_t s = *snap;
  auto ctx = tree.begin_writes(t);
  const uint32_t total = 3 * n;
  for (uint32_t i = tile_id(); i < total; i += tile_count()) {
    if (i < n) {
      uint32_t v = tree.cooperative_find(keys[i], t, s, /*concurrent=*/true);
      if (t.leader()) found[i] = v;
    } else if (i < 2 * n) {
      uint32_t q = i - n;
      uint32_t c = tree.concurrent_cooperative_range_query(keys[q], keys[q] + 2 * L, t, s,
                                                           pairs + (size_t)q * L, L, true);
      if (t.leader()) counts[q] = c;
    } else if (i - 2 * n < n - half) {
      uint32_t k = keys[half + (i - 2 * n)];
      tree.cooperative_insert(k, k, t, ctx);
    }
  }
  tree.end_writes(ctx, t);
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
  MVBT_OK_(mvbt_destroy(tree));
  if (errors) {
    std::printf("FAILED %d checks\n", errors);
    return 1;
  }
  std::printf("OK n=%u snapshot=%u\n", n, snap);
  return 0;
}
