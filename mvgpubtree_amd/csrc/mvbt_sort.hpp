// Stable LSD radix sort of 32-bit keys with 32-bit payloads, for the sorted
// batched form (mvbt_batch.hpp): it groups a call's writes by the leaf they
// route to while keeping operation order inside each group.
//
// The element count is read from device memory (`n`), so the kernels can sit
// in a graph whose parameters never change while the count does. Every pass
// splits the input into gridDim.x contiguous segments, one per workgroup:
//   1. radix_hist_kernel    per-segment histogram of one 8-bit digit
//                           (hist[digit * G + segment]);
//   2. radix_scan_kernel    one workgroup per digit: exclusive scan of its G
//                           segment counts in place, and the digit's total;
//   3. radix_scatter_kernel each segment writes its elements, in input
//                           order, at digit base + earlier segments' count +
//                           rank inside the segment (stable).
// Passes alternate between two buffer pairs. Memory traffic per pass is one
// read of the keys (histogram) and one read plus one scattered write of
// keys and payloads (scatter): 20 bytes per element.
#pragma once
#include "mvbt_common.hpp"

namespace mvbt {

constexpr uint32_t kSortSegments = 512;  // workgroups (segments) per pass
constexpr uint32_t kSortBlock = 256;     // threads per workgroup
constexpr uint32_t kRadix = 256;         // 8-bit digits

MVBT_DEV void segment_of(uint32_t n, uint32_t g, uint32_t G, uint32_t& a, uint32_t& b) {
  const uint32_t seg = (n + G - 1) / G;
  const uint64_t lo = (uint64_t)g * seg;
  a = lo < n ? (uint32_t)lo : n;
  b = (uint64_t)a + seg < n ? a + seg : n;
}

__global__ void __launch_bounds__(kSortBlock)
    radix_hist_kernel(const uint32_t* keys, const uint32_t* n_ptr, uint32_t shift, uint32_t* hist) {
  __shared__ uint32_t h[kRadix];
  const uint32_t tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  uint32_t a, b;
  segment_of(*n_ptr, blockIdx.x, gridDim.x, a, b);
  for (uint32_t i = a + tid; i < b; i += kSortBlock) atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1)], 1u);
  __syncthreads();
  hist[tid * gridDim.x + blockIdx.x] = h[tid];
}

// One workgroup per digit (gridDim.x = 256), G = `segments` counts each.
__global__ void __launch_bounds__(kSortBlock)
    radix_scan_kernel(uint32_t* hist, uint32_t segments, uint32_t* totals) {
  __shared__ uint32_t part[kSortBlock];
  const uint32_t tid = threadIdx.x, d = blockIdx.x;
  uint32_t* row = hist + (size_t)d * segments;
  const uint32_t per = (segments + kSortBlock - 1) / kSortBlock;
  const uint32_t a = tid * per, b = a + per < segments ? a + per : segments;
  uint32_t s = 0;
  for (uint32_t i = a; i < b; ++i) s += row[i];
  part[tid] = s;
  __syncthreads();
  for (uint32_t off = 1; off < kSortBlock; off <<= 1) {  // inclusive scan (Hillis-Steele)
    const uint32_t v = tid >= off ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = part[tid] - s;  // exclusive
  for (uint32_t i = a; i < b; ++i) {
    const uint32_t c = row[i];
    row[i] = run;
    run += c;
  }
  if (tid == kSortBlock - 1) totals[d] = part[tid];
}

__global__ void __launch_bounds__(kSortBlock)
    radix_scatter_kernel(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                         const uint32_t* n_ptr, uint32_t shift, const uint32_t* hist, const uint32_t* totals) {
  constexpr uint32_t kWaves = kSortBlock / 64;
  __shared__ uint32_t base[kRadix];
  __shared__ uint32_t wcnt[kWaves][kRadix];
  __shared__ uint32_t dsum[kRadix];
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6, G = gridDim.x;
  // digit bases: exclusive scan of the digit totals
  dsum[tid] = totals[tid];
  __syncthreads();
  for (uint32_t off = 1; off < kRadix; off <<= 1) {
    const uint32_t v = tid >= off ? dsum[tid - off] : 0u;
    __syncthreads();
    dsum[tid] += v;
    __syncthreads();
  }
  base[tid] = dsum[tid] - totals[tid] + hist[tid * G + blockIdx.x];
#pragma unroll
  for (uint32_t v = 0; v < kWaves; ++v) wcnt[v][tid] = 0;
  __syncthreads();
  uint32_t a, b;
  segment_of(*n_ptr, blockIdx.x, G, a, b);
  const uint64_t below = (1ull << lane) - 1;
  // Per chunk: counts in (thread d owns column d of wcnt), sync, scatter,
  // sync, fold the counts into base and clear column d, sync.
  for (uint32_t c = a; c < b; c += kSortBlock) {
    const uint32_t i = c + tid;
    const bool act = i < b;
    const uint32_t k = act ? kin[i] : 0u;
    const uint32_t d = (k >> shift) & (kRadix - 1);
    // lanes of this wave holding the same digit
    uint64_t same = __builtin_amdgcn_ballot_w64(act);
#pragma unroll
    for (uint32_t bit = 0; bit < 8; ++bit) {
      const uint64_t ones = __builtin_amdgcn_ballot_w64(((d >> bit) & 1u) != 0);
      same &= ((d >> bit) & 1u) ? ones : ~ones;
    }
    const uint32_t rank = (uint32_t)__builtin_popcountll(same & below);
    if (act && rank == 0) wcnt[w][d] = (uint32_t)__builtin_popcountll(same);
    __syncthreads();
    if (act) {
      uint32_t pos = base[d] + rank;
      for (uint32_t v = 0; v < w; ++v) pos += wcnt[v][d];
      kout[pos] = k;
      vout[pos] = vin[i];
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (uint32_t v = 0; v < kWaves; ++v) {
      add += wcnt[v][tid];
      wcnt[v][tid] = 0;
    }
    base[tid] += add;
    __syncthreads();
  }
}

}  // namespace mvbt
