// The plain per-batch launch path as a reference-style host caller would
// write it, in C++ over the C ABI (include/mvbt.h): for every batch take a
// snapshot, launch find and range_query at it, then insert and erase, then
// release the snapshot - one launch per operation type, the pattern of the
// reference's host API (gpu_versioned_blink_tree.hpp:106-302; its benchmarks
// call take_snapshot before each batch, benchmarks/concurrent_insert_range_bench.cu).
// It is the baseline the batched and served forms are measured against, so
// that comparison measures device paths, not an interpreter's loop.
//
// Inputs are device arrays grouped by operation type; batch b's finds are
// f_keys[f_off[b] .. f_off[b+1]) (host offsets), and likewise for ranges,
// inserts and erases. Find results go to f_out at the same indices; range i
// writes counts r_counts[i] and pairs at r_pairs + i * L, or, with
// `pair_window` set, every batch's pairs start at r_pairs (a result buffer
// the caller consumes batch by batch).
#include <hip/hip_runtime.h>

#include <chrono>
#include <thread>

#include "../../include/mvbt.h"

extern "C" {

// Runs batches [0, nb) back to back on `stream` without waiting for the
// device, reclaiming every `reclaim_every` batches (0: never). Returns the
// first failing status (with mvbt_last_error set).
mvbt_status mvbt_driver_per_batch(mvbt_tree* t, uint32_t nb, const uint64_t* f_off, const uint32_t* f_keys,
                                  uint32_t* f_out, const uint64_t* r_off, const uint32_t* r_lo, const uint32_t* r_hi,
                                  mvbt_pair* r_pairs, uint32_t* r_counts, uint32_t L, const uint64_t* i_off,
                                  const uint32_t* i_keys, const uint32_t* i_vals, const uint64_t* e_off,
                                  const uint32_t* e_keys, uint32_t reclaim_every, uint32_t pair_window,
                                  void* stream) {
  for (uint32_t b = 0; b < nb; ++b) {
    uint32_t s = 0;
    if (mvbt_status rc = mvbt_take_snapshot(t, stream, &s)) return rc;
    if (uint64_t n = f_off[b + 1] - f_off[b])
      if (mvbt_status rc = mvbt_find(t, f_keys + f_off[b], f_out + f_off[b], n, s, stream, 0)) return rc;
    if (uint64_t n = r_off[b + 1] - r_off[b])
      if (mvbt_status rc = mvbt_range_query(t, r_lo + r_off[b], r_hi + r_off[b],
                                            r_pairs ? r_pairs + (pair_window ? 0 : r_off[b] * (uint64_t)L) : nullptr,
                                            r_counts ? r_counts + r_off[b] : nullptr, L, n, s, stream, 0))
        return rc;
    if (uint64_t n = i_off[b + 1] - i_off[b])
      if (mvbt_status rc = mvbt_insert(t, i_keys + i_off[b], i_vals + i_off[b], n, stream, 0)) return rc;
    if (uint64_t n = e_off[b + 1] - e_off[b])
      if (mvbt_status rc = mvbt_erase(t, e_keys + e_off[b], n, stream, 0)) return rc;
    if (mvbt_status rc = mvbt_release_snapshot(t, s)) return rc;
    if (reclaim_every && (b + 1) % reclaim_every == 0)
      if (mvbt_status rc = mvbt_reclaim(t, stream)) return rc;
  }
  return MVBT_OK;
}

// The plain per-request server: request b (one batch) is issued when the host
// clock passes arrival_s[b] (seconds after the call starts) and the previous
// request has completed, and is waited for; done_s[b] receives its completion
// time. Open-loop arrivals: a request that arrives while another runs waits.
mvbt_status mvbt_driver_per_request(mvbt_tree* t, uint32_t nb, const double* arrival_s, double* done_s,
                                    const uint64_t* f_off, const uint32_t* f_keys, uint32_t* f_out,
                                    const uint64_t* r_off, const uint32_t* r_lo, const uint32_t* r_hi,
                                    mvbt_pair* r_pairs, uint32_t* r_counts, uint32_t L, const uint64_t* i_off,
                                    const uint32_t* i_keys, const uint32_t* i_vals, uint32_t reclaim_every,
                                    uint32_t pair_window, void* stream) {
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  auto now = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
  static const uint64_t zero[2] = {0, 0};
  for (uint32_t b = 0; b < nb; ++b) {
    while (now() < arrival_s[b]) {
      const double left = arrival_s[b] - now();
      if (left > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(left < 1e-3 ? left : 1e-3));
    }
    // one request = one batch of the per-batch path (offsets shifted to b)
    mvbt_status rc = mvbt_driver_per_batch(t, 1, f_off + b, f_keys, f_out, r_off + b, r_lo, r_hi, r_pairs, r_counts,
                                           L, i_off + b, i_keys, i_vals, zero, nullptr, 0, pair_window, stream);
    if (rc) return rc;
    if (reclaim_every && (b + 1) % reclaim_every == 0)
      if (mvbt_status rc2 = mvbt_reclaim(t, stream)) return rc2;
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return MVBT_ERR_HIP;
    done_s[b] = now();
  }
  return MVBT_OK;
}

}  // extern "C"
