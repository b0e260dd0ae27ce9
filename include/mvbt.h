/*
 * mvbt.h - C ABI of the MI355X-native multiversion B-link tree.
 *
 * One handle owns one tree (plain B-link or multiversion) on one HIP device:
 * its node pool, counters, free stack and retire ring. The caller owns every
 * I/O buffer; buffers are device memory (hipMalloc) or pinned host memory
 * (hipHostMalloc), like the reference's tests. All batch calls are
 * asynchronous on the given stream (`void*` = hipStream_t, NULL = default
 * stream) unless documented otherwise, and never allocate device memory.
 *
 * Each entry point names the reference interface it replaces
 * (owensgroup/MVGpuBTree, paths under include/).
 *
 * Keys and values are uint32. Key 0 and key 0xFFFFFFFF are reserved (the
 * reference uses them as its tombstone and empty sentinel): insert/erase of a
 * reserved key is skipped and raises MVBT_FLAG_INVALID_KEY; find returns
 * MVBT_NOT_FOUND for it. A find miss returns MVBT_NOT_FOUND (0xFFFFFFFF), as in
 * the reference (versioned_node.hpp:267-271).
 */
#ifndef MVBT_H_
#define MVBT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MVBT_NOT_FOUND 0xFFFFFFFFu
#define MVBT_LATEST 0xFFFFFFFFu /* snapshot id meaning "latest version" */

typedef enum mvbt_status {
  MVBT_OK = 0,
  MVBT_ERR_INVALID_ARG = 1,  /* null handle/pointer, bad config, n too large */
  MVBT_ERR_OUT_OF_MEMORY = 2,/* device allocation failed at create */
  MVBT_ERR_HIP = 3,          /* a HIP runtime call failed; see mvbt_last_error() */
  MVBT_ERR_UNSUPPORTED = 4,  /* operation not available for this tree kind */
  MVBT_ERR_DEVICE_FLAGS = 5  /* device error flags are set; see mvbt_error_flags() */
} mvbt_status;

typedef enum mvbt_kind {
  MVBT_PLAIN = 0,    /* GpuBTree::gpu_blink_tree        (gpu_blink_tree.hpp:63-1194) */
  MVBT_VERSIONED = 1 /* GpuBTree::gpu_versioned_btree   (gpu_versioned_blink_tree.hpp:71-1719) */
} mvbt_kind;

/* Device error flags (sticky until mvbt_clear_error_flags). */
#define MVBT_FLAG_POOL_EXHAUSTED 0x1u  /* node pool full: some writes were dropped */
#define MVBT_FLAG_RETIRE_OVERFLOW 0x2u /* retire ring full: some copies are never freed */
#define MVBT_FLAG_INVALID_KEY 0x4u     /* reserved key passed to insert/erase */
#define MVBT_FLAG_VERSION_GAP 0x8u     /* read at a snapshot older than the data */
#define MVBT_FLAG_RANGE_OVERFLOW 0x10u /* a range query matched more than L pairs */
#define MVBT_FLAG_CORRUPT 0x20u        /* a traversal left the pool or hit its step bound */
#define MVBT_FLAG_SNAPSHOT_ID 0x40u    /* a snapshot id returned without waiting differs from the
                                          device clock (the library's own consistency check) */
#define MVBT_FLAG_STAMP_ORDER 0x80u    /* a write stamped by its launch (batched form, mixed calls)
                                          met a leaf stamped later by an overlapping writer (user
                                          kernels through a device view); it was applied at the
                                          later stamp */
#define MVBT_FLAG_BAD_BATCH 0x100u     /* mvbt_run_batches rejected a call (a batch writes one key
                                          twice, an unknown op type, or offsets that disagree
                                          with num_ops); nothing of it was applied */

/* Operation flags. */
#define MVBT_CONCURRENT 0x1u /* readers may run while writers run on another stream
                                (reference `bool concurrent`): coherent loads; a snapshot
                                read validates each leaf against its version word and lock
                                word, a read of the latest version (MVBT_LATEST, and every
                                read of a plain tree) reads each leaf under its lock */

typedef struct mvbt_config {
  uint32_t kind;            /* mvbt_kind */
  uint32_t node_width;      /* pairs per node: 16 (128 B), 32 or 64; 0 = 16 */
  uint64_t pool_nodes;      /* node slots (< 2^29); 0 = 1<<22 */
  uint64_t retire_capacity; /* retire ring entries (power of two); raised to at least
                               pool_nodes (every copy awaiting reclamation holds a pool slot,
                               so the ring then never fills); 0 = that minimum */
  int32_t device;           /* HIP device ordinal; -1 = current device */
  uint32_t reserved[3];
} mvbt_config;

/* {key, value}; same 8-byte layout as the reference pair_type
 * (pair_type.hpp:21-64). */
typedef struct mvbt_pair {
  uint32_t first;
  uint32_t second;
} mvbt_pair;

typedef struct mvbt_memstats {
  uint64_t node_bytes;      /* bytes per node */
  uint64_t capacity_nodes;  /* pool slots */
  uint64_t live_nodes;      /* slots in use (tree nodes + unreclaimed version copies) */
  uint64_t peak_live_nodes; /* max live_nodes seen at reclaim/stat points */
  uint64_t free_nodes;      /* slots available to allocation */
  uint64_t retired_pending; /* retired copies not yet reclaimed */
  uint64_t reclaimed_nodes; /* total slots returned by reclamation */
  uint64_t retired_total;   /* total copies ever retired */
} mvbt_memstats;

/* One reclamation epoch of the memory timeline: the {usage, total allocated} pair the
 * reference records at each epoch advance under COLLECT_RECLAIMER_STATS
 * (memory_reclaimer.hpp:36-41,387-398), with the tree clock and the epoch's own figures. */
typedef struct mvbt_epoch_stat {
  uint64_t epoch;           /* reclamation number since the tree was created (0, 1, ...) */
  uint64_t clock;           /* tree clock (the next snapshot id) at the reclamation */
  uint64_t live_nodes;      /* slots in use after it (reference: usage) */
  uint64_t allocated_total; /* slots ever handed out: live + reclaimed (reference: total allocated) */
  uint64_t freed_nodes;     /* slots this reclamation returned */
  uint64_t retired_pending; /* retired copies still awaiting reclamation after it */
  uint64_t reclaimed_total; /* slots returned by every reclamation so far */
  uint64_t capacity_nodes;  /* pool slots */
} mvbt_epoch_stat;
#define MVBT_SERIES_CAPACITY 4096u /* epochs kept on the device (the latest) */

typedef struct mvbt_tree mvbt_tree;

/* Human-readable text of the last error on the calling thread. */
const char* mvbt_last_error(void);

/* Replaces the constructors gpu_versioned_btree() (gpu_versioned_blink_tree.hpp:88-93)
 * and gpu_blink_tree() (gpu_blink_tree.hpp, initialize_kernel btree_kernels.hpp:525-534).
 * Synchronous. */
mvbt_status mvbt_create(const mvbt_config* config, mvbt_tree** out);
/* Replaces the destructors (shared_ptr + cuda_deleter, host_allocators.hpp:19-22).
 * Synchronous (waits for the device). */
mvbt_status mvbt_destroy(mvbt_tree* tree);

/* Replaces gpu_versioned_btree::insert(keys, values, n, stream, in_place=true)
 * (gpu_versioned_blink_tree.hpp:106-142) and gpu_blink_tree::insert
 * (gpu_blink_tree.hpp:218-225). An existing key gets the new value. */
mvbt_status mvbt_insert(mvbt_tree* tree, const uint32_t* keys, const uint32_t* values,
                        uint64_t n, void* stream, uint32_t flags);

/* Replaces find(keys, values, n, [timestamp,] stream, concurrent)
 * (gpu_versioned_blink_tree.hpp:144-164, gpu_blink_tree.hpp:227-236).
 * values[i] = value of keys[i] visible at `snapshot` (MVBT_LATEST for the
 * latest version; plain trees accept only MVBT_LATEST), else MVBT_NOT_FOUND. */
mvbt_status mvbt_find(mvbt_tree* tree, const uint32_t* keys, uint32_t* values, uint64_t n,
                      uint32_t snapshot, void* stream, uint32_t flags);

/* Replaces erase(keys, n, stream, concurrent) (gpu_versioned_blink_tree.hpp:166-173,
 * gpu_blink_tree.hpp:238-245). On a versioned tree the erase is versioned:
 * snapshots taken before it still see the key (the reference per-batch erase is
 * not; its versioned erase exists only inside concurrent_find_erase). */
mvbt_status mvbt_erase(mvbt_tree* tree, const uint32_t* keys, uint64_t n, void* stream,
                       uint32_t flags);

/* Replaces range_query(lower, upper, result, counts, average_range_length, n,
 * [timestamp,] stream, concurrent) (gpu_versioned_blink_tree.hpp:175-216,
 * gpu_blink_tree.hpp:247-265). Query i covers [lower[i], upper[i]); its pairs,
 * ascending, go to out[i*L ...]; at most L are written (the reference writes
 * past the slot) and counts[i] receives the true count. `out` may be NULL to
 * count only; `counts` may be NULL. */
mvbt_status mvbt_range_query(mvbt_tree* tree, const uint32_t* lower, const uint32_t* upper,
                             mvbt_pair* out, uint32_t* counts, uint32_t L, uint64_t n,
                             uint32_t snapshot, void* stream, uint32_t flags);

/* Replaces take_snapshot(stream) (gpu_versioned_blink_tree.hpp:297-302): returns
 * the id s; writes issued after it (on any stream, through this API) are not
 * visible at s. The reference blocks until its snapshot kernel has run; here
 * the call returns at once with the id from the host's copy of the clock and
 * the snapshot kernel is stream-ordered: work on `stream` and writes on other
 * streams are ordered after it. Once a device view has been handed out (user
 * kernels may take snapshots), the call blocks and reads the id back. The
 * snapshot stays live (its versions are kept) until mvbt_release_snapshot.
 * Versioned trees only. */
mvbt_status mvbt_take_snapshot(mvbt_tree* tree, void* stream, uint32_t* snapshot_id);

/* No reference counterpart (the reference never frees in-place version copies):
 * marks snapshot s as no longer read. */
mvbt_status mvbt_release_snapshot(mvbt_tree* tree, uint32_t snapshot_id);

/* Epoch reclamation (reference DEBR, memory_reclaimer.hpp:244-401, driven here
 * by snapshot lifetimes): returns to the free list every version copy that no
 * live snapshot can read. Stream-ordered; must not overlap writers or readers
 * of older snapshots on other streams. */
mvbt_status mvbt_reclaim(mvbt_tree* tree, void* stream);

/* Memory accounting (reference compute_memory_usage gpu_versioned_blink_tree.hpp:1442-1446,
 * COLLECT_ALLOCATOR_STATS slab_alloc.hpp:489-540). Synchronous. */
mvbt_status mvbt_memory_stats(mvbt_tree* tree, mvbt_memstats* out);

/* The memory timeline (versioned trees): every reclamation - mvbt_reclaim, and the
 * reclamations a batched call makes between its windows - appends one mvbt_epoch_stat on
 * the device, in a ring of the latest MVBT_SERIES_CAPACITY. Copies the latest
 * min(recorded, capacity) of them, oldest first, to `out` (host memory) and their number
 * to *count. Synchronous. Reference: compute_epochs_stats (memory_reclaimer.hpp:114-133). */
mvbt_status mvbt_memory_series(mvbt_tree* tree, mvbt_epoch_stat* out, uint32_t capacity, uint32_t* count);

/* Sticky device error flags (MVBT_FLAG_*). Synchronous. */
mvbt_status mvbt_error_flags(mvbt_tree* tree, uint32_t* flags);
mvbt_status mvbt_clear_error_flags(mvbt_tree* tree);

/* Replaces validate_tree_structure (gpu_versioned_blink_tree.hpp:1098-1259,
 * gpu_blink_tree.hpp:743-900) for every allocator: copies the pool to the host
 * and checks that no node is locked, keys are sorted and below each high key,
 * every level's sibling chain is ordered, and leaves hold exactly the latest
 * key set. Writes the number of latest keys to *num_keys and the tree height
 * to *height (either may be NULL). Synchronous. Returns MVBT_ERR_DEVICE_FLAGS
 * with a message when an invariant fails. */
mvbt_status mvbt_validate(mvbt_tree* tree, uint64_t* num_keys, uint32_t* height);

/* Export of the latest (key, value) pairs in key order (reference
 * copy_tree_to_host, gpu_versioned_blink_tree.hpp:1092-1094). `out` holds
 * `capacity` pairs; *count receives the number of keys. Synchronous. */
mvbt_status mvbt_export(mvbt_tree* tree, mvbt_pair* out, uint64_t capacity, uint64_t* count);

/* Replaces the sorted-input bulk-build constructor gpu_blink_tree(keys, values,
 * num_keys, sorted_input = true, stream) (gpu_blink_tree.hpp:85-136,
 * bulk_build_kernel btree_kernels.hpp:915-1025), for both tree kinds. Loads n
 * (key, value) pairs into a new tree (no write, snapshot or device view yet)
 * with `fill` entries per node (0 = two thirds of the data lanes, rounded up;
 * else 2 .. data lanes). keys must be strictly ascending and exclude the
 * reserved keys; otherwise MVBT_ERR_INVALID_ARG and the tree is unchanged
 * (the reference terminates on unsorted input). The result equals inserting
 * the pairs. Synchronous. */
mvbt_status mvbt_bulk_build(mvbt_tree* tree, const uint32_t* keys, const uint32_t* values, uint64_t n,
                            uint32_t fill, void* stream);

/* Replaces concurrent_insert_range(keys, values, num_insertion, lower_bound,
 * upper_bound, num_ranges, result, average_range_length, stream)
 * (gpu_versioned_blink_tree.hpp:250-294, gpu_blink_tree.hpp:148-187): inserts
 * and range queries of one call run in one launch. Versioned trees: the range
 * queries read the snapshot taken at the start of the call (the clock advances
 * by one; the inserts are stamped after it), so their results are
 * deterministic (the reference snapshots inside the kernel). Plain trees: the
 * queries see the latest state (any linearization, as in the reference). Range
 * i writes at most L pairs to result[i*L ...] (result may be NULL) and its true
 * count to counts[i] (counts may be NULL). The launch runs alone among this
 * handle's writers: it waits for writes issued earlier on other streams, and
 * writes issued later on other streams wait for it. */
mvbt_status mvbt_concurrent_insert_range(mvbt_tree* tree, const uint32_t* keys, const uint32_t* values,
                                         uint64_t num_insertions, const uint32_t* lower, const uint32_t* upper,
                                         uint64_t num_ranges, mvbt_pair* result, uint32_t* counts, uint32_t L,
                                         void* stream);

/* Replaces concurrent_find_erase(find_keys, find_results, num_finds, erase_keys,
 * num_erasures, stream) (gpu_versioned_blink_tree.hpp:218-248,
 * gpu_blink_tree.hpp:189-224): finds and erases of one call run in one launch;
 * snapshot semantics as for mvbt_concurrent_insert_range (versioned erases are
 * versioned). */
mvbt_status mvbt_concurrent_find_erase(mvbt_tree* tree, const uint32_t* find_keys, uint32_t* find_results,
                                       uint64_t num_finds, const uint32_t* erase_keys, uint64_t num_erasures,
                                       void* stream);

/* Replaces plot_dot (gpu_versioned_blink_tree.hpp:1344-1377, gpu_blink_tree.hpp:986):
 * Graphviz text of the latest tree, breadth-first from the root (record nodes
 * with their pairs and high key; child, sibling and, for versioned leaves,
 * version-chain edges), at most max_nodes nodes (0 = all). Writes at most
 * capacity - 1 bytes plus a terminating NUL to buf (buf may be NULL) and the
 * full length to *length, like snprintf; the caller renders it (the reference
 * runs `dot` itself). Synchronous. */
mvbt_status mvbt_export_dot(mvbt_tree* tree, char* buf, uint64_t capacity, uint64_t* length, uint32_t max_nodes);

/* ---- Batched form: many batches of mixed operations in one launch ---------- */

#define MVBT_OP_FIND 0u
#define MVBT_OP_INSERT 1u
#define MVBT_OP_ERASE 2u
#define MVBT_OP_RANGE 3u

/* One operation. key = key or range lower bound; arg = insert value or range
 * upper bound (exclusive); slot = range output slot (pairs at pairs + slot*L). */
typedef struct mvbt_op {
  uint32_t type;
  uint32_t key;
  uint32_t arg;
  uint32_t slot;
} mvbt_op;

#define MVBT_USE_GRAPH 0x2u  /* launch through the handle's captured hipGraph */
#define MVBT_NO_RECLAIM 0x4u /* sorted form: keep every version copy until mvbt_reclaim */

/* No single reference counterpart: it generalises the reference's mixed kernels
 * concurrent_insert_range / concurrent_find_erase (gpu_versioned_blink_tree.hpp:218-294,
 * btree_kernels.hpp:31-306) to any mix, deterministically. Runs num_batches batches in
 * one launch sequence (a persistent work-queue worker in worker modes 0 and 1; the
 * sorted form in mode 2: every write grouped by leaf and applied in batch order, then
 * every read at its batch's snapshot; mode 3, the default, picks by call size, see
 * mvbt_set_worker_mode). batch_offsets (device, num_batches + 1 entries) are
 * absolute indices into ops: batch b is ops[offsets[b] .. offsets[b+1]), results go to
 * the same indices (so `results` must cover index offsets[num_batches] - 1), and
 * offsets[num_batches] - offsets[0] must equal num_ops (so a launch can cover any slice
 * of a larger pre-staged op array). Batch b reads (find, range) at snapshot base + b,
 * i.e. it sees exactly the writes of the batches before it; identical to calling
 * take_snapshot + find/range + insert/erase batch by batch. results[i]: find -> value
 * or MVBT_NOT_FOUND; range -> true count (at most L pairs written); insert -> 1 new
 * key, 0 value replaced; erase -> 1 erased, 0 absent; write of a reserved key ->
 * 0xFFFFFFFF (skipped, MVBT_FLAG_INVALID_KEY).
 * The writes of one batch must name distinct keys (their order inside the batch is
 * not defined); a call in which a batch writes one key twice, an op has an unknown
 * type, or the offsets disagree with num_ops is rejected on the device: nothing is
 * applied, results are left unchanged and MVBT_FLAG_BAD_BATCH is raised.
 * The sorted form runs a large call in windows of about mvbt_set_batch_window operations
 * (whole batches, at most 16 windows), each window's writes before its reads. When no
 * snapshot taken through this API is live and no device view was handed out, the call
 * reclaims between windows (the reference's DEBR frees inside its long mixed kernels,
 * memory_reclaimer.hpp:244-401): every version copy retired so far is freed once the
 * window's reads are done, so a call needs spare slots for about one window's copies, and
 * each such reclamation appends an epoch to the memory timeline (mvbt_memory_series). Like
 * mvbt_reclaim it must then not overlap readers of released snapshots on other streams;
 * MVBT_NO_RECLAIM keeps every copy until mvbt_reclaim instead.
 * The tree clock advances by num_batches (also for a rejected call); *snapshot_base
 * (device pointer, may be NULL) receives base. The launch runs alone among this
 * handle's writers: it waits for writes issued earlier on other streams, and writes
 * issued later on other streams (other run_batches calls included) wait for it. User
 * kernels that write through a device view (mvbt_device_view) must not overlap it
 * either (its reads and its sorted-form leaf writes assume no other writer runs).
 * Versioned trees only; scratch for num_batches and num_ops is allocated on first use
 * and when a call needs more (synchronous then). */
mvbt_status mvbt_run_batches(mvbt_tree* tree, const mvbt_op* ops, uint64_t num_ops,
                             const uint64_t* batch_offsets, uint32_t num_batches,
                             uint32_t* results, mvbt_pair* pairs, uint64_t pair_slots, uint32_t L,
                             uint32_t* snapshot_base, void* stream, uint32_t flags);

/* No counterpart in the reference (a capacity hint, like std::vector::reserve): sizes
 * the batched form's scratch for calls of up to num_ops operations and num_batches
 * batches and builds the launch graphs of the current worker mode (in mode 3, of both
 * forms a call of that size or smaller may run in), so that later such calls allocate
 * nothing and build nothing. Synchronous. */
mvbt_status mvbt_reserve_batches(mvbt_tree* tree, uint64_t num_ops, uint32_t num_batches);

/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * mvbt_run_batches in the sorted form (worker mode 2) with every scratch array in one
 * static device array of the library (about 37 MB on each device it runs on, held
 * until the process ends) instead of handle-owned allocations. Same arguments,
 * results and rejections as mvbt_run_batches (no flags: launched directly, no graph).
 * Limits: num_ops <= 524288 (2^19) and num_batches <= 16384 (2^14) per call, else
 * MVBT_ERR_INVALID_ARG and nothing is launched. Calls no hipMalloc/hipFree (the
 * first write on a stream creates that stream's ordering event, a host-side object).
 * Every call on one device shares that device's scratch: calls that overlap on one
 * device (two streams, or two handles on the same device) race and must be ordered
 * by the caller. */
mvbt_status mvbt_run_batches_noalloc(mvbt_tree* tree, const mvbt_op* ops, uint64_t num_ops,
                                     const uint64_t* batch_offsets, uint32_t num_batches,
                                     uint32_t* results, mvbt_pair* pairs, uint64_t pair_slots,
                                     uint32_t L, uint32_t* snapshot_base, void* stream);
#define MVBT_NOALLOC_MAX_OPS (1u << 19)
#define MVBT_NOALLOC_MAX_BATCHES (1u << 14)

/* Stream-ordered memory accounting: computes the mvbt_memstats of the point in
 * `stream` where it is enqueued and stores it to `out` (device or pinned host
 * memory) without synchronising the host. No reference counterpart. */
mvbt_status mvbt_memory_stats_async(mvbt_tree* tree, mvbt_memstats* out, void* stream);

/* Worker occupancy of the batched form in waves per CU (autotuner knob). */
mvbt_status mvbt_set_worker_waves(mvbt_tree* tree, uint32_t waves_per_cu);

/* Work unit of the batched form (autotuner knob): 0 = one operation per tile of a
 * persistent worker that opens batch after batch; 1 = the same worker with runs of
 * consecutive operations per tile, handled with the chunk routing and batched leaf
 * writes of the per-batch kernels; 2 = the sorted form: the call's writes are routed
 * to their leaves, sorted by leaf (stable radix sort), and applied leaf group by leaf
 * group in batch order by the tile that owns the group, then every find and range runs
 * at its batch's snapshot in one pass; 3 (default) = by call size: mode 0 for calls of
 * fewer than 16384 operations (the sorted form's launch sequence then costs more than
 * it saves), mode 2 from there up. Same results in every mode. */
mvbt_status mvbt_set_worker_mode(mvbt_tree* tree, uint32_t mode);

/* Read tile width (autotuner knob; the reference's tile is always its node width,
 * cg::tiled_partition<branching_factor> in btree_kernels.hpp): lanes per tile of the
 * launches in which no writer runs, i.e. mvbt_find and mvbt_range_query without
 * MVBT_CONCURRENT and the sorted form's read passes. 0 or the node width: one pair per
 * lane (default); 16 or 32 below the node width: each lane reads node_width / lanes
 * (2 or 4) consecutive pairs, and a wave walks 64 / lanes keys at once. Writers and
 * concurrent reads keep one pair per lane. Same results for every width. */
mvbt_status mvbt_set_read_tile(mvbt_tree* tree, uint32_t lanes);

/* Operations per window of the sorted form (autotuner knob; default 2^22; 0 = the whole
 * call in one window). Smaller windows free copies sooner and shorten the version chains
 * that reads walk; each window costs about 20 launches. No reference counterpart. */
mvbt_status mvbt_set_batch_window(mvbt_tree* tree, uint64_t ops_per_window);

/* ---- Device-side cooperative API ------------------------------------------ */

/* Device pointers of a tree for user kernels (include/mvbt_device.hpp turns it
 * into GpuBTree::device::*_tree with cooperative_find/insert/erase/range_query
 * and take_snapshot, the reference's device API, gpu_versioned_blink_tree.hpp:304-559).
 * After it has been called the handle treats every write call as possibly
 * following user-kernel writes. */
typedef struct mvbt_device_handle {
  void* nodes;
  void* state;
  void* free_stack;
  void* spill;
  void* retire_ring;
  uint64_t retire_capacity;
  uint32_t root;
  uint32_t capacity;
  uint32_t node_width;
  uint32_t kind;
} mvbt_device_handle;

mvbt_status mvbt_device_view(mvbt_tree* tree, mvbt_device_handle* out);

/* Tree parameters. */
mvbt_status mvbt_node_width(mvbt_tree* tree, uint32_t* node_width);
mvbt_status mvbt_kind_of(mvbt_tree* tree, uint32_t* kind);

#ifdef __cplusplus
}
#endif
#endif /* MVBT_H_ */
