/*
 * gallatin_amd.h — C-ABI of the MI355X-native Gallatin device allocator.
 *
 * The in-kernel API (Gallatin<SEG,MIN,MAX>::malloc/free, global_malloc/
 * global_free) is C++ device code in gallatin_amd/include/gallatin/ and is
 * inlined into user kernels; it cannot cross a C ABI. This header is the host
 * boundary: construction/teardown, statistics, metadata snapshots, the
 * deterministic single-lane replay, and the benchmark workloads. Every entry
 * point takes plain pointers and sizes and returns an int status
 * (GAL_OK == 0); none of them exits the process.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   gal_create          Gallatin<SEG,MIN,MAX>::generate_on_device
 *                       include/gallatin/allocators/gallatin.cuh:324-421
 *   gal_destroy         Gallatin<SEG,MIN,MAX>::free_on_device  gallatin.cuh:643-668
 *   gal_print_stats     Gallatin::print_info                   gallatin.cuh:1855-1897
 *   gal_init_global_allocator / gal_free_global_allocator / gal_print_global_stats
 *                       include/gallatin/allocators/global_allocator.cuh:48-56, 59-70, 86-97
 *   gal_replay          a kernel in which one lane calls malloc/free in order
 *                       (what tests/src/gallatin_single_thread.cu:38-93 does),
 *                       with the CU id (smid) and thread id supplied explicitly
 *   gal_run_*           the reference's test kernels (tests/src/gallatin_test.cu:58-219,
 *                       tests/src/global_churn.cu:185-300) as fixed workloads
 *
 * Instantiations: templates cannot cross the ABI, so gal_create accepts only
 * the (segment_size, min_size, max_size) triples listed by gal_supported().
 */
#ifndef GALLATIN_AMD_H
#define GALLATIN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define GAL_OK 0
#define GAL_ERR_INVALID 1     /* bad argument (null handle, zero size, ...) */
#define GAL_ERR_UNSUPPORTED 2 /* (segment,min,max) triple not instantiated */
#define GAL_ERR_OOM 3         /* device allocation failed or pool too small to boot */
#define GAL_ERR_HIP 4         /* a HIP runtime call failed */
#define GAL_ERR_BUFFER 5      /* caller buffer too small; *needed says how big */
#define GAL_ERR_STATE 6       /* allocator detected an internal inconsistency, or the handle is
                                 still used by a batch, serving loop or slab table */
#define GAL_ERR_TOO_LARGE 7   /* request above a fixed limit of the entry point */

typedef struct gal_allocator gal_allocator; /* opaque host handle */

/* ---- replay stream ----
 * One op per entry, executed in order by a single lane.
 *   GAL_OP_MALLOC: arg = request size in bytes; result = byte offset of the
 *                  returned pointer from the pool base, or GAL_NULL_OFFSET.
 *   GAL_OP_FREE:   arg = index (into the same stream) of the malloc op whose
 *                  pointer is freed; that malloc must have succeeded and not
 *                  been freed already. result = 0.
 * slot: stands in for the CU id (the reference's %smid): the pinned-block
 *       search starts at slot % slots_of_class.
 * tid:  the thread id fed to the hashed segment search
 *       (reference veb.cuh:946-956 hashes threadIdx.x + blockIdx.x*blockDim.x).
 */
#define GAL_OP_MALLOC 0u
#define GAL_OP_FREE 1u
#define GAL_NULL_OFFSET (~(uint64_t)0)

typedef struct gal_op {
  uint32_t kind;
  uint32_t slot;
  uint64_t arg;
  uint64_t tid;
} gal_op;

/* ---- allocator description ---- */
typedef struct gal_info_t {
  uint64_t segment_size;       /* bytes per segment */
  uint64_t min_size;           /* smallest slice size */
  uint64_t max_size;           /* largest slice size */
  uint64_t num_segments;       /* pool_bytes / segment_size */
  uint64_t blocks_per_segment; /* segment_size / (min_size * 4096) */
  uint64_t num_trees;          /* log2(max) - log2(min) + 1 size classes */
  uint64_t veb_layers;         /* layers of each vEB tree over num_segments */
  uint64_t total_slots;        /* pinned-block slots over all classes */
  uint64_t pool_device_ptr;    /* device address of the pool base */
  uint64_t device_handle;      /* device address of the Gallatin object */
  uint64_t snapshot_bytes;     /* size of a gal_snapshot image */
  uint64_t metadata_bytes;     /* device bytes held by metadata (not pool) */
} gal_info_t;

/*
 * Snapshot image (identical layout from the GPU and from the CPU oracle):
 *   u64 header[8] = {num_segments, blocks_per_segment, num_trees, veb_layers,
 *                    total_slots, locks, 0, 0}
 *   segment tree words, layer 0 (top summary) .. layer L-1 (one bit per segment)
 *   for each class 0..T-1: that class's sub-tree words, same layer order
 *   u16 chunk_ids[num_segments], zero-padded to a multiple of 8 bytes
 *       (0xFFFF free; c < T owned by class c; T+1+n head of an n-segment run)
 *   u32 {malloc_counter, free_counter} per block, num_segments*blocks_per_segment
 *   u32 queue[num_segments*blocks_per_segment]  (0 empty, else block id + 1)
 *   u32 {active, queue_alloc, queue_free, queue_final} per segment (active as i32)
 *   u32 slots[total_slots], class-major (0 empty, else block id + 1)
 *   zero-padded to a multiple of 8 bytes
 */

/* (segment_size, min_size, max_size) triples compiled into the library.
 * Writes up to cap triples (3 u64 each) into out; returns the count. */
int gal_supported(uint64_t* out, int cap);

int gal_create(uint64_t segment_size, uint64_t min_size, uint64_t max_size,
               uint64_t pool_bytes, uint64_t seed, gal_allocator** out);
/* Bytes malloc grants a request of `size` on a (segment, min, max) allocator:
 * a slice, 2^k top-class slices, a whole block or whole segments (host
 * only, no GPU). Replaces Gallatin::get_allocated_size, reference
 * gallatin.cuh:1152-1183, which rounds every request below a segment up to
 * a power of two. */
int gal_allocated_size(uint64_t segment_size, uint64_t min_size, uint64_t max_size, uint64_t size, uint64_t* bytes);

/* The same over a pool of pinned host memory mapped into the GPU (metadata
 * stays in device memory). Replaces generate_on_device_host, reference
 * gallatin.cuh:427-630 (memory_table.cuh:340-366, host_only). */
int gal_create_host(uint64_t segment_size, uint64_t min_size, uint64_t max_size,
                    uint64_t pool_bytes, uint64_t seed, gal_allocator** out);
int gal_destroy(gal_allocator* h);
int gal_get_info(gal_allocator* h, gal_info_t* info);
int gal_print_stats(gal_allocator* h);

/* The figures print_info prints, as numbers (reference gallatin.cuh:1855-2011
 * print_info / print_guided_fill / calculate_overhead, veb.cuh report_fill
 * and report_max). Per class c < num_trees: segments its tree lists,
 * segments formatted for it, and over those, blocks still grantable (the
 * segments' active counts, "available") and blocks granted and not back ("in
 * use"). */
typedef struct gal_fill_report {
  uint64_t num_trees;
  uint64_t free_segments;  /* items of the segment tree (report_fill) */
  uint64_t max_segments;   /* its universe (report_max) */
  uint64_t listed_segments[32];
  uint64_t live_segments[32];
  uint64_t blocks_available[32];
  uint64_t blocks_in_use[32];
  uint64_t overhead_bytes; /* device memory held besides the pool */
  /* per class: pinned slots the allocator counts as holding no block (its
   * fast out-of-memory test; equal to the slots that hold 0 whenever no call
   * is in flight) */
  uint64_t empty_pinned_slots[32];
} gal_fill_report;
int gal_fill(gal_allocator* h, gal_fill_report* out);
int gal_snapshot(gal_allocator* h, void* buf, uint64_t cap, uint64_t* needed);

/* Single-lane deterministic replay; results[i] per op (see above). */
int gal_replay(gal_allocator* h, const gal_op* ops, uint64_t n, uint64_t* results);

/* ---- multi-lane deterministic replay ----
 * One wavefront runs the ops in order, nothing else runs on the allocator.
 * Op i is one call site of malloc (or free) reached by the lanes of `lanes`
 * together, with exactly that exec mask, the way the lanes of a warp that
 * are active at a call in a reference kernel enter Gallatin::malloc
 * together (gallatin.cuh:1186-1276, teams by class at :881-885, one claim
 * per team at block.cuh:245-281; frees grouped by block at :1774-1827).
 * The claims use the reference's pinned-slot walk from `slot` for every
 * group (the overflow-slot spreading that concurrent claims use is off), so
 * the results are a function of the op stream alone and the CPU oracle
 * reproduces them bit for bit.
 *   GAL_OP_MALLOC: arg[l] = request size of lane l. The lanes' sizes are all
 *                  on the slice path (up to min_size * 2048 bytes) or all
 *                  above it (the block and segment requests are served one
 *                  lane at a time in lane order).
 *   GAL_OP_FREE:   arg[l] = index j * 64 + k of the result freed (lane k of an
 *                  earlier malloc op j, which must have succeeded and not
 *                  been freed). A free op of several lanes frees slices or
 *                  whole blocks only; a run of segments is freed alone.
 * results[i * 64 + l]: lane l's byte offset from the pool base (malloc),
 * 0 (free), GAL_NULL_OFFSET (a failed malloc, or a lane not in the op).
 * No counterpart in the reference (whose concurrent results are not
 * reproducible); it pins the group claim semantics against the oracle. */
typedef struct gal_group_op {
  uint32_t kind;    /* GAL_OP_MALLOC or GAL_OP_FREE */
  uint32_t slot;    /* CU id stand-in of the wavefront (the reference's %smid) */
  uint64_t lanes;   /* lanes taking part, nonzero */
  uint64_t arg[64]; /* per lane, see above */
  uint64_t tid[64]; /* per lane: thread id fed to the hashed segment search */
} gal_group_op;
int gal_group_replay(gal_allocator* h, const gal_group_op* ops, uint64_t n, uint64_t* results);

/* gal_group_replay for callers that cannot allocate device memory: code
 * inside a captured hipGraph, tight serving loops, harnesses that forbid
 * hipMalloc. No counterpart in the reference. ops (n entries) and results
 * (n * 64 entries, every one written up to the first rejected op) are device
 * memory that the caller owns; host pointers are refused (GAL_ERR_INVALID).
 * The rules of gal_group_replay are checked on the device, with a bitmap of
 * freed results held in a fixed array of the library in device memory: one
 * copy per GPU, kept until the process ends, room for GAL_NOALLOC_MAX_OPS
 * ops (more: GAL_ERR_TOO_LARGE, nothing runs). Calls on one GPU share that
 * array, so calls that overlap in time on one GPU race. The call clears the
 * part it uses with a kernel, runs on the handle's stream and waits for it;
 * it calls no hipMalloc, hipFree or other allocating routine. Returns
 * GAL_ERR_INVALID if an op breaks the rules: the ops before it have run, it
 * and the later ones have not. */
#define GAL_NOALLOC_MAX_OPS 8192u
int gal_group_replay_noalloc(gal_allocator* h, const gal_group_op* d_ops, uint64_t n, uint64_t* d_results);

/* Device error counters: internal inconsistencies (where the reference would
 * trap) and rejected frees of pointers the allocator does not own. */
int gal_errors(gal_allocator* h, uint64_t* internal_errors, uint64_t* bad_frees);
/* Per-site counts behind internal_errors, in this order: recycle-queue slot
 * empty at a grant; block free count not 4096 at its reset; class id of a
 * released segment changed; a free segment could not be formatted; queue
 * write waited too long for its slot; queue write waited too long for earlier
 * writes; whole-block grant of another class; segment-run id changed under its
 * free; whole-block grant found slices already claimed; impossible counter at
 * a grant; grant from a segment of another class; slice claim from a segment
 * of another class; free of a block carrying another class; more block
 * returns than grants. Writes min(cap, 14) counts. */
int gal_error_sites(gal_allocator* h, uint32_t* counts, uint32_t cap);
/* Cumulative activity of the allocator's rare paths, in this order: spent
 * pinned blocks swapped out; empty pinned slots refilled; block requests;
 * block-request passes that lost a race and retried; segments attached to a
 * class; segments released to the segment tree; blocks recycled after all
 * their slices came back; device wall-clock ticks / 16 during which pinned
 * slots were held busy for replacement and for refill (summed over the
 * slots). Writes min(cap, 9) counts. */
int gal_stats(gal_allocator* h, uint32_t* counts, uint32_t cap);

/* Global allocator (Gallatin<16 MiB, 16, 4096> singleton).
 * Replaces init_global_allocator / free_global_allocator / print_global_stats,
 * reference global_allocator.cuh:48-56, 59-70, 86-97. */
int gal_init_global_allocator(uint64_t num_bytes, uint64_t seed, int print_info);
int gal_free_global_allocator(void);
int gal_print_global_stats(void);
/* one kernel of n threads: global_malloc(size), tag, verify, global_free */
int gal_global_smoke(uint64_t n_threads, uint64_t size, uint64_t* misses, uint64_t* tag_errors);
/* Single-lane replay through global_malloc / global_free (reference
 * global_allocator.cuh:72-82): one lane of one kernel runs the ops in order
 * (gal_op; slot and tid are ignored: global_malloc uses the lane's hardware
 * CU id and its thread id 0). *cu_id reports the CU id the lane ran on, the
 * slot the same stream replays with through gal_replay or the oracle. */
int gal_global_replay(const gal_op* ops, uint64_t n, uint64_t* results, uint32_t* cu_id);

/* Global allocator over mapped host memory, and both together. Replace
 * init_global_allocator_host / free_global_allocator_host /
 * print_global_stats_host and the _combined forms, reference
 * global_allocator.cuh:98-236. */
int gal_init_global_allocator_host(uint64_t num_bytes, uint64_t seed, int print_info);
int gal_free_global_allocator_host(void);
int gal_print_global_stats_host(void);
int gal_init_global_allocator_combined(uint64_t num_bytes, uint64_t host_bytes, uint64_t seed, int print_info);
int gal_free_global_allocator_combined(void);
int gal_print_global_stats_combined(void);
/* the handles of the global allocators (owned by the library: not to be
 * destroyed), for snapshots and counters; NULL when not initialised */
int gal_global_handles(gal_allocator** device, gal_allocator** host);
/* n threads each global_malloc_fused(size) (device pool first, then the
 * host pool), tag; a second kernel checks and global_free_fused. Counts
 * where the allocations landed. Reference global_allocator.cuh:199-216. */
int gal_global_fused(uint64_t n_threads, uint64_t size, uint64_t* on_device, uint64_t* on_host, uint64_t* misses,
                     uint64_t* tag_errors);
/* The reference's poison sequence (tests/src/poison_test.cu:60-128) for one
 * size on the device global allocator, one lane: 7 allocations by
 * global_malloc_poison, each written and freed by global_free_poison after
 * (1) nothing, (2) the header's size word, (3) its extra word, (4) the byte
 * after the caller's bytes, (5) both header words = 17, (6) both = 48 were
 * overwritten, (7) the caller's own last byte was written. codes[k] = the
 * guard check's result (0 intact; 1-5 see global_check_poison; -1 malloc
 * failed). */
int gal_global_poison(uint64_t size, int32_t* codes, uint32_t cap);

/* ---- workloads ----
 * Kernel durations come from HIP events recorded on the library's own stream
 * around each launch (the stream the kernels run on). */
typedef struct gal_run_result {
  uint64_t mallocs;        /* successful mallocs */
  uint64_t frees;          /* frees issued */
  uint64_t misses;         /* mallocs that returned nullptr */
  uint64_t tag_errors;     /* payload tag checks that failed (overlap) */
  uint64_t alloc_errors;   /* growth of the allocator's internal error count */
  uint64_t launches;       /* kernels launched */
  double malloc_kernel_ms; /* summed durations of the kernels that malloc */
  double free_kernel_ms;   /* summed durations of the kernels that only free */
  double total_ms;         /* first launch start to last launch end */
} gal_run_result;

/* Tag check modes for the workloads. */
#define GAL_VERIFY_STORE 0u  /* store the tag, load it back */
#define GAL_VERIFY_ATOMIC 1u /* atomic exchange in (expects 0) and out (expects tag): detects sharing */

/* One kernel of n_threads threads. Each thread: p = malloc(size), write a
 * tag derived from (seed, thread), check it, free(p). The device API's
 * in-kernel interface of the reference test kernels
 * (tests/src/gallatin_test.cu:58-110, tests/src/global_churn.cu:185-245). */
int gal_run_plain(gal_allocator* h, uint64_t n_threads, uint64_t size, uint64_t seed, uint32_t verify,
                  gal_run_result* out);
/* The same kernel through the global allocator (gal_init_global_allocator),
 * timed on its handle's stream: n_threads threads each global_malloc(size),
 * write a tag drawn from (seed, thread), check it (GAL_VERIFY_*) and
 * global_free. The reference's global test kernel
 * (tests/src/global_test.cu:34-110) through global_malloc/global_free,
 * reference global_allocator.cuh:72-82. */
int gal_global_plain(uint64_t n_threads, uint64_t size, uint64_t seed, uint32_t verify, gal_run_result* out);

/* Mixed-size rounds. Round r: thread i mallocs 16 << (k mod 9) bytes with k
 * drawn from (seed, r, i), stores the pointer in the live table and tags the
 * allocation; then every live pointer j is freed (after its tag is checked)
 * when the coin drawn from (seed, r, j) is odd. gal_mixed_reset empties the
 * live table (allocations still live are freed). */
int gal_mixed_reset(gal_allocator* h, uint64_t threads_per_round, uint32_t rounds, uint64_t seed);
int gal_mixed_round(gal_allocator* h, uint32_t round, uint32_t verify, gal_run_result* out);
int gal_mixed_drain(gal_allocator* h, uint32_t verify, gal_run_result* out);
/* live allocations of the table: byte offsets from the pool base, request
 * sizes, and live-table slots (round * threads_per_round + thread) */
int gal_mixed_live(gal_allocator* h, uint64_t* offsets, uint64_t* sizes, uint64_t* slots, uint64_t cap,
                   uint64_t* count);
/* reset + all rounds + drain; one benchmark step */
int gal_run_mixed(gal_allocator* h, uint64_t threads_per_round, uint32_t rounds, uint64_t seed, uint32_t verify,
                  gal_run_result* out);

/* Malloc latency samples of the last mixed run: one per wavefront (its lane
 * 0), the wave's time inside malloc in device wall-clock ticks; tick_ns is
 * the tick length. ticks == NULL just reports the count. */
int gal_latency_samples(gal_allocator* h, uint32_t* ticks, uint64_t cap, uint64_t* count, double* tick_ns);

/* Copy n bytes of the pool starting at byte offset off to host memory. */
int gal_read_pool(gal_allocator* h, uint64_t off, void* dst, uint64_t n);

/* ---- batches of uneven requests ----
 * A batch holds up to `capacity` request sizes. Serving it bins the requests
 * by the class they claim from (a stable counting sort on the device), then
 * one persistent launch serves the binned order, `run` consecutive requests
 * per lane (1..64): a lane whose requests share a slice class claims them
 * with one bulk claim (Gallatin::malloc_run_at: consecutive slices of one
 * block, the lane's share of one group add) and frees them with one
 * free_run. Lanes of one class form claim groups of `width` lanes (1..64, a
 * power of two). No counterpart in the reference, whose callers malloc one
 * request per thread; it serves the same requests from the same blocks and
 * counters (Gallatin::malloc/free, gallatin.cuh:1186-1461). */
typedef struct gal_batch gal_batch;
#define GAL_FORM_PER_THREAD 0u /* one thread per request, request order: malloc, tag, check, free */
#define GAL_FORM_BATCHED 1u    /* bin + persistent malloc, tag, check, free; launched directly */
#define GAL_FORM_GRAPH 2u      /* the same launches captured once into a hipGraph and replayed */
int gal_batch_create(gal_allocator* h, uint64_t capacity, gal_batch** out);
int gal_batch_destroy(gal_batch* b);
/* n sizes 16 << k, k uniform in [0, size_classes), drawn on the device from (seed, request) */
int gal_batch_generate(gal_batch* b, uint64_t n, uint64_t seed, uint32_t size_classes);
/* n sizes copied from host memory; seed feeds the payload tags */
int gal_batch_set_sizes(gal_batch* b, const uint64_t* sizes, uint64_t n, uint64_t seed);
/* serve the batch `replays` times in the given form (each request malloc'd,
 * tagged, checked and freed); out->total_ms covers all replays. width and
 * run are ignored by GAL_FORM_PER_THREAD. */
int gal_batch_run(gal_batch* b, uint32_t form, uint32_t width, uint32_t run, uint32_t replays, uint32_t verify,
                  gal_run_result* out);
/* batched malloc of every request (results kept in the batch), and the
 * batched free of those results, one by one, in the same binned order */
int gal_batch_malloc(gal_batch* b, uint32_t width, uint32_t run, gal_run_result* out);
int gal_batch_free(gal_batch* b, gal_run_result* out);
/* results of gal_batch_malloc as byte offsets from the pool base (GAL_NULL_OFFSET: miss) */
int gal_batch_offsets(gal_batch* b, uint64_t* offsets, uint64_t cap);
/* request indices in the binned order of the last binning */
int gal_batch_order(gal_batch* b, uint32_t* order, uint64_t cap);

/* ---- concurrent churn ----
 * One kernel, one thread per op, all ops of the call running concurrently
 * in any interleaving (the reference's test kernels mix mallocs and frees
 * the same way, tests/src/gallatin_test.cu:58-219):
 *   GAL_OP_MALLOC: arg = size; results[i] = byte offset from the pool base,
 *                  or GAL_NULL_OFFSET when malloc returned nullptr
 *   GAL_OP_FREE:   arg = byte offset of a live allocation (made by an
 *                  earlier call); results[i] = 0
 * slot and tid are ignored (the kernel uses the hardware CU id and its own
 * thread ids). Each malloc tags its allocation's first word with a tag of
 * (seed, offset) by atomic exchange (expects 0) and each free checks and
 * clears it: out->tag_errors counts mismatches (a shared allocation). */
int gal_churn(gal_allocator* h, const gal_op* ops, uint64_t n, uint64_t seed, uint64_t* results,
              gal_run_result* out);

/* Largest number of contiguous segments the segment tree's gather rule
 * (reference veb.cuh:323-452, used by malloc for requests above one block)
 * could claim right now: the fragmentation of the multi-segment path.
 * gal_gather_limit evaluates the rule on caller-supplied item words (host
 * only); gal_largest_run on the allocator's segment tree. No counterpart
 * in the reference (its print_info reports fill, gallatin.cuh:1855-1897). */
int gal_gather_limit(const uint64_t* words, uint64_t nwords, uint64_t* segments);
int gal_largest_run(gal_allocator* h, uint64_t* segments);

/* ---- slab hash client ----
 * A chained hash table of 64-bit keys and values, 7 per 120-byte slab, whose
 * slabs are malloc'd from the handle's allocator inside the insert kernel:
 * the allocator's client workload. Replaces the reference's
 * chaining_table<Key, Val, 7> (include/gallatin/data_structs/
 * chaining_table.cuh:57-300) as driven by tests/src/chaining_test.cu:80-155;
 * the device structure is gallatin_amd/include/gallatin/data_structs/
 * chaining_table.hpp. Key i of a stream is a bijection of (key_seed + i)
 * (never 0, the empty mark); its value is i. */
typedef struct gal_slab_table gal_slab_table;
typedef struct gal_slab_result {
  uint64_t keys;         /* keys inserted or queried */
  uint64_t failed;       /* inserts that found no memory; queries that missed */
  uint64_t wrong;        /* queries that found another value */
  uint64_t slabs;        /* slabs in the table after an insert (or freed by destroy) */
  uint64_t max_chain;    /* longest chain after an insert, slabs */
  uint64_t alloc_errors; /* growth of the allocator's internal error count */
  double kernel_ms;      /* the insert / query kernel (HIP events, handle's stream) */
} gal_slab_result;
int gal_slab_create(gal_allocator* h, uint64_t buckets, uint64_t seed, gal_slab_table** out);
/* one kernel, one thread per key */
int gal_slab_insert(gal_slab_table* t, uint64_t n, uint64_t key_seed, gal_slab_result* out);
int gal_slab_query(gal_slab_table* t, uint64_t n, uint64_t key_seed, gal_slab_result* out);
/* frees every slab back to the allocator, then the table */
int gal_slab_destroy(gal_slab_table* t, uint64_t* slabs_freed);

/* Bytes in free segments (segments in the segment tree) right now: the
 * allocator's own gauge, kept exact at every segment claim and return. No
 * counterpart in the reference (its print_info counts the tree's bits,
 * gallatin.cuh:1855-1897); the serving loop admits requests against it. */
int gal_free_bytes(gal_allocator* h, uint64_t* bytes);

/* ---- serving loop ----
 * Sequences arrive over time and grow by one page per step, then free all
 * their pages. One persistent launch serves every request (continuous
 * batching): each lane of the grid is a sequence slot; an idle slot takes
 * the next request, admits it once it has arrived and the pool has room for
 * all of its pages (free segment bytes >= pages not yet allocated by
 * admitted sequences + the new request's bytes + a headroom for partly used
 * blocks), grows it page by page (malloc, write the page header, read the
 * previous page back), then drains it page by page (read back, free) and
 * reports a checksum of what it read. The checksum of a request depends
 * only on (id, pages, page_bytes), alone or in any mix. No counterpart in the
 * reference; every page is an ordinary Gallatin::malloc/free
 * (gallatin.cuh:1186-1461) made from inside the loop. */
typedef struct gal_request {
  uint64_t arrival_ns; /* arrival after the start of the run */
  uint64_t id;         /* seeds the page contents */
  uint32_t pages;      /* pages the sequence grows to, >= 1 */
  uint32_t page_bytes; /* bytes per page, >= 16 */
} gal_request;

typedef struct gal_serve_workload {
  uint64_t requests;
  double rate_per_s;      /* Poisson arrivals; 0: every request arrives at time 0 */
  double median_pages;    /* lognormal length: median ... */
  double sigma;           /* ... and sigma of log(length) */
  uint32_t max_pages;     /* lengths are capped here */
  uint32_t page_bytes;    /* 0: per-request 16 << k bytes, k uniform in [0, size_classes) */
  uint32_t size_classes;
  uint32_t unused;
  uint64_t seed;
} gal_serve_workload;

/* Fills out[0 .. w->requests) with the workload's requests (host only, no
 * GPU): request i has id i, a length min(max_pages, ceil(median *
 * exp(sigma * z))) with z standard normal, and arrives an exponential gap of
 * mean 1/rate after request i-1. Every draw is a function of (seed, i). */
int gal_serve_generate(const gal_serve_workload* w, gal_request* out);

typedef struct gal_serve gal_serve;
typedef struct gal_serve_result {
  uint64_t requests_done;      /* requests that finished */
  uint64_t requests_aborted;   /* requests cut by the deadline (checksum 0) */
  uint64_t page_mallocs;       /* pages allocated */
  uint64_t page_frees;         /* pages freed */
  uint64_t page_retries;       /* page mallocs that returned nullptr (retried next step) */
  uint64_t admission_refusals; /* admission attempts refused for lack of room */
  uint64_t peak_active;        /* most sequences in flight at once */
  uint64_t peak_committed;     /* most admitted-but-unallocated page bytes */
  uint64_t alloc_errors;       /* growth of the allocator's internal error count */
  uint64_t slots;              /* sequence slots launched */
  uint64_t run;                /* pages claimed per allocator call (at most) */
  uint64_t headroom;           /* headroom bytes used by admission */
  double kernel_ms;            /* the serving launch (HIP events on the handle's stream) */
  double makespan_ms;          /* start of the loop to the last completion (device clock) */
  double page_malloc_p50_ns;   /* per-page malloc latency: the wave's time inside the allocator */
  double page_malloc_p99_ns;   /*   call that claimed the page, one sample per page (1/16 octave) */
  double page_malloc_max_ns;
  double queue_p50_ns;         /* arrival -> admission */
  double queue_p99_ns;
  double latency_p50_ns;       /* arrival -> last page freed */
  double latency_p99_ns;
  uint64_t window_pages;       /* pages written plus pages read back and freed before window_s */
  double window_ms;            /* the measuring window (0: none asked) */
  double first_dry_ms;         /* when a slot first found every request queue empty (-1: never):
                                  a window that ends before it measures the loop saturated */
} gal_serve_result;

#define GAL_SERVE_DEFAULT_HEADROOM (~(uint64_t)0)
int gal_serve_create(gal_allocator* h, uint64_t capacity, gal_serve** out);
int gal_serve_destroy(gal_serve* s);
int gal_serve_load(gal_serve* s, const gal_request* reqs, uint64_t n);
/* Serve the loaded requests in one persistent launch of `slots` sequence
 * slots (rounded up to 256, capped at what the GPU holds at once). A growing
 * sequence claims its next min(run, pages left) pages (run in 1..64) with one
 * bulk claim (consecutive slices of one block, each page still a separate
 * allocation) and frees them with one free_run once they are all read back;
 * run 1 is one malloc and one free per page. headroom
 * GAL_SERVE_DEFAULT_HEADROOM: for each class the page sizes claim from, its
 * pinned and overflow slots' blocks and two segments; plus a tenth of the
 * pool when such a class has several blocks per segment (recycled blocks of
 * a segment stay hidden until a tenth of them are back, reference
 * gallatin.cuh:1646). Requests not finished deadline_s
 * seconds after the start are aborted (their pages are freed). window_s > 0
 * also counts the pages written and freed in the first window_s seconds
 * (window_pages), the sustained rate while the queues are not empty. Returns
 * GAL_ERR_INVALID if a request could never be admitted. */
int gal_serve_run(gal_serve* s, uint32_t slots, uint32_t run, uint64_t headroom, double deadline_s, double window_s,
                  gal_serve_result* out);
/* per request of the last run, indexed as the requests passed to
 * gal_serve_load: checksum (0 if aborted), admission and completion times in
 * ns after the start (~0 if aborted) */
int gal_serve_results(gal_serve* s, uint64_t* checksums, uint64_t* admit_ns, uint64_t* done_ns, uint64_t cap);

/* Agent-scope 32-bit atomic add rate of this GPU, in returned adds per
 * second. GAL_ATOMIC_DISTINCT: `threads` lanes each add `iters` times to
 * their own word. GAL_ATOMIC_WAVE_HOT: the first lane of each wave of
 * `threads` adds `iters` times to one of `addresses` words, the pattern of a
 * wave-aggregated claim on a pinned block's counter. No counterpart in the
 * reference: it measures the rate that bounds the slice path, for the
 * roofline. Allocates and frees its own buffer. */
#define GAL_ATOMIC_DISTINCT 0u
#define GAL_ATOMIC_WAVE_HOT 1u
int gal_atomic_rate(uint32_t mode, uint64_t threads, uint32_t iters, uint32_t addresses, double* ops_per_s);

#ifdef __cplusplus
}
#endif

#endif /* GALLATIN_AMD_H */
