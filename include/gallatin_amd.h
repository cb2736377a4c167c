/*
 * gallatin_amd.h — C-ABI of the MI355X-native Gallatin device allocator.
 *
 * The in-kernel API (Gallatin<SEG,MIN,MAX>::malloc/free, global_malloc/
 * global_free) is C++ device code in gallatin_amd/include/gallatin/ and is
 * inlined into user kernels; it cannot cross a C ABI. This header is the host
 * boundary: construction/teardown, statistics, metadata snapshots and the
 * deterministic replays. Every entry point takes plain pointers and sizes and
 * returns an int status (GAL_OK == 0); none of them exits the process. The
 * benchmark and test workloads that drive the allocator are a separate
 * library on top of this one (gallatin_workloads.h).
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
 *   u32 ahead[total_slots], same order: the block fetched ahead for each slot
 *       (0 none, else block id + 1); always 0 in the oracle, which models no
 *       fetch ahead (one lane never fetches ahead)
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
 * GAL_OP_MALLOC claims use the reference's pinned-slot walk from `slot` for
 * every group; GAL_OP_MALLOC_SPREAD claims are the ones malloc makes under
 * load (when more than one lane claims, a group's first pass starts at slot
 * `slot` % usable of its class, pinned or overflow, and fills it if empty
 * unless the class is out of memory), without the blocks that malloc fetches
 * ahead for its slots under load. Either way the results are a function
 * of the op stream alone and the CPU oracle reproduces them bit for bit.
 *   GAL_OP_MALLOC, GAL_OP_MALLOC_SPREAD:
 *                  arg[l] = request size of lane l. The lanes' sizes are all
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
#define GAL_OP_MALLOC_SPREAD 2u
typedef struct gal_group_op {
  uint32_t kind;    /* GAL_OP_MALLOC, GAL_OP_MALLOC_SPREAD or GAL_OP_FREE */
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
/* Single-lane replay through global_malloc / global_free (reference
 * global_allocator.cuh:72-82): one lane of one kernel runs the ops in order
 * (gal_op; slot and tid are ignored: global_malloc uses the slot hint of the
 * lane's CU, its XCD in the low three bits, and its thread id 0). *cu_id
 * reports that hint, the slot the same stream replays with through
 * gal_replay or the oracle. */
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
/* The HIP stream (a hipStream_t) every call on the handle runs its device
 * work on. Work a caller orders with the handle's calls goes on it. */
int gal_stream(gal_allocator* h, void** stream);
/* Holds on a handle. An object built on a handle (a table, a workload of
 * another library) takes a hold while it lives and lets it go when it is
 * destroyed; gal_destroy and the global free calls return GAL_ERR_STATE
 * while any hold is taken. A handle is not safe for concurrent calls from
 * several host threads. No counterpart in the reference (its objects hold
 * raw device pointers). */
int gal_retain(gal_allocator* h);
int gal_release(gal_allocator* h);

/* Copy n bytes of the pool starting at byte offset off to host memory. */
int gal_read_pool(gal_allocator* h, uint64_t off, void* dst, uint64_t n);

/* Largest number of contiguous segments the segment tree's gather rule
 * (reference veb.cuh:323-452, used by malloc for requests above one block)
 * could claim right now: the fragmentation of the multi-segment path.
 * gal_gather_limit evaluates the rule on caller-supplied item words (host
 * only); gal_largest_run on the allocator's segment tree. No counterpart
 * in the reference (its print_info reports fill, gallatin.cuh:1855-1897). */
int gal_gather_limit(const uint64_t* words, uint64_t nwords, uint64_t* segments);
int gal_largest_run(gal_allocator* h, uint64_t* segments);

/* Bytes in free segments (segments in the segment tree) right now: the
 * allocator's own gauge, kept exact at every segment claim and return. No
 * counterpart in the reference (its print_info counts the tree's bits,
 * gallatin.cuh:1855-1897); the serving loop admits requests against it. */
int gal_free_bytes(gal_allocator* h, uint64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* GALLATIN_AMD_H */
