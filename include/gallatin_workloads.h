/*
 * gallatin_workloads.h — benchmark and test workloads of the MI355X-native
 * Gallatin device allocator, a library of their own
 * (gallatin_amd/lib/libgallatin_workloads.so) on top of the allocator's
 * C-ABI (gallatin_amd.h, libgallatin_amd.so).
 *
 * A workload object is built on an allocator handle and holds it (gal_retain)
 * until it is destroyed: its kernels run on the handle's stream and call the
 * same device interface a user kernel calls (Gallatin<SEG,MIN,MAX>::malloc/
 * free, global_malloc/global_free). Batches, serving loops and slab tables
 * are built on a workload object and hold it in turn. Every entry point
 * takes plain pointers and sizes and returns an int status (GAL_OK == 0).
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   gal_run_*, gal_mixed_*, gal_churn
 *                       the reference's test kernels (tests/src/gallatin_test.cu:58-219,
 *                       tests/src/global_churn.cu:185-300) as fixed workloads
 *   gal_global_*        its global-allocator tests (tests/src/global_test.cu:34-110,
 *                       tests/src/poison_test.cu:60-128)
 *   gal_slab_*          data_structs/chaining_table.cuh:57-300 as driven by
 *                       tests/src/chaining_test.cu:80-155
 */
#ifndef GALLATIN_WORKLOADS_H
#define GALLATIN_WORKLOADS_H

#include <stdint.h>

#include "gallatin_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gal_workload gal_workload; /* opaque host handle */

/* A workload object on allocator handle h (device or host pool, any
 * supported triple); it holds h until gal_workload_destroy. Destroy returns
 * GAL_ERR_STATE while batches, serving loops or slab tables made on it are
 * alive. */
int gal_workload_create(gal_allocator* h, gal_workload** out);
int gal_workload_destroy(gal_workload* w);

/* ---- workloads ----
 * Kernel durations come from HIP events recorded around each launch on the
 * allocator handle's stream, the stream the kernels run on. */
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
int gal_run_plain(gal_workload* w, uint64_t n_threads, uint64_t size, uint64_t seed, uint32_t verify,
                  gal_run_result* out);

/* Mixed-size rounds. Round r: thread i mallocs 16 << (k mod 9) bytes with k
 * drawn from (seed, r, i), stores the pointer in the live table and tags the
 * allocation; then every live pointer j is freed (after its tag is checked)
 * when the coin drawn from (seed, r, j) is odd. gal_mixed_reset empties the
 * live table (allocations still live are freed). */
int gal_mixed_reset(gal_workload* w, uint64_t threads_per_round, uint32_t rounds, uint64_t seed);
int gal_mixed_round(gal_workload* w, uint32_t round, uint32_t verify, gal_run_result* out);
int gal_mixed_drain(gal_workload* w, uint32_t verify, gal_run_result* out);
/* live allocations of the table: byte offsets from the pool base, request
 * sizes, and live-table slots (round * threads_per_round + thread) */
int gal_mixed_live(gal_workload* w, uint64_t* offsets, uint64_t* sizes, uint64_t* slots, uint64_t cap,
                   uint64_t* count);
/* reset + all rounds + drain; one benchmark step */
int gal_run_mixed(gal_workload* w, uint64_t threads_per_round, uint32_t rounds, uint64_t seed, uint32_t verify,
                  gal_run_result* out);

/* Malloc latency samples of the last mixed run (or of the runs kept, see
 * gal_latency_keep): one per wavefront (its lane 0) and round, the wave's
 * time inside malloc in device wall-clock ticks; tick_ns is the tick
 * length. ticks == NULL just reports the count. */
int gal_latency_samples(gal_workload* w, uint32_t* ticks, uint64_t cap, uint64_t* count, double* tick_ns);
/* Keep the samples of the next `runs` gal_run_mixed calls one after another
 * (runs of one shape; later runs record none) instead of the last run's:
 * gal_latency_samples then returns them all, so percentiles cover several
 * runs. gal_mixed_round records no samples meanwhile. runs == 0 restores
 * the default. Clears the samples held. */
int gal_latency_keep(gal_workload* w, uint32_t runs);

/* ---- workloads of the global allocator ----
 * They run on the global allocators made by gal_init_global_allocator(_host)
 * (GAL_ERR_INVALID if absent), through global_malloc / global_free compiled
 * into this library, whose global pointers each call binds to them first. */
/* one kernel of n threads: global_malloc(size), tag, verify, global_free */
int gal_global_smoke(uint64_t n_threads, uint64_t size, uint64_t* misses, uint64_t* tag_errors);
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

/* The same kernel through the global allocator (gal_init_global_allocator),
 * timed on its handle's stream: n_threads threads each global_malloc(size),
 * write a tag drawn from (seed, thread), check it (GAL_VERIFY_*) and
 * global_free. The reference's global test kernel
 * (tests/src/global_test.cu:34-110) through global_malloc/global_free,
 * reference global_allocator.cuh:72-82. */
int gal_global_plain(uint64_t n_threads, uint64_t size, uint64_t seed, uint32_t verify, gal_run_result* out);

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
int gal_batch_create(gal_workload* w, uint64_t capacity, gal_batch** out);
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
 * slot and tid are ignored (the kernel uses its CU's slot hint and its own
 * thread ids). Each malloc tags its allocation's first word with a tag of
 * (seed, offset) by atomic exchange (expects 0) and each free checks and
 * clears it: out->tag_errors counts mismatches (a shared allocation). */
int gal_churn(gal_workload* w, const gal_op* ops, uint64_t n, uint64_t seed, uint64_t* results,
              gal_run_result* out);

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
int gal_slab_create(gal_workload* w, uint64_t buckets, uint64_t seed, gal_slab_table** out);
/* one kernel, one thread per key */
int gal_slab_insert(gal_slab_table* t, uint64_t n, uint64_t key_seed, gal_slab_result* out);
int gal_slab_query(gal_slab_table* t, uint64_t n, uint64_t key_seed, gal_slab_result* out);
/* frees every slab back to the allocator, then the table */
int gal_slab_destroy(gal_slab_table* t, uint64_t* slabs_freed);

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
int gal_serve_create(gal_workload* w, uint64_t capacity, gal_serve** out);
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

#endif /* GALLATIN_WORKLOADS_H */
