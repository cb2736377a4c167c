/*
 * gallatin_oracle.h — C-ABI of the CPU oracle.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE. The oracle is a sequential host
 * restatement of the reference allocator's algorithm. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as
 * the checker (or as the timed CPU baseline). The product path never calls it.
 */
#ifndef GALLATIN_ORACLE_H
#define GALLATIN_ORACLE_H

#include <stdint.h>

#include "../include/gallatin_amd.h" /* gal_op, gal_info_t, snapshot layout */

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_alloc orc_alloc;
typedef struct orc_veb orc_veb;

/* ---- primitives ---- */
uint64_t orc_murmur64a(uint64_t key, uint64_t seed);
int orc_first_bit_bigger(uint64_t x);
int orc_cfcll(uint64_t bits);
uint64_t orc_splitmix64(uint64_t x);
/* the re-register test return_block applies to a returned block's queue slot */
int orc_reregister(uint32_t slot, uint64_t nb);

/* Size routing: path 0 = slice, 1 = block, 2 = segments, 3 = unsupported.
 * tree = class used, count = slices (slice path) or segments (segment path). */
int orc_route(uint64_t seg, uint64_t min, uint64_t max, uint64_t size,
              uint32_t* path, uint32_t* tree, uint64_t* count);

/* ---- standalone vEB tree (brute-force tests) ---- */
orc_veb* orc_veb_create(uint64_t universe, uint64_t seed, int start_full);
void orc_veb_destroy(orc_veb* t);
int orc_veb_layers(orc_veb* t);
uint64_t orc_veb_layer_words(orc_veb* t, int layer);
void orc_veb_get_words(orc_veb* t, int layer, uint64_t* out);
void orc_veb_set_words(orc_veb* t, int layer, const uint64_t* in);
int orc_veb_insert(orc_veb* t, uint64_t v); /* insert_force_update */
int orc_veb_remove(orc_veb* t, uint64_t v);
int orc_veb_query(orc_veb* t, uint64_t v);
uint64_t orc_veb_successor(orc_veb* t, uint64_t v); /* successor_thorough */
uint64_t orc_veb_lock_offset(orc_veb* t, uint64_t start);
uint64_t orc_veb_malloc_first(orc_veb* t);
uint64_t orc_veb_find_random(orc_veb* t, uint64_t tid);
uint64_t orc_veb_gather_multiple(orc_veb* t, uint64_t n);
void orc_veb_return_multiple(orc_veb* t, uint64_t start, uint64_t n);

/* ---- whole allocator ---- */
int orc_create(uint64_t segment_size, uint64_t min_size, uint64_t max_size,
               uint64_t pool_bytes, uint64_t seed, orc_alloc** out);
void orc_destroy(orc_alloc* a);
int orc_get_info(orc_alloc* a, gal_info_t* info);
/* byte offset from the pool base, or GAL_NULL_OFFSET */
uint64_t orc_malloc(orc_alloc* a, uint64_t size, uint64_t tid, uint32_t slot);
/* frees the allocation at byte offset; returns GAL_OK or GAL_ERR_INVALID */
int orc_free(orc_alloc* a, uint64_t byte_offset);
int orc_replay(orc_alloc* a, const gal_op* ops, uint64_t n, uint64_t* results);
/* Multi-lane replay with the semantics of gal_group_replay: op i is one warp
 * call site reached by the lanes of ops[i].lanes together (teams by class,
 * one claim per team, frees grouped by block); results[i * 64 + l]. */
int orc_group_replay(orc_alloc* a, const gal_group_op* ops, uint64_t n, uint64_t* results);
/* One such call site at a time: malloc by the lanes of `lanes` (sizes,
 * tids and out indexed by lane; spread != 0: the claims malloc makes under
 * load, GAL_OP_MALLOC_SPREAD), or free of the byte offsets of `lanes`. */
int orc_group_malloc(orc_alloc* a, uint64_t lanes, const uint64_t* sizes, const uint64_t* tids, uint32_t slot,
                     int spread, uint64_t* out);
int orc_group_free(orc_alloc* a, uint64_t lanes, const uint64_t* byte_offsets);
int orc_snapshot(orc_alloc* a, void* buf, uint64_t cap, uint64_t* needed);
/* count of internal consistency faults seen (reference traps) */
uint64_t orc_faults(orc_alloc* a);

/* ---- CPU baselines (bench.py cpu_baseline leg) ---- */
/* Mixed-size rounds through the oracle on one core (same request stream as
 * the GPU mixed workload); returns successful mallocs + frees, and seconds. */
int orc_bench_mixed(uint64_t pool_bytes, uint64_t seed, uint64_t threads_per_round,
                    uint32_t rounds, uint64_t* ops_done, uint64_t* misses,
                    double* seconds);
/* glibc malloc/free on `threads` host threads (0 = all hardware threads):
 * each thread draws log-uniform power-of-two sizes 16..4096, mallocs, writes
 * and checks a tag, frees. p50/p99 from 1-in-64 sampled op latencies. */
int orc_bench_sysmalloc(uint64_t ops_per_thread, uint32_t threads, uint64_t seed,
                        uint64_t* ops_done, double* seconds, double* p50_ns,
                        double* p99_ns, uint32_t* threads_used);
/* the same with every request fixed_size bytes (0: the log-uniform law) */
int orc_bench_sysmalloc_size(uint64_t ops_per_thread, uint32_t threads, uint64_t seed, uint64_t fixed_size,
                             uint64_t* ops_done, double* seconds, double* p50_ns, double* p99_ns,
                             uint32_t* threads_used);

#ifdef __cplusplus
}
#endif

#endif /* GALLATIN_ORACLE_H */
