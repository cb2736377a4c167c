// workloads.hip — the benchmark and test workloads of the MI355X Gallatin
// allocator (include/gallatin_workloads.h), a library of their own on top of
// the allocator's C-ABI (libgallatin_amd.so).
//
// A workload is built on an allocator handle: it reads the handle's
// description (gal_get_info: the instantiated triple and the allocator's
// device address), launches on the handle's stream (gal_stream) and holds
// the handle while it lives (gal_retain / gal_release). Its kernels call the
// same device interface a user kernel calls (Gallatin<...>::malloc/free,
// global_malloc/global_free), instantiated here for each supported triple.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include <gallatin/allocators/global_allocator.hpp>
#include <gallatin/data_structs/chaining_table.hpp>
#include <gallatin/workloads.hpp>

#include "../../include/gallatin_workloads.h"

namespace ga = gallatin::allocators;
namespace gw = gallatin::workloads;
namespace gu = gallatin::utils;
namespace wave = gallatin::wave;
namespace atom = gallatin::atom;

namespace {

// Activity counters of a run, striped over cache lines: a wave adds to the
// stripe of its index, so that the adds of a large grid do not all queue on
// one line (same-address atomics on one line run at about 1e8/s, which a
// kernel of 2^16 waves adding twice each would spend 1.3 ms on); the host
// sums the stripes.
constexpr int kCounterStripes = 64;
struct alignas(128) counter_stripe {
  unsigned long long mallocs, frees, misses, tag_errors;
};
struct run_counters {
  counter_stripe s[kCounterStripes];
  uint32_t errors_before, errors_after;  // the allocator's error count around the run (copied on the stream)
};

constexpr int kEventPairs = 96;

struct wl_ops;
struct serve_bins;

}  // namespace

struct gal_workload {
  gal_allocator* owner = nullptr;  // the allocator handle it is built on
  bool holds = false;              // took a hold on the owner (gal_retain)
  const wl_ops* ops = nullptr;
  void* dev = nullptr;  // Gallatin<...>* in device memory
  gal_info_t info{};
  hipStream_t stream = nullptr;  // the owner's stream
  uint32_t* control = nullptr;   // device control words of the allocator
  hipEvent_t ev[2 * kEventPairs] = {};
  run_counters* d_counters = nullptr;
  run_counters* h_counters = nullptr;  // pinned host copy, written on the stream
  double tick_ns = 10.0;  // device wall-clock tick
  // mixed workload state: the round's malloc results (one pointer per
  // thread), and the survivors of the rounds so far in per-wave regions of
  // packed entries (double buffered, a count per region; mixed_free_kernel)
  uint64_t* d_live = nullptr;
  uint64_t live_cap = 0;
  uint64_t* d_regions[2] = {nullptr, nullptr};
  uint32_t* d_region_counts[2] = {nullptr, nullptr};
  uint64_t region_cap = 0;  // regions per buffer
  uint64_t n_regions = 0;   // regions of buffer `cur` that hold the live set
  int cur = 0;
  uint64_t threads_per_round = 0;
  uint32_t rounds = 0;
  uint32_t rounds_done = 0;
  uint64_t mixed_seed = 0;
  // per-wave malloc latency samples (device wall-clock ticks) of the last
  // run, or of the runs kept since gal_latency_keep (lat_keep of them, each
  // lat_per_run samples long, lat_kept recorded so far)
  uint32_t* d_lat = nullptr;
  uint64_t lat_cap = 0;
  uint64_t lat_count = 0;
  uint32_t lat_keep = 0;
  uint32_t lat_kept = 0;
  uint64_t lat_per_run = 0;
  // batches, serving loops and slab tables made on this workload and not
  // yet destroyed: they use its counters and events, so it outlives them
  int children = 0;
};

// a batch of requests and the device buffers that bin and serve it
struct gal_batch {
  gal_workload* h = nullptr;
  uint64_t capacity = 0;
  uint64_t n = 0;            // requests in the batch
  uint64_t seed = 0;         // tags are drawn from (seed, request)
  uint64_t* d_sizes = nullptr;
  uint32_t* d_perm = nullptr;     // binned order of the last binning
  uint32_t* d_counts = nullptr;   // per binning wave and bin
  uint32_t* d_offsets = nullptr;  // per binning wave and bin
  uint32_t* d_totals = nullptr;   // per bin
  uint64_t** d_ptrs = nullptr;    // results of gal_batch_malloc
};

// a slab hash table over the handle's allocator
struct gal_slab_table {
  gal_workload* h = nullptr;
  void* dev = nullptr;     // chaining_table<u64, u64, 7, A> in device memory
  void* heads = nullptr;   // bucket heads
  uint64_t buckets = 0;
  void* counters = nullptr;  // slab_counters + 2 chain words
};

// a loaded serving workload and the device buffers of its runs
struct gal_serve {
  gal_workload* h = nullptr;
  uint64_t capacity = 0;
  uint64_t n = 0;
  // requests bin after bin; the per-request outputs follow the same order
  void* d_reqs = nullptr;  // serve_req[capacity]
  uint64_t* d_checksums = nullptr;
  uint64_t* d_admit = nullptr;  // wall-clock ticks after the start, ~0: aborted
  uint64_t* d_done = nullptr;
  void* d_state = nullptr;                // serve_state
  unsigned long long* d_hist = nullptr;   // per-page malloc latency histogram
  std::vector<uint64_t> arrival_ns;       // host copy of the arrivals, in device order
  std::vector<uint32_t> page_sizes;       // distinct page sizes of the loaded requests
  std::vector<uint32_t> order;            // caller's index of the request at each device position
  std::vector<uint32_t> bin_start, bin_count;  // per non-empty bin
  std::vector<double> bin_work;           // pages grown and drained per bin (for dealing waves)
  uint64_t max_need = 0;                  // largest request, bytes
  bool ran = false;
};
namespace {

struct wl_ops {
  uint64_t seg, min, max;
  uint32_t* (*control_of)(void* dev);  // device control words (nullptr on failure)
  void (*launch_plain)(gal_workload* h, uint64_t n, uint64_t size, uint64_t seed, uint32_t verify);
  void (*launch_mixed_malloc)(gal_workload* h, uint64_t base, uint64_t n, uint32_t round, uint64_t seed,
                              uint32_t verify, uint32_t* lat);
  void (*launch_mixed_free)(gal_workload* h, uint64_t n_fresh, uint64_t fresh_base, uint32_t round, uint64_t seed,
                            uint32_t verify, bool all);
  void (*launch_churn)(gal_workload* h, const gal_op* d_ops, uint64_t n, uint64_t seed, uint64_t* d_results);
  // batches (batch.inl)
  void (*batch_bin)(gal_batch* b);
  void (*batch_pass)(gal_batch* b, uint32_t mode, uint32_t width_log2, uint32_t run, uint32_t verify);
  void (*batch_per_thread)(gal_batch* b, uint32_t verify);
  // slab hash client (slab.inl)
  size_t slab_table_bytes;
  void (*slab_init)(gal_slab_table* t, uint64_t seed);
  void (*slab_launch)(gal_slab_table* t, int what, uint64_t n, uint64_t key_seed);
  // serving loop (serve.inl)
  uint64_t (*serve_headroom)(gal_workload* h, const std::vector<uint32_t>& page_sizes);
  unsigned (*serve_resident_blocks)();
  void (*serve_launch)(gal_serve* s, unsigned blocks, uint32_t run, uint64_t headroom, uint64_t deadline_ticks,
                       const serve_bins& bins, uint64_t window_ticks);
  uint32_t (*serve_bin)(uint64_t page_bytes);
};

// ---------------------------------------------------------------- kernels
__device__ inline bool tag_in(uint64_t* p, uint64_t tag, uint32_t verify) {
  if (verify == GAL_VERIFY_ATOMIC) return atom::exch(p, tag) == 0ull;
  *(volatile uint64_t*)p = tag;
  return *(volatile uint64_t*)p == tag;
}
__device__ inline bool tag_out(uint64_t* p, uint64_t tag, uint32_t verify) {
  if (verify == GAL_VERIFY_ATOMIC) return atom::exch(p, 0ull) == tag;
  bool ok = *(volatile uint64_t*)p == tag;
  *(volatile uint64_t*)p = 0ull;
  return ok;
}

// the calling wave's stripe of the run counters
__device__ inline counter_stripe* stripe_of(run_counters* c) {
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  return &c->s[wave % kCounterStripes];
}

// wave-aggregated counter update: one atomic per wave and counter
__device__ inline void count(run_counters* rc, bool m, bool f, bool miss, bool bad) {
  const uint64_t active = wave::active_mask();
  const uint32_t nm = wave::popc(wave::ballot(m)), nf = wave::popc(wave::ballot(f));
  const uint32_t nx = wave::popc(wave::ballot(miss)), nb = wave::popc(wave::ballot(bad));
  if ((int)wave::lane_id() == wave::first_lane(active)) {
    counter_stripe* c = stripe_of(rc);
    if (nm) atom::add(&c->mallocs, (unsigned long long)nm);
    if (nf) atom::add(&c->frees, (unsigned long long)nf);
    if (nx) atom::add(&c->misses, (unsigned long long)nx);
    if (nb) atom::add(&c->tag_errors, (unsigned long long)nb);
  }
}


template <class A>
__global__ void plain_kernel(A* a, uint64_t n, uint64_t size, uint64_t seed, uint32_t verify, run_counters* c) {
  const uint64_t tid = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  bool got = false, bad = false, miss = false;
  if (tid < n) {
    uint64_t* p = (uint64_t*)a->malloc(size);
    if (p) {
      got = true;
      const uint64_t tag = gw::payload_tag(seed, tid);
      bad = !tag_in(p, tag, verify);
      bad |= !tag_out(p, tag, verify);
      a->free(p);
    } else {
      miss = true;
    }
  }
  count(c, got, got, miss, bad);
}

// Agent-scope atomic rate probe. Mode GAL_ATOMIC_DISTINCT: every lane adds
// to its own word. Mode GAL_ATOMIC_WAVE_HOT: the first lane of every wave adds
// to one of `addresses` words (one per 128-B line), the access pattern of a
// wave-aggregated claim on a pinned block's counter. Returned values are
// consumed so that each add is a full round trip, as in the allocator.
__global__ void atomic_rate_kernel(uint32_t* words, uint64_t n, uint32_t iters, uint32_t mode, uint32_t addresses,
                                   uint32_t* sink) {
  const uint64_t tid = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  if (tid >= n) return;
  uint32_t acc = 0;
  if (mode == GAL_ATOMIC_DISTINCT) {
    for (uint32_t i = 0; i < iters; i++) acc += atom::add(&words[tid], 1u);
  } else if ((tid & 63) == 0) {
    uint32_t* w = &words[(uint64_t)((tid >> 6) % addresses) * 32];
    for (uint32_t i = 0; i < iters; i++) acc += atom::add(w, 64u);
  }
  if (acc == 0xFFFFFFFFu) *sink = acc;  // keeps the results live
}

// device wall clock (constant rate, see hipDeviceAttributeWallClockRate)
__device__ inline uint64_t wall_ticks() { return __builtin_amdgcn_s_memrealtime(); }

template <class A>
__global__ void mixed_malloc_kernel(A* a, uint64_t* live, uint64_t base, uint64_t n, uint32_t round, uint64_t seed,
                                    uint32_t verify, run_counters* c, uint32_t* lat) {
  // 32 bits: a round has fewer than 2^32 threads (gal_mixed_reset). After
  // malloc the index is rebuilt from the wave's first index (a scalar) and
  // the lane id: kept in a vector register across the call, it made the
  // kernel spill one (1-D blocks of whole waves: lane = threadIdx.x % 64).
  const uint32_t tid = threadIdx.x + blockIdx.x * blockDim.x;
  const uint32_t wave_first = __builtin_amdgcn_readfirstlane(tid) & ~63u;
  bool got = false, bad = false, miss = false;
  if (tid < n) {
    const uint64_t t0 = wall_ticks();
    uint64_t* p = (uint64_t*)a->malloc(gw::mixed_size(seed, round, tid));
    const uint64_t t1 = wall_ticks();
    const uint32_t me = wave_first + wave::lane_id();
    // one latency sample per wave: the wave's time inside malloc
    if (lat && (me & 63) == 0) lat[(base + me) >> 6] = (uint32_t)(t1 - t0);
    live[me] = (uint64_t)p;
    if (p) {
      got = true;
      bad = !tag_in(p, gw::payload_tag(seed, base + me), verify);
    } else {
      miss = true;
    }
  }
  count(c, got, false, miss, bad);
}

// The live table of the mixed workload. The malloc kernel of a round writes
// one pointer per thread (0 for a miss) into the round's fresh array; the
// free kernel that follows scans the survivors of earlier rounds and the
// fresh array, frees what the round's coin picks, and writes what it keeps
// into regions of kRegionEntries entries, one region per free-kernel wave,
// packed as (pool offset / 16) << 32 | live-table slot, with a count per
// region. The next free kernel reads those regions: every round scans only
// what is still live plus the fresh round, not every slot of every round so
// far. Regions are double buffered; a wave writes its survivors in scan
// order, with no atomics.
//
// Each wave packs the pointers it frees into full 64-lane batches in LDS
// before calling free, so that every free call runs with a full wave (one
// release fence per 64 frees) however sparse its entries are.
constexpr int kFreeScanIters = 16;
constexpr uint64_t kRegionEntries = 64ull * kFreeScanIters;
// entries each lane loads at once, ahead of the scan steps that use them
// (one round trip per kFreeAhead steps instead of one per step)
constexpr int kFreeAhead = 8;
static_assert(kFreeScanIters % kFreeAhead == 0, "whole batches of loads ahead");
constexpr int kFreeWavesPerBlock = 4;

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// packed survivor entry <-> (pointer, live-table slot)
__device__ inline uint64_t pack_live(uint64_t p, const char* base, uint64_t j) {
  return (((p - (uint64_t)base) >> 4) << 32) | j;
}

// waves 0 .. n_in-1 take one input region each; waves n_in .. take
// kRegionEntries fresh entries each (slot fresh_base + index)
template <class A>
__global__ void __launch_bounds__(64 * kFreeWavesPerBlock)
    mixed_free_kernel(A* a, const uint64_t* in_regions, const uint32_t* in_counts, uint64_t n_in,
                      const uint64_t* fresh, uint64_t n_fresh, uint64_t fresh_base, uint64_t* out_regions,
                      uint32_t* out_counts, uint32_t round, uint64_t seed, uint32_t verify, bool all,
                      run_counters* c) {
  __shared__ uint64_t pend_p[kFreeWavesPerBlock][128];
  __shared__ uint64_t pend_j[kFreeWavesPerBlock][128];
  const uint32_t lane = wave::lane_id();
  const uint32_t w = threadIdx.x >> 6;
  const uint64_t wave_index = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t fresh_waves = (n_fresh + kRegionEntries - 1) / kRegionEntries;
  if (wave_index >= n_in + fresh_waves) return;  // the whole wave
  const char* base = a->table.memory;
  const bool from_region = wave_index < n_in;
  const uint64_t first = from_region ? wave_index * kRegionEntries : (wave_index - n_in) * kRegionEntries;
  const uint64_t count = from_region ? (uint64_t)in_counts[wave_index]
                                     : (n_fresh - first < kRegionEntries ? n_fresh - first : kRegionEntries);
  uint64_t* out = out_regions + wave_index * kRegionEntries;
  uint32_t kept = 0;     // wave-uniform count of survivors written
  uint32_t pending = 0;  // wave-uniform count of packed frees
  uint64_t freed = 0, bad = 0;
  // The tags were stored by the malloc kernel, which has ended: in store
  // mode the tag is only read here, nothing is stored to the freed memory
  // and free_released applies. Atomic mode clears the tag (the next owner's
  // exchange expects 0), so its frees release.
  auto flush = [&](uint32_t k) {
    bool b = false;
    if (lane < k) {
      uint64_t* p = (uint64_t*)pend_p[w][lane];
      const uint64_t j = pend_j[w][lane];
      const uint64_t tag = gw::payload_tag(seed, j);
      if (verify == GAL_VERIFY_ATOMIC) {
        b = !tag_out(p, tag, verify);
        a->free(p);
      } else {
        b = *(volatile uint64_t*)p != tag;
        a->free_released(p);
      }
    }
    freed += k;
    bad += wave::popc(wave::ballot(b));
  };
  const uint32_t iters = (uint32_t)((count + 63) / 64);
  uint64_t ahead[kFreeAhead];
  for (uint32_t it = 0; it < iters; it++) {
    if (it % kFreeAhead == 0) {
#pragma unroll
      for (int q = 0; q < kFreeAhead; q++) {
        const uint64_t i = (uint64_t)(it + q) * 64 + lane;
        ahead[q] = i >= count ? 0ull : from_region ? in_regions[first + i] : fresh[first + i];
      }
    }
    const uint64_t i = (uint64_t)it * 64 + lane;
    const uint64_t e = ahead[it % kFreeAhead];
    bool live;
    uint64_t p, j;
    if (from_region) {
      live = i < count;
      p = (uint64_t)base + ((e >> 32) << 4);
      j = e & 0xFFFFFFFFull;
    } else {
      live = e != 0;
      p = e;
      j = fresh_base + first + i;
    }
    const bool pick = live && (all || gw::mixed_coin(seed, round, j));
    const bool keep = live && !pick;
    const uint64_t km = wave::ballot(keep);
    if (keep) out[kept + wave::popc(km & wave::lanemask_lt())] = from_region ? e : pack_live(p, base, j);
    kept += wave::popc(km);
    const uint64_t m = wave::ballot(pick);
    if (pick) {
      const uint32_t pos = pending + wave::popc(m & wave::lanemask_lt());
      pend_p[w][pos] = p;
      pend_j[w][pos] = j;
    }
    pending += wave::popc(m);
    wave_lds_sync();
    if (pending >= 64) {
      flush(64);
      pending -= 64;
      uint64_t mp = 0, mj = 0;
      if (lane < pending) {
        mp = pend_p[w][64 + lane];
        mj = pend_j[w][64 + lane];
      }
      wave_lds_sync();
      if (lane < pending) {
        pend_p[w][lane] = mp;
        pend_j[w][lane] = mj;
      }
      wave_lds_sync();
    }
  }
  if (pending > 0) flush(pending);
  if (lane == 0) out_counts[wave_index] = kept;
  if (lane == 0 && freed) atom::add(&stripe_of(c)->frees, (unsigned long long)freed);
  if (lane == 0 && bad) atom::add(&stripe_of(c)->tag_errors, (unsigned long long)bad);
}

// the reference's global test kernel shape (tests/src/global_test.cu:34-110):
// global_malloc, tag, check, global_free
__global__ void global_plain_kernel(uint64_t n, uint64_t size, uint64_t seed, uint32_t verify, run_counters* c) {
  const uint64_t tid = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  bool got = false, bad = false, miss = false;
  if (tid < n) {
    uint64_t* p = (uint64_t*)ga::global_malloc(size);
    if (p) {
      got = true;
      const uint64_t tag = gw::payload_tag(seed, tid);
      bad = !tag_in(p, tag, verify);
      bad |= !tag_out(p, tag, verify);
      ga::global_free(p);
    } else {
      miss = true;
    }
  }
  count(c, got, got, miss, bad);
}

// Concurrent mixed malloc/free: thread i runs op i (any interleaving with
// the others). A malloc tags the first word of its allocation with a tag of
// its offset (atomic exchange, expects 0); a free checks and clears it first.
template <class A>
__global__ void churn_kernel(A* a, const gal_op* ops, uint64_t n, uint64_t seed, uint64_t* results,
                             run_counters* c) {
  const uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  bool got = false, freed = false, miss = false, bad = false;
  if (i < n) {
    const gal_op op = ops[i];
    char* base = a->table.memory;
    if (op.kind == GAL_OP_MALLOC) {
      uint64_t* p = (uint64_t*)a->malloc(op.arg);
      if (p) {
        const uint64_t off = (uint64_t)((char*)p - base);
        bad = !tag_in(p, gw::payload_tag(seed, off), GAL_VERIFY_ATOMIC);
        results[i] = off;
        got = true;
      } else {
        results[i] = GAL_NULL_OFFSET;
        miss = true;
      }
    } else {
      uint64_t* p = (uint64_t*)(base + op.arg);
      bad = !tag_out(p, gw::payload_tag(seed, op.arg), GAL_VERIFY_ATOMIC);
      a->free(p);
      results[i] = 0;
      freed = true;
    }
  }
  count(c, got, freed, miss, bad);
}


// device pool first, host pool when it has none (global_malloc_fused); counts
// where each allocation landed
__global__ void global_fused_malloc_kernel(uint64_t n, uint64_t size, uint64_t** ptrs, run_counters* c,
                                           unsigned long long* on_host) {
  const uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  bool got = false, bad = false, miss = false, host = false;
  if (i < n) {
    uint64_t* p = (uint64_t*)ga::global_malloc_fused(size);
    ptrs[i] = p;
    if (p) {
      got = true;
      host = !ga::global_gallatin->owns_allocation(p);
      bad = !tag_in(p, gw::payload_tag(11, i), GAL_VERIFY_ATOMIC);
    } else {
      miss = true;
    }
  }
  count(c, got, false, miss, bad);
  const uint64_t h = wave::ballot(host);
  if (h && (int)wave::lane_id() == wave::first_lane(h)) atom::add(on_host, (unsigned long long)wave::popc(h));
}
__global__ void global_fused_free_kernel(uint64_t n, uint64_t** ptrs, run_counters* c) {
  const uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  bool freed = false, bad = false;
  if (i < n && ptrs[i]) {
    uint64_t* p = ptrs[i];
    bad = !tag_out(p, gw::payload_tag(11, i), GAL_VERIFY_ATOMIC);
    ga::global_free_fused(p);
    freed = true;
  }
  count(c, false, freed, false, bad);
}

// The reference's poison sequence (tests/src/poison_test.cu:60-128) for one
// size, in one lane: an intact allocation, then guard words and the byte
// after the caller's bytes overwritten. codes[k] = the check's result.
__device__ inline char* poison_alloc_and_write(uint64_t size) {
  char* p = (char*)ga::global_malloc_poison(size);
  if (p)
    for (uint64_t i = 0; i < size; i++) p[i] = (char)((i * i + 1) % 256);
  return p;
}
__global__ void global_poison_kernel(uint64_t size, int* codes) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int k = 0; k < 7; k++) {
    char* p = poison_alloc_and_write(size);
    if (!p) {
      codes[k] = -1;
      continue;
    }
    uint64_t* w = (uint64_t*)p;
    if (k == 1) w[-2] = 17;  // the header's size word
    if (k == 2) w[-1] = 17;  // the header's extra word
    if (k == 3) p[size] = 'c';  // the byte after the caller's bytes
    if (k == 4) w[-1] = w[-2] = 17;
    if (k == 5) w[-1] = w[-2] = 48;
    if (k == 6) p[size - 1] = 'c';  // the caller's own last byte: allowed
    codes[k] = ga::global_free_poison(p);
  }
}

inline unsigned blocks_for(uint64_t n, unsigned threads) { return (unsigned)((n + threads - 1) / threads); }


// Threads per block of the one-request-per-thread kernels. A block's waves
// hold their CU slots until the last of them ends, so larger blocks keep
// fewer claimers in flight at once; on the 8 GiB mixed workload 1024-thread
// blocks ran the full rounds 1.3 ms instead of 1.9 ms (256 threads), the
// claims queueing less on each block counter.
constexpr unsigned kClaimBlock = 1024;

#include "batch.inl"
#include "serve.inl"
#include "slab.inl"


// ------------------------------------------------------ per-type workloads
template <uint64_t S, uint64_t M, uint64_t B>
struct wl_impl {
  using A = ga::Gallatin<S, M, B>;

  static uint32_t* control_of(void* dev) {
    A host;
    return A::host_copy((const A*)dev, &host) ? host.control : nullptr;
  }

  static void launch_plain(gal_workload* h, uint64_t n, uint64_t size, uint64_t seed, uint32_t verify) {
    plain_kernel<A><<<blocks_for(n, kClaimBlock), kClaimBlock, 0, h->stream>>>((A*)h->dev, n, size, seed, verify, h->d_counters);
  }
  static void launch_mixed_malloc(gal_workload* h, uint64_t base, uint64_t n, uint32_t round, uint64_t seed,
                                  uint32_t verify, uint32_t* lat) {
    mixed_malloc_kernel<A><<<blocks_for(n, kClaimBlock), kClaimBlock, 0, h->stream>>>((A*)h->dev, h->d_live, base, n, round, seed,
                                                                       verify, h->d_counters, lat);
  }
  // the live set (buffer cur's regions) and n_fresh fresh entries in, the
  // survivors out into the other buffer, which becomes the live set
  static void launch_mixed_free(gal_workload* h, uint64_t n_fresh, uint64_t fresh_base, uint32_t round,
                                uint64_t seed, uint32_t verify, bool all) {
    const uint64_t n_in = h->n_regions;
    const uint64_t waves = n_in + (n_fresh + kRegionEntries - 1) / kRegionEntries;
    const int in = h->cur, out = h->cur ^ 1;
    if (waves > 0)
      mixed_free_kernel<A><<<(unsigned)((waves + kFreeWavesPerBlock - 1) / kFreeWavesPerBlock),
                             64 * kFreeWavesPerBlock, 0, h->stream>>>(
          (A*)h->dev, h->d_regions[in], h->d_region_counts[in], n_in, h->d_live, n_fresh, fresh_base,
          h->d_regions[out], h->d_region_counts[out], round, seed, verify, all, h->d_counters);
    h->cur = out;
    h->n_regions = all ? 0 : waves;  // a drain keeps nothing
  }
  static void launch_churn(gal_workload* h, const gal_op* d_ops, uint64_t n, uint64_t seed, uint64_t* d_results) {
    churn_kernel<A><<<blocks_for(n, 256), 256, 0, h->stream>>>((A*)h->dev, d_ops, n, seed, d_results, h->d_counters);
  }
  using table_t = gallatin::data_structs::chaining_table<uint64_t, uint64_t, kSlabSlots, A>;
  static void slab_init(gal_slab_table* t, uint64_t seed) {
    table_t host;
    host.heads = (typename table_t::block_type**)t->heads;
    host.nbuckets = t->buckets;
    host.seed = seed;
    host.empty = 0;
    host.alloc = (A*)t->h->dev;
    (void)hipMemcpy(t->dev, &host, sizeof(host), hipMemcpyHostToDevice);
  }
  // what: 0 insert, 1 query, 2 chain lengths, 3 free every slab
  static void slab_launch(gal_slab_table* t, int what, uint64_t n, uint64_t key_seed) {
    table_t* d = (table_t*)t->dev;
    slab_counters* c = (slab_counters*)t->counters;
    unsigned long long* chains = (unsigned long long*)((char*)t->counters + sizeof(slab_counters));
    hipStream_t st = t->h->stream;
    if (what == 0) slab_insert_kernel<table_t><<<blocks_for(n, 256), 256, 0, st>>>(d, n, key_seed, c);
    if (what == 1) slab_query_kernel<table_t><<<blocks_for(n, 256), 256, 0, st>>>(d, n, key_seed, c);
    if (what == 2) slab_chain_kernel<table_t><<<blocks_for(t->buckets, 256), 256, 0, st>>>(d, chains);
    if (what == 3) slab_free_kernel<table_t><<<blocks_for(t->buckets, 256), 256, 0, st>>>(d, chains);
  }

  static void batch_bin(gal_batch* b) {
    const uint64_t waves = (b->n + kBinWaveRequests - 1) / kBinWaveRequests;
    const unsigned grid = (unsigned)((waves + 3) / 4);
    bin_count_kernel<A><<<grid, 256, 0, b->h->stream>>>(b->d_sizes, b->n, b->d_counts);
    bin_scan_kernel<<<kBins, 256, 0, b->h->stream>>>(b->d_counts, waves, b->d_offsets, b->d_totals);
    bin_scatter_kernel<A><<<grid, 256, 0, b->h->stream>>>(b->d_sizes, b->n, b->d_offsets, b->d_totals, b->d_perm);
  }
  // persistent grid: the blocks that fit the GPU at once, no more than the chunks need
  static unsigned persistent_blocks(uint64_t n) {
    static int resident = 0;
    if (resident == 0) {
      int per_cu = 0, cus = 0, device = 0;
      if (hipGetDevice(&device) != hipSuccess ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<A>, 256, 0) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || per_cu < 1)
        per_cu = 1, cus = 256;
      resident = per_cu * cus;
    }
    const uint64_t need = (n + 255) / 256;
    return (unsigned)(need < (uint64_t)resident ? (need ? need : 1) : (uint64_t)resident);
  }
  static void batch_pass(gal_batch* b, uint32_t mode, uint32_t width_log2, uint32_t run, uint32_t verify) {
    batch_kernel<A><<<persistent_blocks((b->n + run - 1) / run), 256, 0, b->h->stream>>>(
        (A*)b->h->dev, b->d_sizes, b->d_perm, b->n, b->d_ptrs, mode, width_log2, run, b->seed, verify,
        b->h->d_counters);
  }
  static void batch_per_thread(gal_batch* b, uint32_t verify) {
    batch_plain_kernel<A><<<blocks_for(b->n, kClaimBlock), kClaimBlock, 0, b->h->stream>>>((A*)b->h->dev, b->d_sizes, b->n, b->seed,
                                                                           verify, b->h->d_counters);
  }

  // Memory the admission rule keeps back for the classes the page sizes
  // claim from: every slot of such a class holding a partly filled block and
  // a block fetched ahead, two segments per class attached and not yet used
  // up, and, when a class
  // has several blocks per segment, a tenth of the pool for recycled blocks
  // of segments not listed again yet (a segment is listed again once a
  // tenth of its blocks are back). Pages of the segment path need none.
  static uint64_t serve_headroom(gal_workload* h, const std::vector<uint32_t>& page_sizes) {
    bool used[A::kNumTrees] = {};
    bool slots_used[A::kNumTrees] = {};
    for (uint32_t size : page_sizes) {
      const uint32_t tree = size < M ? 0u : (uint32_t)(gu::get_first_bit_bigger(size) - A::kSmallestBits);
      if (tree < A::kNumTrees) {
        used[tree] = slots_used[tree] = true;
        continue;
      }
      const int block_tree = gu::get_first_bit_bigger(size) - A::kSmallestTreeBits;
      if (block_tree < 0) {
        used[A::kNumTrees - 1] = slots_used[A::kNumTrees - 1] = true;  // big slices
      } else if ((uint32_t)block_tree < A::kNumTrees) {
        used[block_tree] = true;  // whole blocks
      }
    }
    uint64_t bytes = 0;
    bool hidden = false;
    for (uint32_t c = 0; c < A::kNumTrees; c++) {
      if (!used[c]) continue;
      if (slots_used[c]) {
        const uint64_t block = A::table_type::tree_alloc_size(c) * ga::kSlicesPerBlock;
        bytes += 2 * (uint64_t)ga::usable_slots_of(c, h->info.num_segments * S, block) * block;
      }
      bytes += 2 * S;
      hidden |= A::table_type::blocks_in_segment(c) > 1;
    }
    if (hidden) bytes += h->info.num_segments * S / 10;
    return bytes;
  }
  static unsigned serve_resident_blocks() {
    static int resident = 0;
    if (resident == 0) {
      int per_cu = 0, cus = 0, device = 0;
      if (hipGetDevice(&device) != hipSuccess ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_kernel<A>, 256, 0) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || per_cu < 1)
        per_cu = 1, cus = 256;
      resident = per_cu * cus;
    }
    return (unsigned)resident;
  }
  static void serve_launch(gal_serve* s, unsigned blocks, uint32_t run, uint64_t headroom, uint64_t deadline_ticks,
                           const serve_bins& bins, uint64_t window_ticks) {
    serve_state* st = (serve_state*)s->d_state;
    serve_start_kernel<<<1, 64, 0, s->h->stream>>>(st);
    serve_kernel<A><<<blocks, 256, 0, s->h->stream>>>((A*)s->h->dev, (const serve_req*)s->d_reqs, s->n, st,
                                                      s->d_checksums, s->d_admit, s->d_done, s->d_hist, headroom,
                                                      deadline_ticks, run, bins, window_ticks);
  }
  // the bin of a page size: the class its pages claim from (big slices by
  // slice count), block and segment pages in the last bin
  static uint32_t serve_bin(uint64_t page_bytes) {
    const uint32_t t = page_bytes < M ? 0u : (uint32_t)(gu::get_first_bit_bigger(page_bytes) - A::kSmallestBits);
    return t < kServeBins - 1 ? t : kServeBins - 1;
  }


  static constexpr wl_ops ops() {
    return wl_ops{S,
                  M,
                  B,
                  &control_of,
                  &launch_plain,
                  &launch_mixed_malloc,
                  &launch_mixed_free,
                  &launch_churn,
                  &batch_bin,
                  &batch_pass,
                  &batch_per_thread,
                  sizeof(table_t),
                  &slab_init,
                  &slab_launch,
                  &serve_headroom,
                  &serve_resident_blocks,
                  &serve_launch,
                  &serve_bin};
  }
};

// The triples libgallatin_amd.so instantiates (gal_supported), each with
// its workload kernels.
constexpr uint64_t MiB = 1ull << 20;
const wl_ops kImpls[] = {
    wl_impl<16 * MiB, 16, 4096>::ops(), wl_impl<1 * MiB, 16, 256>::ops(),   wl_impl<4 * MiB, 128, 1024>::ops(),
    wl_impl<16 * MiB, 128, 4096>::ops(), wl_impl<64 * MiB, 128, 4096>::ops(), wl_impl<64 * MiB, 16, 16384>::ops(),
};

const wl_ops* find_impl(uint64_t seg, uint64_t min, uint64_t max) {
  for (const auto& o : kImpls)
    if (o.seg == seg && o.min == min && o.max == max) return &o;
  return nullptr;
}


// Events around launches on the handle's stream. A launch is timed from
// the event before it to the event after it; with `chain` set, launches
// that follow each other with nothing between them on the stream share
// their boundary event (one event a launch instead of two).
struct launch_timer {
  gal_workload* h;
  bool chain = false;
  int used = 0;              // launches timed
  int events = 0;            // events recorded
  std::vector<int> kinds;    // 0 malloc kernel, 1 free-only kernel
  std::vector<int> starts;   // event before each launch (the one after is next)
  explicit launch_timer(gal_workload* hh) : h(hh) {}
  bool begin() {
    if (events + 2 > 2 * kEventPairs) return false;
    if (chain && events > 0) {
      starts.push_back(events - 1);
      return true;
    }
    starts.push_back(events);
    return hipEventRecord(h->ev[events++], h->stream) == hipSuccess;
  }
  void end(int kind) {
    (void)hipEventRecord(h->ev[events++], h->stream);
    kinds.push_back(kind);
    used++;
  }
  int finish(gal_run_result* out) {
    if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
    if (hipGetLastError() != hipSuccess) return GAL_ERR_HIP;
    for (int i = 0; i < used; i++) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, h->ev[starts[i]], h->ev[starts[i] + 1]) != hipSuccess) return GAL_ERR_HIP;
      if (kinds[i] == 0)
        out->malloc_kernel_ms += ms;
      else
        out->free_kernel_ms += ms;
    }
    if (used > 0) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, h->ev[0], h->ev[events - 1]) != hipSuccess) return GAL_ERR_HIP;
      out->total_ms += ms;
    }
    out->launches += (uint64_t)used;
    return GAL_OK;
  }
};

// Counters zeroed and the allocator's error count taken before a run, the
// error count taken again after it: one small launch each, on the stream.
// The error word is read with a read-modify-write that changes nothing, so
// it counts every error an atomic recorded before.
__global__ void run_begin_kernel(run_counters* c, uint32_t* errors) {
  unsigned long long* w = (unsigned long long*)c->s;
  const uint32_t n = (uint32_t)(sizeof(c->s) / sizeof(unsigned long long));
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) w[i] = 0ull;
  if (threadIdx.x == 0) {
    c->errors_before = atom::read_current(errors);
    c->errors_after = 0u;
  }
}
__global__ void run_end_kernel(run_counters* c, uint32_t* errors) {
  if (threadIdx.x == 0) c->errors_after = atom::read_current(errors);
}

// run `body` (which launches through the timer) with zeroed counters; fill out
template <class F>
int timed_run(gal_workload* h, gal_run_result* out, F&& body) {
  gal_run_result r;
  std::memset(&r, 0, sizeof(r));
  // the error word is taken before and after the kernels on the same stream,
  // and the counters come back to pinned host memory on the stream, so that
  // the one synchronisation of the run returns them all
  uint32_t* errors = ga::ctl_word(h->control, ga::kCtlErrors);
  run_begin_kernel<<<1, 256, 0, h->stream>>>(h->d_counters, errors);
  if (hipGetLastError() != hipSuccess) return GAL_ERR_HIP;
  launch_timer t(h);
  int rc = body(t);
  if (rc != GAL_OK) {
    (void)hipStreamSynchronize(h->stream);
    return rc;
  }
  run_end_kernel<<<1, 64, 0, h->stream>>>(h->d_counters, errors);
  if (hipMemcpyAsync(h->h_counters, h->d_counters, sizeof(run_counters), hipMemcpyDeviceToHost, h->stream) !=
      hipSuccess)
    return GAL_ERR_HIP;
  rc = t.finish(&r);
  if (rc != GAL_OK) return rc;
  const run_counters& c = *h->h_counters;
  const uint64_t err0 = c.errors_before, err1 = c.errors_after;
  for (const counter_stripe& k : c.s) {
    r.mallocs += k.mallocs;
    r.frees += k.frees;
    r.misses += k.misses;
    r.tag_errors += k.tag_errors;
  }
  r.alloc_errors = err1 - err0;
  if (out) *out = r;
  return GAL_OK;
}

// This library's device code has its own global_gallatin and
// global_host_gallatin (a module-scope device variable is per code object).
// Before a workload of the global API they are pointed at the global
// allocators of libgallatin_amd.so, as any separately built module that
// calls global_malloc would point its own.
int bind_globals(gal_allocator** device, gal_allocator** host) {
  gal_allocator *d = nullptr, *hst = nullptr;
  if (gal_global_handles(&d, &hst) != GAL_OK) return GAL_ERR_STATE;
  ga::global_allocator_type* pd = nullptr;
  ga::global_allocator_type* ph = nullptr;
  gal_info_t info;
  if (d) {
    if (gal_get_info(d, &info) != GAL_OK) return GAL_ERR_STATE;
    pd = (ga::global_allocator_type*)info.device_handle;
  }
  if (hst) {
    if (gal_get_info(hst, &info) != GAL_OK) return GAL_ERR_STATE;
    ph = (ga::global_allocator_type*)info.device_handle;
  }
  if (hipMemcpyToSymbol(HIP_SYMBOL(ga::global_gallatin), &pd, sizeof(pd)) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(ga::global_host_gallatin), &ph, sizeof(ph)) != hipSuccess)
    return GAL_ERR_HIP;
  if (device) *device = d;
  if (host) *host = hst;
  return GAL_OK;
}

int make_workload(gal_allocator* h, bool hold, gal_workload** out);

// A workload on the global device allocator for one call of the global API,
// with this library's global pointers bound first. It takes no hold: the
// call returns before anything can free the allocator.
struct global_call {
  gal_workload* w = nullptr;
  gal_allocator* host = nullptr;
  int rc = GAL_OK;
  explicit global_call(bool need_host) {
    gal_allocator* device = nullptr;
    rc = bind_globals(&device, &host);
    if (rc == GAL_OK && (!device || (need_host && !host))) rc = GAL_ERR_INVALID;
    if (rc == GAL_OK) rc = make_workload(device, false, &w);
  }
  ~global_call() {
    if (w) (void)gal_workload_destroy(w);
  }
};

// a workload on h; `hold` takes a hold on h for the workload's lifetime
int make_workload(gal_allocator* h, bool hold, gal_workload** out) {
  *out = nullptr;
  gal_info_t info;
  if (!h || gal_get_info(h, &info) != GAL_OK) return GAL_ERR_INVALID;
  const wl_ops* ops = find_impl(info.segment_size, info.min_size, info.max_size);
  if (!ops) return GAL_ERR_UNSUPPORTED;
  void* stream = nullptr;
  if (gal_stream(h, &stream) != GAL_OK) return GAL_ERR_INVALID;
  gal_workload* w = new gal_workload();
  w->owner = h;
  w->ops = ops;
  w->info = info;
  w->dev = (void*)info.device_handle;
  w->stream = (hipStream_t)stream;
  int rc = GAL_OK;
  w->control = ops->control_of(w->dev);
  if (!w->control) rc = GAL_ERR_HIP;
  for (auto& e : w->ev)
    if (rc == GAL_OK && hipEventCreate(&e) != hipSuccess) rc = GAL_ERR_HIP;
  if (rc == GAL_OK && hipMalloc(&w->d_counters, sizeof(run_counters)) != hipSuccess) rc = GAL_ERR_OOM;
  if (rc == GAL_OK && hipHostMalloc((void**)&w->h_counters, sizeof(run_counters), hipHostMallocDefault) != hipSuccess)
    rc = GAL_ERR_OOM;
  int khz = 0, device = 0;
  if (hipGetDevice(&device) == hipSuccess &&
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
    w->tick_ns = 1e6 / (double)khz;
  if (rc == GAL_OK && hold) {
    if (gal_retain(h) != GAL_OK) rc = GAL_ERR_INVALID;
    w->holds = rc == GAL_OK;
  }
  if (rc != GAL_OK) {
    gal_workload_destroy(w);
    return rc;
  }
  *out = w;
  return GAL_OK;
}

}  // namespace

extern "C" {

int gal_workload_create(gal_allocator* h, gal_workload** out) {
  if (!h || !out) return GAL_ERR_INVALID;
  return make_workload(h, true, out);
}

int gal_workload_destroy(gal_workload* w) {
  if (!w) return GAL_ERR_INVALID;
  if (w->children > 0) return GAL_ERR_STATE;  // a batch, serving loop or slab table still uses it
  if (w->stream) (void)hipStreamSynchronize(w->stream);
  if (w->d_live) (void)hipFree(w->d_live);
  for (int b = 0; b < 2; b++) {
    if (w->d_regions[b]) (void)hipFree(w->d_regions[b]);
    if (w->d_region_counts[b]) (void)hipFree(w->d_region_counts[b]);
  }
  if (w->d_lat) (void)hipFree(w->d_lat);
  if (w->d_counters) (void)hipFree(w->d_counters);
  if (w->h_counters) (void)hipHostFree(w->h_counters);
  for (auto& e : w->ev)
    if (e) (void)hipEventDestroy(e);
  if (w->holds) (void)gal_release(w->owner);
  delete w;
  return GAL_OK;
}

int gal_run_plain(gal_workload* h, uint64_t n_threads, uint64_t size, uint64_t seed, uint32_t verify,
                  gal_run_result* out) {
  if (!h || n_threads == 0 || verify > GAL_VERIFY_ATOMIC) return GAL_ERR_INVALID;
  return timed_run(h, out, [&](launch_timer& t) {
    if (!t.begin()) return GAL_ERR_INVALID;
    h->ops->launch_plain(h, n_threads, size, seed, verify);
    t.end(0);
    return GAL_OK;
  });
}

int gal_mixed_reset(gal_workload* h, uint64_t threads_per_round, uint32_t rounds, uint64_t seed) {
  if (!h || threads_per_round == 0 || rounds == 0) return GAL_ERR_INVALID;
  // survivors are packed as (pool offset / 16) << 32 | live-table slot
  const uint64_t need = threads_per_round * rounds;
  if (need >= (1ull << 32) || h->info.num_segments * h->info.segment_size > (16ull << 32)) return GAL_ERR_INVALID;
  // free whatever the previous run left live
  if (h->d_live && h->n_regions > 0) {
    h->ops->launch_mixed_free(h, 0, 0, 0, h->mixed_seed, GAL_VERIFY_STORE, true);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  }
  h->rounds_done = 0;  // nothing of that run is live any more, whatever fails below
  h->n_regions = 0;
  // the fresh array is written by each round's malloc kernel before its free
  // kernel reads it, and a region's entries below its count are written
  // before any kernel reads them: nothing needs clearing
  if (threads_per_round > h->live_cap) {
    if (h->d_live) (void)hipFree(h->d_live);
    h->d_live = nullptr;
    h->live_cap = 0;
    if (hipMalloc(&h->d_live, threads_per_round * 8) != hipSuccess) return GAL_ERR_OOM;
    h->live_cap = threads_per_round;
  }
  const uint64_t regions = (uint64_t)rounds * ((threads_per_round + kRegionEntries - 1) / kRegionEntries);
  if (regions > h->region_cap) {
    for (int b = 0; b < 2; b++) {
      if (h->d_regions[b]) (void)hipFree(h->d_regions[b]);
      if (h->d_region_counts[b]) (void)hipFree(h->d_region_counts[b]);
      h->d_regions[b] = nullptr;
      h->d_region_counts[b] = nullptr;
    }
    h->region_cap = 0;
    for (int b = 0; b < 2; b++)
      if (hipMalloc(&h->d_regions[b], regions * kRegionEntries * 8) != hipSuccess ||
          hipMalloc(&h->d_region_counts[b], regions * 4) != hipSuccess)
        return GAL_ERR_OOM;
    h->region_cap = regions;
  }
  const uint64_t per_run = (need + 63) / 64;
  if (h->lat_keep && h->lat_kept > 0 && per_run != h->lat_per_run) return GAL_ERR_INVALID;  // kept runs differ in shape
  const uint64_t lat_need = per_run * (h->lat_keep ? h->lat_keep : 1) + 1;
  if (lat_need > h->lat_cap) {
    if (h->d_lat) (void)hipFree(h->d_lat);
    h->d_lat = nullptr;
    h->lat_cap = 0;
    if (hipMalloc(&h->d_lat, lat_need * 4) != hipSuccess) return GAL_ERR_OOM;
    h->lat_cap = lat_need;
  }
  if (!h->lat_keep) h->lat_count = 0;
  h->lat_per_run = per_run;
  h->threads_per_round = threads_per_round;
  h->rounds = rounds;
  h->rounds_done = 0;
  h->mixed_seed = seed;
  return GAL_OK;
}

int gal_mixed_round(gal_workload* h, uint32_t round, uint32_t verify, gal_run_result* out) {
  const uint64_t seed = h ? h->mixed_seed : 0;
  if (!h || !h->d_live || round != h->rounds_done || round >= h->rounds || verify > GAL_VERIFY_ATOMIC)
    return GAL_ERR_INVALID;
  int rc = timed_run(h, out, [&](launch_timer& t) {
    if (!t.begin()) return GAL_ERR_INVALID;
    h->ops->launch_mixed_malloc(h, (uint64_t)round * h->threads_per_round, h->threads_per_round, round, seed, verify,
                                h->lat_keep ? nullptr : h->d_lat);
    t.end(0);
    if (!t.begin()) return GAL_ERR_INVALID;
    h->ops->launch_mixed_free(h, h->threads_per_round, (uint64_t)round * h->threads_per_round, round, seed, verify,
                              false);
    t.end(1);
    return GAL_OK;
  });
  if (rc == GAL_OK) {
    h->rounds_done++;
    if (!h->lat_keep) h->lat_count = ((uint64_t)h->rounds_done * h->threads_per_round + 63) / 64;
  }
  return rc;
}

int gal_mixed_drain(gal_workload* h, uint32_t verify, gal_run_result* out) {
  if (!h || !h->d_live || verify > GAL_VERIFY_ATOMIC) return GAL_ERR_INVALID;
  const uint64_t seed = h->mixed_seed;
  int rc = timed_run(h, out, [&](launch_timer& t) {
    if (h->n_regions == 0) return GAL_OK;
    if (!t.begin()) return GAL_ERR_INVALID;
    h->ops->launch_mixed_free(h, 0, 0, 0, seed, verify, true);
    t.end(1);
    return GAL_OK;
  });
  if (rc == GAL_OK) h->rounds_done = 0;
  return rc;
}

int gal_mixed_live(gal_workload* h, uint64_t* offsets, uint64_t* sizes, uint64_t* slots, uint64_t cap,
                   uint64_t* count) {
  if (!h || !count) return GAL_ERR_INVALID;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  const uint64_t n = h->n_regions;
  std::vector<uint32_t> counts(n);
  std::vector<uint64_t> entries(n * kRegionEntries);
  if (n && (hipMemcpy(counts.data(), h->d_region_counts[h->cur], n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(entries.data(), h->d_regions[h->cur], n * kRegionEntries * 8, hipMemcpyDeviceToHost) !=
                hipSuccess))
    return GAL_ERR_HIP;
  uint64_t k = 0;
  for (uint64_t r = 0; r < n; r++) {
    for (uint32_t i = 0; i < counts[r] && i < kRegionEntries; i++) {
      const uint64_t e = entries[r * kRegionEntries + i];
      const uint64_t j = e & 0xFFFFFFFFull;
      if (k < cap) {
        if (offsets) offsets[k] = (e >> 32) << 4;
        if (sizes) sizes[k] = gw::mixed_size(h->mixed_seed, j / h->threads_per_round, j % h->threads_per_round);
        if (slots) slots[k] = j;
      }
      k++;
    }
  }
  *count = k;
  return (offsets && k > cap) ? GAL_ERR_BUFFER : GAL_OK;
}

int gal_run_mixed(gal_workload* h, uint64_t threads_per_round, uint32_t rounds, uint64_t seed, uint32_t verify,
                  gal_run_result* out) {
  int rc = gal_mixed_reset(h, threads_per_round, rounds, seed);
  if (rc != GAL_OK) return rc;
  gal_run_result total;
  std::memset(&total, 0, sizeof(total));
  // samples at the start of the buffer, or after the runs kept so far (none
  // once gal_latency_keep's count is reached)
  uint32_t* lat = h->d_lat;
  if (h->lat_keep) lat = h->lat_kept < h->lat_keep ? h->d_lat + (uint64_t)h->lat_kept * h->lat_per_run : nullptr;
  rc = timed_run(h, &total, [&](launch_timer& t) {
    t.chain = true;  // the launches follow each other with nothing between
    for (uint32_t r = 0; r < rounds; r++) {
      if (!t.begin()) return GAL_ERR_INVALID;
      h->ops->launch_mixed_malloc(h, (uint64_t)r * threads_per_round, threads_per_round, r, seed, verify, lat);
      t.end(0);
      if (!t.begin()) return GAL_ERR_INVALID;
      h->ops->launch_mixed_free(h, threads_per_round, (uint64_t)r * threads_per_round, r, seed, verify, false);
      t.end(1);
    }
    if (!t.begin()) return GAL_ERR_INVALID;
    h->ops->launch_mixed_free(h, 0, 0, 0, seed, verify, true);
    t.end(1);
    return GAL_OK;
  });
  if (rc == GAL_OK) {
    h->rounds_done = 0;  // drained
    if (!h->lat_keep) {
      h->lat_count = h->lat_per_run;
    } else if (lat) {
      h->lat_kept++;
      h->lat_count = (uint64_t)h->lat_kept * h->lat_per_run;
    }
  }
  if (out) *out = total;
  return rc;
}

int gal_latency_keep(gal_workload* h, uint32_t runs) {
  if (!h) return GAL_ERR_INVALID;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  h->lat_keep = runs;
  h->lat_kept = 0;
  h->lat_count = 0;
  return GAL_OK;
}

int gal_latency_samples(gal_workload* h, uint32_t* ticks, uint64_t cap, uint64_t* count, double* tick_ns) {
  if (!h || !count) return GAL_ERR_INVALID;
  *count = h->lat_count;
  if (tick_ns) *tick_ns = h->tick_ns;
  if (!ticks) return GAL_OK;
  if (cap < h->lat_count) return GAL_ERR_BUFFER;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  if (h->lat_count && hipMemcpy(ticks, h->d_lat, h->lat_count * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return GAL_ERR_HIP;
  return GAL_OK;
}

static int width_log2_of(uint32_t width, uint32_t* out) {
  for (uint32_t l = 0; l <= 6; l++)
    if ((1u << l) == width) {
      *out = l;
      return GAL_OK;
    }
  return GAL_ERR_INVALID;
}

int gal_batch_destroy(gal_batch* b) {
  if (!b) return GAL_ERR_INVALID;
  if (b->h) (void)hipStreamSynchronize(b->h->stream);
  void* bufs[] = {b->d_sizes, b->d_perm, b->d_counts, b->d_offsets, b->d_totals, b->d_ptrs};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (b->h) b->h->children--;
  delete b;
  return GAL_OK;
}

int gal_batch_create(gal_workload* h, uint64_t capacity, gal_batch** out) {
  if (!h || !out || capacity == 0 || capacity > 0xFFFFFFFFull) return GAL_ERR_INVALID;
  *out = nullptr;
  gal_batch* b = new gal_batch();
  b->h = h;
  h->children++;
  b->capacity = capacity;
  const uint64_t waves = (capacity + kBinWaveRequests - 1) / kBinWaveRequests;
  const bool ok = hipMalloc(&b->d_sizes, capacity * 8) == hipSuccess &&
                  hipMalloc(&b->d_perm, capacity * 4) == hipSuccess &&
                  hipMalloc(&b->d_counts, waves * kBins * 4) == hipSuccess &&
                  hipMalloc(&b->d_offsets, waves * kBins * 4) == hipSuccess &&
                  hipMalloc(&b->d_totals, kBins * 4) == hipSuccess && hipMalloc(&b->d_ptrs, capacity * 8) == hipSuccess &&
                  hipMemsetAsync(b->d_ptrs, 0, capacity * 8, h->stream) == hipSuccess &&
                  hipStreamSynchronize(h->stream) == hipSuccess;
  if (!ok) {
    gal_batch_destroy(b);
    return GAL_ERR_OOM;
  }
  *out = b;
  return GAL_OK;
}

int gal_batch_generate(gal_batch* b, uint64_t n, uint64_t seed, uint32_t size_classes) {
  if (!b || n == 0 || n > b->capacity || size_classes == 0 || size_classes > 40) return GAL_ERR_INVALID;
  b->n = n;
  b->seed = seed;
  batch_sizes_kernel<<<blocks_for(n, 256), 256, 0, b->h->stream>>>(b->d_sizes, n, seed, size_classes);
  return hipStreamSynchronize(b->h->stream) == hipSuccess ? GAL_OK : GAL_ERR_HIP;
}

int gal_batch_set_sizes(gal_batch* b, const uint64_t* sizes, uint64_t n, uint64_t seed) {
  if (!b || !sizes || n == 0 || n > b->capacity) return GAL_ERR_INVALID;
  if (hipMemcpy(b->d_sizes, sizes, n * 8, hipMemcpyHostToDevice) != hipSuccess) return GAL_ERR_HIP;
  b->n = n;
  b->seed = seed;
  return GAL_OK;
}

int gal_batch_run(gal_batch* b, uint32_t form, uint32_t width, uint32_t run, uint32_t replays, uint32_t verify,
                  gal_run_result* out) {
  uint32_t wl = 6;
  if (!b || b->n == 0 || replays == 0 || verify > GAL_VERIFY_ATOMIC || form > GAL_FORM_GRAPH ||
      width_log2_of(width, &wl) != GAL_OK || run == 0 || run > kMaxRun)
    return GAL_ERR_INVALID;
  gal_workload* h = b->h;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (form == GAL_FORM_GRAPH) {
    if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed) != hipSuccess) return GAL_ERR_HIP;
    h->ops->batch_bin(b);
    h->ops->batch_pass(b, kBatchMalloc | kBatchFree, wl, run, verify);
    if (hipStreamEndCapture(h->stream, &graph) != hipSuccess) return GAL_ERR_HIP;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return GAL_ERR_HIP;
    }
  }
  int rc = timed_run(h, out, [&](launch_timer& t) {
    if (!t.begin()) return GAL_ERR_INVALID;
    for (uint32_t r = 0; r < replays; r++) {
      if (form == GAL_FORM_PER_THREAD) {
        h->ops->batch_per_thread(b, verify);
      } else if (form == GAL_FORM_BATCHED) {
        h->ops->batch_bin(b);
        h->ops->batch_pass(b, kBatchMalloc | kBatchFree, wl, run, verify);
      } else if (hipGraphLaunch(exec, h->stream) != hipSuccess) {
        return GAL_ERR_HIP;
      }
    }
    t.end(0);
    return GAL_OK;
  });
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  return rc;
}

int gal_batch_malloc(gal_batch* b, uint32_t width, uint32_t run, gal_run_result* out) {
  uint32_t wl = 6;
  if (!b || b->n == 0 || width_log2_of(width, &wl) != GAL_OK || run == 0 || run > kMaxRun) return GAL_ERR_INVALID;
  return timed_run(b->h, out, [&](launch_timer& t) {
    if (!t.begin()) return GAL_ERR_INVALID;
    b->h->ops->batch_bin(b);
    b->h->ops->batch_pass(b, kBatchMalloc, wl, run, GAL_VERIFY_STORE);
    t.end(0);
    return GAL_OK;
  });
}

int gal_batch_free(gal_batch* b, gal_run_result* out) {
  if (!b || b->n == 0) return GAL_ERR_INVALID;
  return timed_run(b->h, out, [&](launch_timer& t) {
    if (!t.begin()) return GAL_ERR_INVALID;
    b->h->ops->batch_pass(b, kBatchFree, 6, 1, GAL_VERIFY_STORE);
    t.end(1);
    return GAL_OK;
  });
}

int gal_batch_offsets(gal_batch* b, uint64_t* offsets, uint64_t cap) {
  if (!b || !offsets || cap < b->n) return GAL_ERR_INVALID;
  if (hipStreamSynchronize(b->h->stream) != hipSuccess) return GAL_ERR_HIP;
  if (hipMemcpy(offsets, b->d_ptrs, b->n * 8, hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
  const uint64_t base = b->h->info.pool_device_ptr;
  for (uint64_t i = 0; i < b->n; i++) offsets[i] = offsets[i] ? offsets[i] - base : GAL_NULL_OFFSET;
  return GAL_OK;
}

int gal_batch_order(gal_batch* b, uint32_t* order, uint64_t cap) {
  if (!b || !order || cap < b->n) return GAL_ERR_INVALID;
  if (hipStreamSynchronize(b->h->stream) != hipSuccess) return GAL_ERR_HIP;
  return hipMemcpy(order, b->d_perm, b->n * 4, hipMemcpyDeviceToHost) == hipSuccess ? GAL_OK : GAL_ERR_HIP;
}

int gal_churn(gal_workload* h, const gal_op* ops, uint64_t n, uint64_t seed, uint64_t* results,
              gal_run_result* out) {
  if (!h || !ops || !results || n == 0) return GAL_ERR_INVALID;
  const uint64_t pool = h->info.num_segments * h->info.segment_size;
  for (uint64_t i = 0; i < n; i++) {
    if (ops[i].kind == GAL_OP_MALLOC) {
      if (ops[i].arg == 0) return GAL_ERR_INVALID;
    } else if (ops[i].kind != GAL_OP_FREE || ops[i].arg >= pool || (ops[i].arg & 7) != 0) {
      return GAL_ERR_INVALID;
    }
  }
  gal_op* d_ops = nullptr;
  uint64_t* d_res = nullptr;
  int rc = GAL_OK;
  if (hipMalloc(&d_ops, n * sizeof(gal_op)) != hipSuccess || hipMalloc(&d_res, n * 8) != hipSuccess) rc = GAL_ERR_OOM;
  if (rc == GAL_OK && hipMemcpy(d_ops, ops, n * sizeof(gal_op), hipMemcpyHostToDevice) != hipSuccess) rc = GAL_ERR_HIP;
  if (rc == GAL_OK)
    rc = timed_run(h, out, [&](launch_timer& t) {
      if (!t.begin()) return GAL_ERR_INVALID;
      h->ops->launch_churn(h, d_ops, n, seed, d_res);
      t.end(0);
      return GAL_OK;
    });
  if (rc == GAL_OK && hipMemcpy(results, d_res, n * 8, hipMemcpyDeviceToHost) != hipSuccess) rc = GAL_ERR_HIP;
  if (d_ops) (void)hipFree(d_ops);
  if (d_res) (void)hipFree(d_res);
  return rc;
}

// ---- slab hash client
int gal_slab_create(gal_workload* h, uint64_t buckets, uint64_t seed, gal_slab_table** out) {
  if (!h || !out || buckets == 0) return GAL_ERR_INVALID;
  *out = nullptr;
  gal_slab_table* t = new gal_slab_table();
  t->h = h;
  t->buckets = buckets;
  const size_t ctr = sizeof(slab_counters) + 2 * sizeof(unsigned long long);
  if (hipMalloc(&t->dev, h->ops->slab_table_bytes) != hipSuccess || hipMalloc(&t->heads, buckets * 8) != hipSuccess ||
      hipMalloc(&t->counters, ctr) != hipSuccess || hipMemset(t->heads, 0, buckets * 8) != hipSuccess) {
    if (t->dev) (void)hipFree(t->dev);
    if (t->heads) (void)hipFree(t->heads);
    if (t->counters) (void)hipFree(t->counters);
    delete t;
    return GAL_ERR_OOM;
  }
  h->children++;
  h->ops->slab_init(t, seed);
  *out = t;
  return GAL_OK;
}

// run one slab kernel with zeroed counters; fills the result's counts
static int slab_run(gal_slab_table* t, int what, uint64_t n, uint64_t key_seed, gal_slab_result* out) {
  gal_workload* h = t->h;
  const size_t ctr = sizeof(slab_counters) + 2 * sizeof(unsigned long long);
  gal_run_result base;
  std::memset(&base, 0, sizeof(base));
  int rc = timed_run(h, &base, [&](launch_timer& tm) {
    if (hipMemsetAsync(t->counters, 0, ctr, h->stream) != hipSuccess) return GAL_ERR_HIP;
    if (!tm.begin()) return GAL_ERR_INVALID;
    h->ops->slab_launch(t, what, n, key_seed);
    tm.end(0);
    return GAL_OK;
  });
  if (rc != GAL_OK) return rc;
  unsigned char buf[sizeof(slab_counters) + 16];
  if (hipMemcpy(buf, t->counters, ctr, hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
  slab_counters c;
  unsigned long long chains[2];
  std::memcpy(&c, buf, sizeof(c));
  std::memcpy(chains, buf + sizeof(c), 16);
  if (out) {
    std::memset(out, 0, sizeof(*out));
    out->keys = n;
    out->failed = c.failed;
    out->wrong = c.wrong;
    out->slabs = chains[0];
    out->max_chain = chains[1];
    out->alloc_errors = base.alloc_errors;
    out->kernel_ms = base.malloc_kernel_ms;
  }
  return GAL_OK;
}

int gal_slab_insert(gal_slab_table* t, uint64_t n, uint64_t key_seed, gal_slab_result* out) {
  if (!t || n == 0) return GAL_ERR_INVALID;
  gal_slab_result r, chains;
  int rc = slab_run(t, 0, n, key_seed, &r);
  if (rc == GAL_OK) rc = slab_run(t, 2, 0, 0, &chains);
  if (rc != GAL_OK) return rc;
  r.slabs = chains.slabs;
  r.max_chain = chains.max_chain;
  if (out) *out = r;
  return GAL_OK;
}

int gal_slab_query(gal_slab_table* t, uint64_t n, uint64_t key_seed, gal_slab_result* out) {
  if (!t || n == 0) return GAL_ERR_INVALID;
  return slab_run(t, 1, n, key_seed, out);
}

int gal_slab_destroy(gal_slab_table* t, uint64_t* slabs_freed) {
  if (!t) return GAL_ERR_INVALID;
  gal_slab_result r;
  int rc = slab_run(t, 3, 0, 0, &r);
  if (slabs_freed) *slabs_freed = rc == GAL_OK ? r.slabs : 0;
  (void)hipFree(t->dev);
  (void)hipFree(t->heads);
  (void)hipFree(t->counters);
  t->h->children--;
  delete t;
  return rc;
}

// ---- serving loop
static double unit_draw(uint64_t seed, uint64_t i, uint64_t salt) {
  const uint64_t x = gu::splitmix64(seed ^ gu::splitmix64((i << 4) | salt));
  return ((double)(x >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
}

int gal_serve_generate(const gal_serve_workload* w, gal_request* out) {
  if (!w || !out || w->requests == 0 || w->max_pages == 0 || !(w->median_pages >= 1.0) || !(w->sigma >= 0.0) ||
      !(w->rate_per_s >= 0.0))
    return GAL_ERR_INVALID;
  if (w->page_bytes == 0 && (w->size_classes == 0 || w->size_classes > 28)) return GAL_ERR_INVALID;
  if (w->page_bytes != 0 && w->page_bytes < 16) return GAL_ERR_INVALID;
  const double two_pi = 6.283185307179586476925;
  double t = 0.0;
  for (uint64_t i = 0; i < w->requests; i++) {
    if (w->rate_per_s > 0.0) t += -std::log(unit_draw(w->seed, i, 0)) / w->rate_per_s;
    // Box-Muller: z standard normal
    const double z = std::sqrt(-2.0 * std::log(unit_draw(w->seed, i, 1))) * std::cos(two_pi * unit_draw(w->seed, i, 2));
    double len = std::ceil(w->median_pages * std::exp(w->sigma * z));
    if (!(len >= 1.0)) len = 1.0;
    if (len > (double)w->max_pages) len = (double)w->max_pages;
    gal_request& r = out[i];
    r.arrival_ns = (uint64_t)(t * 1e9);
    r.id = i;
    r.pages = (uint32_t)len;
    r.page_bytes = w->page_bytes ? w->page_bytes
                                 : (uint32_t)(16u << (gu::splitmix64(w->seed ^ gu::splitmix64((i << 4) | 3)) %
                                                      w->size_classes));
  }
  return GAL_OK;
}

int gal_serve_destroy(gal_serve* s) {
  if (!s) return GAL_ERR_INVALID;
  if (s->h) (void)hipStreamSynchronize(s->h->stream);
  void* bufs[] = {s->d_reqs, s->d_checksums, s->d_admit, s->d_done, s->d_state, s->d_hist};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (s->h) s->h->children--;
  delete s;
  return GAL_OK;
}

int gal_serve_create(gal_workload* h, uint64_t capacity, gal_serve** out) {
  if (!h || !out || capacity == 0) return GAL_ERR_INVALID;
  *out = nullptr;
  gal_serve* s = new gal_serve();
  s->h = h;
  h->children++;
  s->capacity = capacity;
  const bool ok = hipMalloc(&s->d_reqs, capacity * sizeof(serve_req)) == hipSuccess &&
                  hipMalloc(&s->d_checksums, capacity * 8) == hipSuccess &&
                  hipMalloc(&s->d_admit, capacity * 8) == hipSuccess &&
                  hipMalloc(&s->d_done, capacity * 8) == hipSuccess &&
                  hipMalloc(&s->d_state, sizeof(serve_state)) == hipSuccess &&
                  hipMalloc(&s->d_hist, kHistBuckets * sizeof(unsigned long long)) == hipSuccess;
  if (!ok) {
    gal_serve_destroy(s);
    return GAL_ERR_OOM;
  }
  *out = s;
  return GAL_OK;
}

int gal_serve_load(gal_serve* s, const gal_request* reqs, uint64_t n) {
  if (!s || !reqs || n == 0 || n > s->capacity) return GAL_ERR_INVALID;
  // every request is checked before anything of the previous load changes
  for (uint64_t i = 0; i < n; i++)
    if (reqs[i].pages == 0 || reqs[i].page_bytes < 16) return GAL_ERR_INVALID;
  s->ran = false;
  s->n = 0;  // nothing loaded until the copy below has succeeded
  s->page_sizes.clear();
  s->max_need = 0;
  for (uint64_t i = 0; i < n; i++) {
    if (std::find(s->page_sizes.begin(), s->page_sizes.end(), reqs[i].page_bytes) == s->page_sizes.end())
      s->page_sizes.push_back(reqs[i].page_bytes);
    const uint64_t need = (uint64_t)reqs[i].pages * reqs[i].page_bytes;
    if (need > s->max_need) s->max_need = need;
  }
  // stable counting sort by bin; empty bins dropped
  std::vector<uint32_t> raw(n), counts(kServeBins, 0);
  std::vector<double> work(kServeBins, 0.0);
  for (uint64_t i = 0; i < n; i++) {
    raw[i] = s->h->ops->serve_bin(reqs[i].page_bytes);
    counts[raw[i]]++;
    work[raw[i]] += 2.0 * reqs[i].pages + 2.0;
  }
  std::vector<uint32_t> next(kServeBins, 0);
  s->bin_start.clear();
  s->bin_count.clear();
  s->bin_work.clear();
  uint32_t pos = 0;
  for (uint32_t b = 0; b < kServeBins; b++) {
    if (!counts[b]) continue;
    next[b] = pos;
    s->bin_start.push_back(pos);
    s->bin_count.push_back(counts[b]);
    s->bin_work.push_back(work[b]);
    pos += counts[b];
  }
  s->order.assign(n, 0);
  for (uint64_t i = 0; i < n; i++) s->order[next[raw[i]]++] = (uint32_t)i;
  // the device copy in that order
  std::vector<serve_req> dev(n);
  s->arrival_ns.resize(n);
  const double tick_ns = s->h->tick_ns;
  for (uint64_t p = 0; p < n; p++) {
    const gal_request& q = reqs[s->order[p]];
    dev[p].arrival = (uint64_t)((double)q.arrival_ns / tick_ns);
    dev[p].id = q.id;
    dev[p].pages = q.pages;
    dev[p].page_bytes = q.page_bytes;
    s->arrival_ns[p] = q.arrival_ns;
  }
  if (hipMemcpy(s->d_reqs, dev.data(), n * sizeof(serve_req), hipMemcpyHostToDevice) != hipSuccess)
    return GAL_ERR_HIP;
  s->n = n;
  s->ran = false;
  return GAL_OK;
}

static double nearest_rank(std::vector<double>& v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)std::ceil(q * (double)v.size());
  return v[k == 0 ? 0 : k - 1];
}

// q-quantile of a latency histogram, linear inside the bucket
static double hist_quantile(const std::vector<unsigned long long>& hist, double q, double tick_ns) {
  unsigned long long total = 0;
  for (auto c : hist) total += c;
  if (total == 0) return 0.0;
  const double target = q * (double)total;
  double cum = 0.0;
  for (uint32_t b = 0; b < kHistBuckets; b++) {
    if (hist[b] == 0) continue;
    if (cum + (double)hist[b] >= target) {
      double lo, width;
      hist_range(b, lo, width);
      return (lo + width * (target - cum) / (double)hist[b]) * tick_ns;
    }
    cum += (double)hist[b];
  }
  return 0.0;
}

int gal_serve_run(gal_serve* s, uint32_t slots, uint32_t run, uint64_t headroom, double deadline_s, double window_s,
                  gal_serve_result* out) {
  if (!s || s->n == 0 || slots == 0 || run == 0 || run > kServeMaxRun || !(deadline_s > 0.0) || deadline_s > 3600.0 ||
      !(window_s >= 0.0))
    return GAL_ERR_INVALID;
  gal_workload* h = s->h;
  if (headroom == GAL_SERVE_DEFAULT_HEADROOM) headroom = h->ops->serve_headroom(h, s->page_sizes);
  uint64_t room = 0;
  int rc = gal_free_bytes(h->owner, &room);
  if (rc != GAL_OK) return rc;
  if (s->max_need + headroom > room) return GAL_ERR_INVALID;  // some request could never be admitted
  const unsigned resident = h->ops->serve_resident_blocks();
  unsigned blocks = (unsigned)((slots + 255u) / 256u);
  if (blocks > resident) blocks = resident;
  const uint64_t deadline_ticks = (uint64_t)(deadline_s * 1e9 / h->tick_ns);
  const uint64_t window_ticks = window_s > 0.0 ? (uint64_t)(window_s * 1e9 / h->tick_ns) : ~0ull;
  // deal the waves to the bins in proportion to their work, one at least
  // each while there are enough
  serve_bins bins;
  std::memset(&bins, 0, sizeof(bins));
  bins.nbins = (uint32_t)s->bin_start.size();
  const uint32_t waves = blocks * 4;
  {
    const uint32_t nb = bins.nbins;
    double total = 0.0;
    for (double w : s->bin_work) total += w;
    std::vector<uint32_t> share(nb, 0);
    std::vector<std::pair<double, uint32_t>> frac;
    uint32_t dealt = 0;
    for (uint32_t b = 0; b < nb; b++) {
      const double exact = total > 0 ? waves * s->bin_work[b] / total : 0.0;
      share[b] = (uint32_t)exact;
      dealt += share[b];
      frac.push_back({exact - share[b], b});
    }
    std::sort(frac.begin(), frac.end(), [](auto& x, auto& y) { return x.first > y.first; });
    for (uint32_t k = 0; dealt < waves && k < nb; k++, dealt++) share[frac[k].second]++;
    if (waves >= nb)
      for (uint32_t b = 0; b < nb; b++)
        if (share[b] == 0) {
          uint32_t big = (uint32_t)(std::max_element(share.begin(), share.end()) - share.begin());
          share[big]--;
          share[b]++;
        }
    uint32_t acc = 0;
    for (uint32_t b = 0; b < nb; b++) {
      bins.start[b] = s->bin_start[b];
      bins.count[b] = s->bin_count[b];
      bins.wave_start[b] = acc;
      acc += share[b];
    }
    bins.wave_start[nb] = acc;
  }

  gal_run_result base;
  std::memset(&base, 0, sizeof(base));
  rc = timed_run(h, &base, [&](launch_timer& t) {
    if (hipMemsetAsync(s->d_state, 0, sizeof(serve_state), h->stream) != hipSuccess ||
        hipMemsetAsync(s->d_hist, 0, kHistBuckets * sizeof(unsigned long long), h->stream) != hipSuccess)
      return GAL_ERR_HIP;
    if (!t.begin()) return GAL_ERR_INVALID;
    h->ops->serve_launch(s, blocks, run, headroom, deadline_ticks, bins, window_ticks);
    t.end(0);
    return GAL_OK;
  });
  if (rc != GAL_OK) return rc;
  s->ran = true;

  serve_state st;
  std::vector<unsigned long long> hist(kHistBuckets);
  std::vector<uint64_t> admit(s->n), done(s->n);
  if (hipMemcpy(&st, s->d_state, sizeof(st), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hist.data(), s->d_hist, kHistBuckets * 8, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(admit.data(), s->d_admit, s->n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(done.data(), s->d_done, s->n * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return GAL_ERR_HIP;
  if (!out) return GAL_OK;
  gal_serve_result& r = *out;
  std::memset(&r, 0, sizeof(r));
  r.requests_done = st.done;
  r.requests_aborted = st.aborted;
  r.page_mallocs = st.page_mallocs;
  r.page_frees = st.page_frees;
  r.page_retries = st.page_retries;
  r.admission_refusals = st.refused;
  r.peak_committed = st.peak_committed.v;
  r.window_pages = st.window_pages;
  r.window_ms = window_s > 0.0 ? window_s * 1e3 : 0.0;
  r.first_dry_ms = st.first_dry == ~0ull ? -1.0 : (double)st.first_dry * h->tick_ns * 1e-6;
  r.alloc_errors = base.alloc_errors;
  r.slots = (uint64_t)blocks * 256;
  r.run = run;
  r.headroom = headroom;
  r.kernel_ms = base.malloc_kernel_ms;
  const double tick_ns = h->tick_ns;
  std::vector<double> queue, latency;
  queue.reserve(s->n);
  latency.reserve(s->n);
  uint64_t last = 0;
  for (uint64_t i = 0; i < s->n; i++) {
    if (done[i] == ~0ull || admit[i] == ~0ull) continue;
    if (done[i] > last) last = done[i];
    const double arr = (double)s->arrival_ns[i];
    queue.push_back(std::max(0.0, (double)admit[i] * tick_ns - arr));
    latency.push_back(std::max(0.0, (double)done[i] * tick_ns - arr));
  }
  r.makespan_ms = (double)last * tick_ns * 1e-6;
  // most sequences in flight at once: at each admission, the requests
  // admitted so far minus those finished strictly before it
  {
    std::vector<uint64_t> a_t, d_t;
    a_t.reserve(s->n);
    d_t.reserve(s->n);
    for (uint64_t i = 0; i < s->n; i++) {
      if (admit[i] == ~0ull) continue;
      a_t.push_back(admit[i]);
      d_t.push_back(done[i] == ~0ull ? ~0ull : done[i]);
    }
    std::sort(a_t.begin(), a_t.end());
    std::sort(d_t.begin(), d_t.end());
    uint64_t peak = 0;
    size_t finished = 0;
    for (size_t k = 0; k < a_t.size(); k++) {
      while (finished < d_t.size() && d_t[finished] < a_t[k]) finished++;
      peak = std::max<uint64_t>(peak, (uint64_t)(k + 1 - finished));
    }
    r.peak_active = peak;
  }
  r.page_malloc_p50_ns = hist_quantile(hist, 0.50, tick_ns);
  r.page_malloc_p99_ns = hist_quantile(hist, 0.99, tick_ns);
  for (uint32_t b = kHistBuckets; b-- > 0;)
    if (hist[b]) {
      double lo, width;
      hist_range(b, lo, width);
      r.page_malloc_max_ns = (lo + width) * tick_ns;
      break;
    }
  r.queue_p50_ns = nearest_rank(queue, 0.50);
  r.queue_p99_ns = nearest_rank(queue, 0.99);
  r.latency_p50_ns = nearest_rank(latency, 0.50);
  r.latency_p99_ns = nearest_rank(latency, 0.99);
  return GAL_OK;
}

int gal_serve_results(gal_serve* s, uint64_t* checksums, uint64_t* admit_ns, uint64_t* done_ns, uint64_t cap) {
  if (!s || !s->ran || cap < s->n) return GAL_ERR_INVALID;
  if (hipStreamSynchronize(s->h->stream) != hipSuccess) return GAL_ERR_HIP;
  const double tick_ns = s->h->tick_ns;
  std::vector<uint64_t> tmp(s->n);
  // device order -> the caller's order; times from ticks to ns
  auto fetch = [&](uint64_t* dst, const uint64_t* src, bool ticks) {
    if (!dst) return true;
    if (hipMemcpy(tmp.data(), src, s->n * 8, hipMemcpyDeviceToHost) != hipSuccess) return false;
    for (uint64_t p = 0; p < s->n; p++) {
      const uint64_t v = tmp[p];
      dst[s->order[p]] = !ticks || v == ~0ull ? v : (uint64_t)((double)v * tick_ns);
    }
    return true;
  };
  if (!fetch(checksums, s->d_checksums, false) || !fetch(admit_ns, s->d_admit, true) ||
      !fetch(done_ns, s->d_done, true))
    return GAL_ERR_HIP;
  return GAL_OK;
}

int gal_atomic_rate(uint32_t mode, uint64_t threads, uint32_t iters, uint32_t addresses, double* ops_per_s) {
  if (!ops_per_s || threads == 0 || iters == 0 || mode > GAL_ATOMIC_WAVE_HOT) return GAL_ERR_INVALID;
  if (mode == GAL_ATOMIC_WAVE_HOT && addresses == 0) return GAL_ERR_INVALID;
  const uint64_t words = mode == GAL_ATOMIC_DISTINCT ? threads : (uint64_t)addresses * 32;
  uint32_t* d = nullptr;
  uint32_t* sink = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GAL_ERR_HIP;
  if (hipMalloc(&d, words * sizeof(uint32_t)) == hipSuccess && hipMalloc(&sink, sizeof(uint32_t)) == hipSuccess &&
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
      hipEventCreate(&e1) == hipSuccess && hipMemsetAsync(d, 0, words * sizeof(uint32_t), st) == hipSuccess) {
    const unsigned grid = blocks_for(threads, 256);
    atomic_rate_kernel<<<grid, 256, 0, st>>>(d, threads, 1, mode, addresses, sink);  // warm-up
    (void)hipEventRecord(e0, st);
    atomic_rate_kernel<<<grid, 256, 0, st>>>(d, threads, iters, mode, addresses, sink);
    (void)hipEventRecord(e1, st);
    float ms = 0.f;
    if (hipGetLastError() == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
        hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.f) {
      const double ops =
          mode == GAL_ATOMIC_DISTINCT ? (double)threads * iters : (double)((threads + 63) / 64) * iters;
      *ops_per_s = ops / (ms * 1e-3);
      rc = GAL_OK;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (sink) (void)hipFree(sink);
  if (d) (void)hipFree(d);
  return rc;
}

int gal_global_fused(uint64_t n_threads, uint64_t size, uint64_t* on_device, uint64_t* on_host, uint64_t* misses,
                     uint64_t* tag_errors) {
  if (n_threads == 0) return GAL_ERR_INVALID;
  global_call g(true);
  if (g.rc != GAL_OK) return g.rc;
  gal_workload* w = g.w;
  uint64_t** ptrs = nullptr;
  unsigned long long* host_count = nullptr;
  if (hipMalloc(&ptrs, n_threads * 8) != hipSuccess || hipMalloc(&host_count, 8) != hipSuccess ||
      hipMemset(host_count, 0, 8) != hipSuccess) {
    if (ptrs) (void)hipFree(ptrs);
    if (host_count) (void)hipFree(host_count);
    return GAL_ERR_OOM;
  }
  gal_run_result a, b;
  int rc = timed_run(w, &a, [&](launch_timer& t) {
    if (!t.begin()) return GAL_ERR_INVALID;
    global_fused_malloc_kernel<<<blocks_for(n_threads, 256), 256, 0, w->stream>>>(n_threads, size, ptrs,
                                                                                 w->d_counters, host_count);
    t.end(0);
    return GAL_OK;
  });
  if (rc == GAL_OK)
    rc = timed_run(w, &b, [&](launch_timer& t) {
      if (!t.begin()) return GAL_ERR_INVALID;
      global_fused_free_kernel<<<blocks_for(n_threads, 256), 256, 0, w->stream>>>(n_threads, ptrs, w->d_counters);
      t.end(1);
      return GAL_OK;
    });
  unsigned long long hc = 0;
  if (rc == GAL_OK && hipMemcpy(&hc, host_count, 8, hipMemcpyDeviceToHost) != hipSuccess) rc = GAL_ERR_HIP;
  (void)hipFree(ptrs);
  (void)hipFree(host_count);
  if (rc != GAL_OK) return rc;
  if (on_device) *on_device = a.mallocs - hc;
  if (on_host) *on_host = hc;
  if (misses) *misses = a.misses;
  if (tag_errors) *tag_errors = a.tag_errors + b.tag_errors + (b.frees == a.mallocs ? 0 : 1);
  return GAL_OK;
}

int gal_global_poison(uint64_t size, int32_t* codes, uint32_t cap) {
  if (!codes || cap < 7 || size == 0) return GAL_ERR_INVALID;
  global_call g(false);
  if (g.rc != GAL_OK) return g.rc;
  int* d = nullptr;
  if (hipMalloc(&d, 7 * sizeof(int)) != hipSuccess) return GAL_ERR_OOM;
  global_poison_kernel<<<1, 64, 0, g.w->stream>>>(size, d);
  int rc = hipStreamSynchronize(g.w->stream) == hipSuccess &&
                   hipMemcpy(codes, d, 7 * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess
               ? GAL_OK
               : GAL_ERR_HIP;
  (void)hipFree(d);
  return rc;
}

int gal_global_plain(uint64_t n_threads, uint64_t size, uint64_t seed, uint32_t verify, gal_run_result* out) {
  if (n_threads == 0 || size == 0 || verify > GAL_VERIFY_ATOMIC) return GAL_ERR_INVALID;
  global_call g(false);
  if (g.rc != GAL_OK) return g.rc;
  gal_workload* w = g.w;
  return timed_run(w, out, [&](launch_timer& t) {
    if (!t.begin()) return GAL_ERR_INVALID;
    global_plain_kernel<<<blocks_for(n_threads, kClaimBlock), kClaimBlock, 0, w->stream>>>(
        n_threads, size, seed, verify, w->d_counters);
    t.end(0);
    return GAL_OK;
  });
}

int gal_global_smoke(uint64_t n_threads, uint64_t size, uint64_t* misses, uint64_t* tag_errors) {
  gal_run_result r;
  std::memset(&r, 0, sizeof(r));
  const int rc = gal_global_plain(n_threads, size, 7, GAL_VERIFY_ATOMIC, &r);
  if (misses) *misses = r.misses;
  if (tag_errors) *tag_errors = r.tag_errors;
  return rc;
}

}  // extern "C"
