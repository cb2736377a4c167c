/*
 * cpu_baseline.cpp — CPU baselines timed beside the GPU by bench.py.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *  orc_bench_mixed     the oracle replaying the mixed-size workload on one core
 *  orc_bench_sysmalloc glibc malloc/free on all host cores, same size law
 *  orc_bench_sysmalloc_size the same at one fixed size (the plain workload's 16 B)
 *
 * The mixed-size request stream is defined by mixed_key() below; the GPU
 * workload kernels generate the identical stream (request size of thread i in
 * round r, and the coin that decides whether live pointer j is freed after
 * round r).
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "gallatin_oracle.h"

namespace {

uint64_t sm64(uint64_t x) { return orc_splitmix64(x); }

// salt 0: size draw of thread i in round r; salt 1: free coin of live slot i after round r
uint64_t mixed_key(uint64_t seed, uint64_t round, uint64_t i, uint64_t salt) {
  return sm64(seed + sm64((round << 32) ^ i ^ (salt << 60)));
}

uint64_t mixed_size(uint64_t seed, uint64_t round, uint64_t i) { return 16ull << (mixed_key(seed, round, i, 0) % 9); }

double now_s() {
  using namespace std::chrono;
  return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" int orc_bench_mixed(uint64_t pool_bytes, uint64_t seed, uint64_t threads_per_round, uint32_t rounds,
                               uint64_t* ops_done, uint64_t* misses, double* seconds) {
  orc_alloc* a = nullptr;
  int rc = orc_create(16ull << 20, 16, 4096, pool_bytes, seed, &a);
  if (rc != GAL_OK) return rc;
  std::vector<uint64_t> live;  // byte offsets, GAL_NULL_OFFSET when free
  live.reserve(threads_per_round * rounds);
  uint64_t ops = 0, miss = 0;
  double t0 = now_s();
  for (uint32_t r = 0; r < rounds; r++) {
    for (uint64_t i = 0; i < threads_per_round; i++) {
      // a wave of 64 threads shares one CU; 256 CUs
      uint64_t off = orc_malloc(a, mixed_size(seed, r, i), i, (uint32_t)((i / 64) % 256));
      if (off == GAL_NULL_OFFSET) {
        miss++;
      } else {
        ops++;
      }
      live.push_back(off);
    }
    for (uint64_t j = 0; j < live.size(); j++) {
      if (live[j] == GAL_NULL_OFFSET) continue;
      if (mixed_key(seed, r, j, 1) & 1) {
        orc_free(a, live[j]);
        live[j] = GAL_NULL_OFFSET;
        ops++;
      }
    }
  }
  for (uint64_t j = 0; j < live.size(); j++) {
    if (live[j] != GAL_NULL_OFFSET) {
      orc_free(a, live[j]);
      ops++;
    }
  }
  double t1 = now_s();
  orc_destroy(a);
  if (ops_done) *ops_done = ops;
  if (misses) *misses = miss;
  if (seconds) *seconds = t1 - t0;
  return GAL_OK;
}

extern "C" int orc_bench_sysmalloc(uint64_t ops_per_thread, uint32_t threads, uint64_t seed, uint64_t* ops_done,
                                   double* seconds, double* p50_ns, double* p99_ns, uint32_t* threads_used) {
  return orc_bench_sysmalloc_size(ops_per_thread, threads, seed, 0, ops_done, seconds, p50_ns, p99_ns,
                                  threads_used);
}

extern "C" int orc_bench_sysmalloc_size(uint64_t ops_per_thread, uint32_t threads, uint64_t seed, uint64_t fixed_size,
                                        uint64_t* ops_done, double* seconds, double* p50_ns, double* p99_ns,
                                        uint32_t* threads_used) {
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  std::vector<std::vector<float>> lat(threads);
  std::vector<uint64_t> bad(threads, 0);
  std::atomic<int> ready{0};
  std::atomic<bool> go{false};
  auto worker = [&](uint32_t t) {
    lat[t].reserve(ops_per_thread / 64 + 1);
    uint64_t state = sm64(seed ^ ((uint64_t)t << 48));
    ready.fetch_add(1);
    while (!go.load(std::memory_order_acquire)) {
    }
    for (uint64_t i = 0; i < ops_per_thread; i++) {
      state = sm64(state);
      size_t size = fixed_size ? (size_t)fixed_size : (size_t)16 << (state % 9);
      uint64_t tag = state | 1ull;
      bool sample = (i & 63) == 0;
      std::chrono::steady_clock::time_point a;
      if (sample) a = std::chrono::steady_clock::now();
      uint64_t* p = (uint64_t*)std::malloc(size);
      if (sample) {
        auto b = std::chrono::steady_clock::now();
        lat[t].push_back((float)std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count());
      }
      if (!p) {
        bad[t]++;
        continue;
      }
      __atomic_store_n(p, tag, __ATOMIC_RELAXED);
      if (__atomic_load_n(p, __ATOMIC_RELAXED) != tag) bad[t]++;
      std::free(p);
    }
  };
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < threads; t++) pool.emplace_back(worker, t);
  while (ready.load() < (int)threads) {
  }
  double t0 = now_s();
  go.store(true, std::memory_order_release);
  for (auto& th : pool) th.join();
  double t1 = now_s();
  std::vector<float> all;
  for (auto& v : lat) all.insert(all.end(), v.begin(), v.end());
  double p50 = 0, p99 = 0;
  if (!all.empty()) {
    size_t k50 = all.size() / 2, k99 = std::min(all.size() - 1, (size_t)(all.size() * 0.99));
    std::nth_element(all.begin(), all.begin() + k50, all.end());
    p50 = all[k50];
    std::nth_element(all.begin(), all.begin() + k99, all.end());
    p99 = all[k99];
  }
  uint64_t fails = 0;
  for (auto b : bad) fails += b;
  if (ops_done) *ops_done = 2 * (ops_per_thread * threads - fails);
  if (seconds) *seconds = t1 - t0;
  if (p50_ns) *p50_ns = p50;
  if (p99_ns) *p99_ns = p99;
  if (threads_used) *threads_used = threads;
  return fails ? GAL_ERR_STATE : GAL_OK;
}
