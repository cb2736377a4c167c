// device_utils.hpp — wave64, atomic and bit primitives for gfx950.
//
// Every piece of allocator metadata is shared by all 8 XCDs, whose L2 caches
// are not coherent with each other. Metadata is therefore only touched through
// agent-scope atomics: read-modify-writes, and loads that the compiler emits
// with the sc1 (coherent) bit. Plain loads are never used on metadata.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gallatin {

namespace utils {

constexpr uint64_t kFail = ~0ull;

// ---- bit helpers (reference alloc_utils.cuh:382-398, :440-453; veb.cuh:66-69)
__host__ __device__ constexpr inline uint64_t bitmask(uint64_t n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }

__host__ __device__ inline int popc64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __popcll(x);
#else
  return __builtin_popcountll(x);
#endif
}

// 1-based index of the lowest set bit, 0 for zero.
__host__ __device__ inline int ffs64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __ffsll((unsigned long long)x);
#else
  return x ? __builtin_ctzll(x) + 1 : 0;
#endif
}

__host__ __device__ inline int clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return x ? __clzll((long long)x) : 64;
#else
  return x ? __builtin_clzll(x) : 64;
#endif
}

// ceil(log2(x)) for x >= 1; 0 for x == 0.
__host__ __device__ inline int get_first_bit_bigger(uint64_t x) { return 63 - clz64(x) + (popc64(x) != 1); }

// number of trailing one bits of `bits`
__host__ __device__ inline int trailing_ones(uint64_t bits) { return ffs64(~bits) == 0 ? 64 : ffs64(~bits) - 1; }

// MurmurHash64A (Austin Appleby) of one 8-byte key: the hash that seeds the
// hashed segment search (reference murmurhash.cuh:20-61 with len == 8).
__host__ __device__ inline uint64_t murmur64a(uint64_t key, uint64_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ull;
  uint64_t h = seed ^ (8ull * m);
  uint64_t k = key * m;
  k ^= k >> 47;
  k *= m;
  h ^= k;
  h *= m;
  h ^= h >> 47;
  h *= m;
  h ^= h >> 47;
  return h;
}

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace utils

#if defined(__HIPCC__)

namespace atom {

// the value argument never takes part in deduction: `add(&u64, 1ull)` works
template <class T>
struct same {
  using type = T;
};
template <class T>
using same_t = typename same<T>::type;

// Agent-scope atomics. `relaxed` is enough wherever the returned value alone
// decides what happens next; the few publish/consume points use fences.
template <class T>
__device__ inline T load(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline void store(T* p, same_t<T> v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline T add(T* p, same_t<T> v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// release: the caller's earlier stores (to any address) are performed before
// this read-modify-write is.
template <class T>
__device__ inline T add_release(T* p, same_t<T> v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline T sub(T* p, same_t<T> v) {
  return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline T fetch_or(T* p, same_t<T> v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline T fetch_and(T* p, same_t<T> v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline T exch(T* p, same_t<T> v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The word's current value, read with a read-modify-write that changes
// nothing: it is performed where the word's atomics are (beyond the XCD
// L2s), so it sees every atomic performed on the word before it. For the
// few reads whose result nothing else validates. The zero it ORs in is
// opaque to the compiler, which otherwise turns an OR with a known zero into
// a plain agent-scope load.
template <class T>
__device__ inline T read_current(T* p) {
  uint32_t zero;
  asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
  return __hip_atomic_fetch_or(p, (T)zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// returns the value found (== expected on success), like CUDA atomicCAS
template <class T>
__device__ inline T cas(T* p, same_t<T> expected, same_t<T> desired) {
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expected;
}
// Agent-scope fences (fence() is the reference's __threadfence()). On
// MI355X the release half writes back the calling XCD's whole L2 and the
// acquire half invalidates it: they are for ordinary (non-atomic) data, i.e.
// the payload a free hands back. Metadata is ordered with order() below.
__device__ inline void fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent"); }
__device__ inline void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ inline void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
// Orders the calling lane's metadata accesses, which are all agent-scope
// atomics, loads and stores (performed coherently across the XCDs): waits
// until every earlier vector memory access has completed, with none of the
// L2 write-back or invalidation that an agent-scope fence adds for ordinary
// (non-atomic) data. The signal fences keep the compiler from moving memory
// accesses across the wait.
__device__ inline void order() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0), expcnt and lgkmcnt not waited for
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

}  // namespace atom

namespace wave {

// wave64 helpers. "group" masks are 64-bit lane masks over the wavefront.
__device__ inline uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ inline uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
// The exec mask at this point of the program. A ballot of `true` has no
// operands, so the optimizer may evaluate it earlier (outside the branch or
// loop that narrowed the mask); a volatile read of exec stays where it is.
// Every ballot taken inside a narrowed region is ANDed with such a mask.
__device__ inline uint64_t active_mask() {
  uint64_t m;
  asm volatile("s_mov_b64 %0, exec" : "=s"(m));
  return m;
}
// lanes strictly below this one
__device__ inline uint64_t lanemask_lt() {
  uint32_t l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}
__device__ inline uint32_t popc(uint64_t m) { return (uint32_t)__popcll(m); }
__device__ inline int first_lane(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }
__device__ inline int last_lane(uint64_t m) { return 63 - __clzll((long long)m); }
__device__ inline uint32_t readlane(uint32_t v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ inline uint64_t readlane64(uint64_t v, int lane) {
  uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, lane);
  uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// value of `v` in lane `src` (per-lane source; the source lane must be active)
__device__ inline uint32_t shfl(uint32_t v, int src) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}
__device__ inline uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = shfl((uint32_t)v, src);
  const uint32_t hi = shfl((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// Partition the lanes of `lanes` by equal `key`; returns this lane's group
// (a lane mask). A uniform loop with one iteration per distinct key, using
// only register and scalar work.
__device__ inline uint64_t match_groups(uint64_t lanes, uint64_t key) {
  uint64_t mine = 0;
  const uint32_t me = lane_id();
  for (uint64_t todo = lanes; todo;) {
    const uint64_t k = readlane64(key, first_lane(todo));
    const uint64_t g = ballot(key == k) & todo;
    mine = ((g >> me) & 1ull) ? g : mine;
    todo &= ~g;
  }
  return mine;
}

// Exclusive prefix sum and total of a per-lane count over the lanes of
// `group`, in lane order, built from ballots of the count's bits so that no
// lane outside the group is ever read. Counts are < 2^bits.
// `group` may differ per lane (each lane passes its own group's mask), so
// several groups are scanned at once.
__device__ inline void group_scan(uint64_t group, uint32_t count, int bits, uint32_t& prefix, uint32_t& total) {
  uint64_t below = group & lanemask_lt();
  uint32_t p = 0, t = 0;
  for (int b = 0; b < bits; b++) {
    uint64_t m = ballot((count >> b) & 1u) & group;
    p += popc(m & below) << b;
    t += popc(m) << b;
  }
  prefix = p;
  total = t;
}

// cheap back-off inside bounded retry loops
__device__ inline void backoff(int iteration) {
  if (iteration > 2) __builtin_amdgcn_s_sleep(2);
}

// A value that differs between the waves resident on the GPU at one time:
// the hardware wave slot (HW_ID: wave, SIMD, CU, shader array and engine
// ids) and the XCD (XCC_ID), mixed with the wall clock.
__device__ inline uint64_t slot_salt() {
  const uint32_t hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);    // HW_REG_HW_ID, 32 bits
  const uint32_t xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20);   // HW_REG_XCC_ID, bits 3:0
  return ((uint64_t)(hw & 0xFFFFu) | ((uint64_t)(xcc & 0xFu) << 16) | ((uint64_t)lane_id() << 20)) ^
         (__builtin_amdgcn_s_memrealtime() << 26);
}

}  // namespace wave

#endif  // __HIPCC__

}  // namespace gallatin
