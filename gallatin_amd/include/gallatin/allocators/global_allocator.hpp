// global_allocator.hpp — one process-wide allocator reachable from any kernel
// without passing a pointer (reference include/gallatin/allocators/
// global_allocator.cuh:41-97).
#pragma once

#include "gallatin.hpp"

namespace gallatin {
namespace allocators {

using global_allocator_type = Gallatin<16ull * 1024 * 1024, 16ull, 4096ull>;

__device__ inline global_allocator_type* global_gallatin = nullptr;

// reference global_allocator.cuh:48-56; returns false if construction failed
__host__ inline bool init_global_allocator(uint64_t num_bytes, uint64_t seed, bool print_info = true,
                                           bool running_calloc = false) {
  global_allocator_type* local_copy =
      global_allocator_type::generate_on_device(num_bytes, seed, print_info, running_calloc);
  if (!local_copy) return false;
  if (hipMemcpyToSymbol(HIP_SYMBOL(global_gallatin), &local_copy, sizeof(local_copy)) != hipSuccess) {
    global_allocator_type::free_on_device(local_copy);
    return false;
  }
  return hipDeviceSynchronize() == hipSuccess;
}

__host__ inline global_allocator_type* get_global_allocator() {
  global_allocator_type* local_copy = nullptr;
  if (hipMemcpyFromSymbol(&local_copy, HIP_SYMBOL(global_gallatin), sizeof(local_copy)) != hipSuccess) return nullptr;
  return local_copy;
}

// reference global_allocator.cuh:59-70
__host__ inline void free_global_allocator() {
  global_allocator_type* local_copy = get_global_allocator();
  if (!local_copy) return;
  global_allocator_type::free_on_device(local_copy);
  global_allocator_type* null_copy = nullptr;
  (void)hipMemcpyToSymbol(HIP_SYMBOL(global_gallatin), &null_copy, sizeof(null_copy));
}

// reference global_allocator.cuh:86-97
__host__ inline void print_global_stats() {
  global_allocator_type* local_copy = get_global_allocator();
  if (local_copy) local_copy->print_info();
}

#if defined(__HIPCC__)
// reference global_allocator.cuh:72-82
__device__ inline void* global_malloc(uint64_t num_bytes) { return global_gallatin->malloc(num_bytes); }
__device__ inline void global_free(void* ptr) { global_gallatin->free(ptr); }
#endif

// ---- a second global allocator whose pool is mapped host memory
// (reference global_allocator.cuh:98-150)
__device__ inline global_allocator_type* global_host_gallatin = nullptr;

__host__ inline bool init_global_allocator_host(uint64_t num_bytes, uint64_t seed, bool print_info = true,
                                                bool running_calloc = false) {
  global_allocator_type* local_copy =
      global_allocator_type::generate_on_device_host(num_bytes, seed, print_info, running_calloc);
  if (!local_copy) return false;
  if (hipMemcpyToSymbol(HIP_SYMBOL(global_host_gallatin), &local_copy, sizeof(local_copy)) != hipSuccess) {
    global_allocator_type::free_on_device(local_copy);
    return false;
  }
  return hipDeviceSynchronize() == hipSuccess;
}

__host__ inline global_allocator_type* get_global_allocator_host() {
  global_allocator_type* local_copy = nullptr;
  if (hipMemcpyFromSymbol(&local_copy, HIP_SYMBOL(global_host_gallatin), sizeof(local_copy)) != hipSuccess)
    return nullptr;
  return local_copy;
}

__host__ inline void free_global_allocator_host() {
  global_allocator_type* local_copy = get_global_allocator_host();
  if (!local_copy) return;
  global_allocator_type::free_on_device(local_copy);
  global_allocator_type* null_copy = nullptr;
  (void)hipMemcpyToSymbol(HIP_SYMBOL(global_host_gallatin), &null_copy, sizeof(null_copy));
}

__host__ inline void print_global_stats_host() {
  global_allocator_type* local_copy = get_global_allocator_host();
  if (local_copy) local_copy->print_info();
}

// both at once (reference global_allocator.cuh:153-236)
__host__ inline bool init_global_allocator_combined(uint64_t num_bytes, uint64_t host_bytes, uint64_t seed,
                                                    bool print_info = true, bool running_calloc = false) {
  if (!init_global_allocator(num_bytes, seed, print_info, running_calloc)) return false;
  if (!init_global_allocator_host(host_bytes, seed, print_info, running_calloc)) {
    free_global_allocator();
    return false;
  }
  return true;
}

__host__ inline void free_global_allocator_combined() {
  free_global_allocator();
  free_global_allocator_host();
}

__host__ inline void print_global_stats_combined() {
  printf("Device Allocator:\n");
  print_global_stats();
  printf("Host Allocator:\n");
  print_global_stats_host();
}

#if defined(__HIPCC__)
__device__ inline void* global_malloc_host(uint64_t num_bytes) { return global_host_gallatin->malloc(num_bytes); }
__device__ inline void global_free_host(void* ptr) { global_host_gallatin->free(ptr); }

__device__ inline void* global_malloc_combined(uint64_t num_bytes, bool on_host = false) {
  return on_host ? global_malloc_host(num_bytes) : global_malloc(num_bytes);
}
__device__ inline void global_free_combined(void* ptr, bool on_host = false) {
  if (on_host)
    global_free_host(ptr);
  else
    global_free(ptr);
}

// device memory first, host memory when the device pool has none; the free
// goes to whichever pool owns the pointer. Pointers stay where they were
// allocated (reference global_allocator.cuh:199-216).
__device__ inline void* global_malloc_fused(uint64_t num_bytes) {
  void* p = global_malloc(num_bytes);
  return p ? p : global_malloc_host(num_bytes);
}
__device__ inline void global_free_fused(void* ptr) {
  if (global_gallatin->owns_allocation(ptr))
    global_free(ptr);
  else
    global_free_host(ptr);
}

// ---- poison (guard) allocations (reference global_allocator.cuh:240-357)
// Layout of an allocation of S bytes (S = what malloc grants for n + 32):
//   [0, 16)          header {n, extra}
//   [16, 16 + n)     the caller's bytes (the returned pointer is base + 16)
//   [16 + n, S - 16) extra = S - 32 - n bytes, byte i holding (char) i
//   [S - 16, S)      trailer {n, extra}
// S is a multiple of 16, so header and trailer are 16-byte aligned.
__device__ inline void* global_malloc_poison(uint64_t bytes_needed) {
  const uint64_t total = global_allocator_type::allocated_size(bytes_needed + 32);
  const uint64_t extra = total - (bytes_needed + 32);
  char* base = (char*)global_malloc(bytes_needed + 32);
  if (!base) return nullptr;
  char* fill = base + 16 + bytes_needed;
  for (uint64_t i = 0; i < extra; i++) fill[i] = (char)i;
  uint64_t* head = (uint64_t*)base;
  uint64_t* tail = (uint64_t*)(base + total - 16);
  atom::exch(&head[0], bytes_needed);
  atom::exch(&head[1], extra);
  atom::exch(&tail[0], bytes_needed);
  atom::exch(&tail[1], extra);
  return base + 16;
}

// 0 when the guards are intact; otherwise which check failed: 1 the header
// does not describe an allocation malloc would grant (or one that would run
// past the pool), 2 / 3 the trailer's size / extra words differ from the
// header's, 4 a byte of the extra region was overwritten, 5 the pointer is
// not in the pool. The trailer is read only where a consistent header puts
// it, so every read stays inside the pool (the reference reads wherever the
// header points).
__device__ inline int global_check_poison(void* allocation) {
  global_allocator_type* a = global_gallatin;
  if (!allocation) return 5;
  const char* base = (const char*)allocation - 16;
  const uint64_t pool = a->table.num_segments * global_allocator_type::kBytesPerSegment;
  if (base < a->table.memory || (uint64_t)(base - a->table.memory) >= pool) return 5;
  const uint64_t* head = (const uint64_t*)base;
  const uint64_t bytes_needed = atom::load(&head[0]);
  const uint64_t extra = atom::load(&head[1]);
  if (bytes_needed > pool) return 1;
  const uint64_t total = global_allocator_type::allocated_size(bytes_needed + 32);
  if (extra != total - (bytes_needed + 32)) return 1;
  if ((uint64_t)(base - a->table.memory) + total > pool) return 1;
  const uint64_t* tail = (const uint64_t*)(base + total - 16);
  if (atom::load(&tail[0]) != bytes_needed) return 2;
  if (atom::load(&tail[1]) != extra) return 3;
  const char* fill = base + 16 + bytes_needed;
  for (uint64_t i = 0; i < extra; i++)
    if (fill[i] != (char)i) return 4;
  return 0;
}

// checks the guards, then frees; returns the check's result
__device__ inline int global_free_poison(void* allocation) {
  const int check = global_check_poison(allocation);
  if (check != 5) global_free((char*)allocation - 16);
  return check;
}
#endif

}  // namespace allocators
}  // namespace gallatin
