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

}  // namespace allocators
}  // namespace gallatin
