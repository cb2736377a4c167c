// nodes.hip — one translation unit of a user program built against the
// headers alone (the way the reference's README.md:15-25 has users include
// the global allocator). Device functions here allocate and free through
// global_malloc/global_free; the kernels that call them are in main.hip, so
// the program needs relocatable device code (-fgpu-rdc): both units then
// share the one global_gallatin that init_global_allocator sets.
#include <gallatin/allocators/global_allocator.cuh>

#include <stdint.h>

using namespace gallatin::allocators;

// a 16-byte node {value, check}, or nullptr when the allocator has none
__device__ uint64_t* make_node(uint64_t value) {
  if (!global_gallatin) return nullptr;  // init_global_allocator not called (or not seen by this unit)
  uint64_t* p = (uint64_t*)global_malloc(16);
  if (p) {
    p[0] = value;
    p[1] = ~value;
  }
  return p;
}

// true when the node still holds what make_node wrote; frees it
__device__ bool drop_node(uint64_t* p, uint64_t value) {
  const bool ok = p[0] == value && p[1] == ~value;
  global_free(p);
  return ok;
}
