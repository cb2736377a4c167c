// main.hip — the other translation unit of the drop-in example: kernels that
// call device functions of nodes.hip, and the host side of the reference's
// API (init_global_allocator, print_global_stats, free_global_allocator,
// Gallatin<...>::generate_on_device / free_on_device). Prints "dropin ok"
// and exits 0 when every allocation was served and read back intact.
#include <gallatin/allocators/gallatin.cuh>
#include <gallatin/allocators/global_allocator.cuh>

#include <stdint.h>
#include <stdio.h>

using namespace gallatin::allocators;

__device__ uint64_t* make_node(uint64_t value);
__device__ bool drop_node(uint64_t* p, uint64_t value);

// every thread: a node from the global allocator, then a lane-dependent size
// from the allocator passed in (the reference's non-global usage)
template <class A>
__global__ void roundtrip(A* local, uint64_t n, unsigned long long* counts) {
  const uint64_t tid = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  if (tid >= n) return;
  uint64_t* node = make_node(tid * 3 + 1);
  if (!node) {
    atomicAdd(&counts[0], 1ull);
  } else if (!drop_node(node, tid * 3 + 1)) {
    atomicAdd(&counts[1], 1ull);
  }
  const uint64_t size = 16ull << (tid % 9);
  uint64_t* p = (uint64_t*)local->malloc(size);
  if (!p) {
    atomicAdd(&counts[0], 1ull);
    return;
  }
  p[0] = tid;
  if (p[0] != tid) atomicAdd(&counts[1], 1ull);
  local->free(p);
}

int main() {
  using local_type = Gallatin<16ull * 1024 * 1024, 16ull, 4096ull>;
  // 4 GiB each: 2^20 threads per launch claim and free at once, and a 1 GiB
  // pool (64 segments for nine classes) could run a class out of blocks for
  // a moment while its spent blocks wait for their last frees
  init_global_allocator(4ull << 30, 42, false);
  local_type* local = local_type::generate_on_device(4ull << 30, 43, false);
  unsigned long long* counts = nullptr;
  if (!local || hipMalloc(&counts, 16) != hipSuccess || hipMemset(counts, 0, 16) != hipSuccess) {
    printf("dropin: setup failed\n");
    return 1;
  }
  const uint64_t n = 1ull << 20;
  for (int rep = 0; rep < 3; rep++) roundtrip<<<(unsigned)(n / 256), 256>>>(local, n, counts);
  unsigned long long host[2] = {1, 1};
  if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(host, counts, 16, hipMemcpyDeviceToHost) != hipSuccess) {
    printf("dropin: kernel failed\n");
    return 1;
  }
  print_global_stats();
  local->print_info();
  local_type::free_on_device(local);
  free_global_allocator();
  (void)hipFree(counts);
  printf("dropin: misses %llu damaged %llu\n", host[0], host[1]);
  if (host[0] != 0 || host[1] != 0) return 1;
  printf("dropin ok\n");
  return 0;
}
