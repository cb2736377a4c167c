// slab.inl — the slab hash client (part of workloads.hip): a chaining_table of
// 64-bit keys and values, 7 per slab, whose slabs come from the handle's
// allocator (reference chaining_table.cuh, tests/src/chaining_test.cu).

constexpr int kSlabSlots = 7;

// key i of a key stream: a bijection of (seed + i), so distinct i give
// distinct keys; 0 marks an empty slot and is mapped to ~0
__host__ __device__ inline uint64_t slab_key(uint64_t seed, uint64_t i) {
  const uint64_t k = gu::splitmix64(seed + i);
  return k ? k : ~0ull;
}

struct slab_counters {
  unsigned long long failed, wrong;
};

template <class T>
__global__ void slab_insert_kernel(T* t, uint64_t n, uint64_t key_seed, slab_counters* c) {
  const uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  bool failed = false;
  if (i < n) failed = !t->insert(slab_key(key_seed, i), i);
  const uint64_t m = wave::ballot(failed);
  if (m && (int)wave::lane_id() == wave::first_lane(m)) atom::add(&c->failed, (unsigned long long)wave::popc(m));
}

template <class T>
__global__ void slab_query_kernel(const T* t, uint64_t n, uint64_t key_seed, slab_counters* c) {
  const uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  bool missing = false, wrong = false;
  if (i < n) {
    uint64_t v = 0;
    missing = !t->query(slab_key(key_seed, i), v);
    wrong = !missing && v != i;
  }
  const uint64_t m = wave::ballot(missing), w = wave::ballot(wrong);
  if ((m | w) && (int)wave::lane_id() == wave::first_lane(m | w)) {
    if (m) atom::add(&c->failed, (unsigned long long)wave::popc(m));
    if (w) atom::add(&c->wrong, (unsigned long long)wave::popc(w));
  }
}

// chain lengths: sum and maximum over the buckets
template <class T>
__global__ void slab_chain_kernel(const T* t, unsigned long long* sum_max) {
  const uint64_t b = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  const uint64_t len = b < t->nbuckets ? t->chain_length(b) : 0;
  if (len) {
    atom::add(&sum_max[0], (unsigned long long)len);
    __hip_atomic_fetch_max(&sum_max[1], (unsigned long long)len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <class T>
__global__ void slab_free_kernel(T* t, unsigned long long* freed) {
  const uint64_t b = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  const uint64_t n = b < t->nbuckets ? t->free_chain(b) : 0;
  if (n) atom::add(freed, (unsigned long long)n);
}
