// batch.inl — batched service of uneven request sizes (part of workloads.hip).
//
// A batch is an array of request sizes. It is served in three steps, all on
// the device: the requests are binned by the class they claim from (a stable
// counting sort: per-wave bin counts, a per-bin scan over waves, a scatter),
// then one persistent launch walks the binned order 64 requests at a time,
// so that a wave's lanes nearly always ask for one class and claim with one
// atomic instead of one per class. The steps can be captured once into a
// hipGraph and replayed. The per-thread form (each thread mallocs its own
// request, in request order) is kept beside it as the baseline.

constexpr uint32_t kBins = 32;       // slice classes, then big-slice counts, last bin: blocks and segments
constexpr uint32_t kBinChunks = 16;  // 64-request chunks per wave in the binning passes
constexpr uint64_t kBinWaveRequests = 64ull * kBinChunks;

constexpr uint32_t kBatchMalloc = 1u;  // persistent pass mallocs (and tags)
constexpr uint32_t kBatchFree = 2u;    // persistent pass frees (after checking the tag)

template <class A>
__device__ inline uint32_t request_bin(uint64_t size) {
  const uint32_t t = A::get_tree_id_from_size(size);
  return t < kBins - 1 ? t : kBins - 1;
}

__device__ inline uint64_t global_wave() { return ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; }

// sizes 16 << k, k uniform in [0, classes), drawn from (seed, request)
__global__ void batch_sizes_kernel(uint64_t* sizes, uint64_t n, uint64_t seed, uint32_t classes) {
  const uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  if (i < n) sizes[i] = 16ull << (gw::mixed_key(seed, 0x5eed, i, 3) % classes);
}

// counts[w * kBins + b] = requests of bin b among wave w's kBinWaveRequests
template <class A>
__global__ void bin_count_kernel(const uint64_t* sizes, uint64_t n, uint32_t* counts) {
  const uint64_t w = global_wave();
  const uint32_t lane = wave::lane_id();
  const uint64_t base = w * kBinWaveRequests;
  if (base >= n) return;  // whole wave
  uint32_t mine = 0;      // lane b counts bin b
  for (uint32_t c = 0; c < kBinChunks; c++) {
    const uint64_t i = base + c * 64 + lane;
    const bool valid = i < n;
    const uint32_t bin = valid ? request_bin<A>(sizes[i]) : kBins;
    for (uint64_t todo = wave::ballot(valid); todo;) {
      const uint32_t k = wave::readlane(bin, wave::first_lane(todo));
      const uint64_t g = wave::ballot(bin == k) & todo;
      if (lane == k) mine += wave::popc(g);
      todo &= ~g;
    }
  }
  if (lane < kBins) counts[w * kBins + lane] = mine;
}

// inclusive prefix sum over the 64 lanes (all lanes active)
__device__ inline uint32_t wave_inclusive_sum(uint32_t x) {
  const uint32_t lane = wave::lane_id();
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t y = wave::shfl(x, (int)(lane >= d ? lane - d : lane));
    if (lane >= d) x += y;
  }
  return x;
}

// one block per bin: offsets[w * kBins + b] = requests of bin b in waves
// before w; totals[b] = requests of bin b
__global__ void __launch_bounds__(256) bin_scan_kernel(const uint32_t* counts, uint64_t waves, uint32_t* offsets,
                                                       uint32_t* totals) {
  __shared__ uint32_t wave_sums[4];
  const uint32_t bin = blockIdx.x;
  const uint32_t lane = wave::lane_id(), wv = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (uint64_t tile = 0; tile < waves; tile += 256) {
    const uint64_t w = tile + threadIdx.x;
    const uint32_t v = w < waves ? counts[w * kBins + bin] : 0u;
    const uint32_t inc = wave_inclusive_sum(v);
    if (lane == 63) wave_sums[wv] = inc;
    __syncthreads();
    uint32_t before = 0, tile_total = 0;
    for (uint32_t k = 0; k < 4; k++) {
      if (k < wv) before += wave_sums[k];
      tile_total += wave_sums[k];
    }
    if (w < waves) offsets[w * kBins + bin] = carry + before + inc - v;
    carry += tile_total;
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[bin] = carry;
}

// perm[position] = request index, requests ordered by bin, then by index
template <class A>
__global__ void bin_scatter_kernel(const uint64_t* sizes, uint64_t n, const uint32_t* offsets, const uint32_t* totals,
                                   uint32_t* perm) {
  const uint64_t w = global_wave();
  const uint32_t lane = wave::lane_id();
  const uint64_t base = w * kBinWaveRequests;
  if (base >= n) return;
  uint32_t next = 0;  // lane b: next position of bin b for this wave
  if (lane < kBins) {
    uint32_t start = 0;
    for (uint32_t b = 0; b < lane; b++) start += totals[b];
    next = start + offsets[w * kBins + lane];
  }
  for (uint32_t c = 0; c < kBinChunks; c++) {
    const uint64_t i = base + c * 64 + lane;
    const bool valid = i < n;
    const uint32_t bin = valid ? request_bin<A>(sizes[i]) : kBins;
    uint32_t pos = 0;
    for (uint64_t todo = wave::ballot(valid); todo;) {
      const uint32_t k = wave::readlane(bin, wave::first_lane(todo));
      const uint64_t g = wave::ballot(bin == k) & todo;
      const uint32_t s = wave::readlane(next, (int)k);
      if (bin == k) pos = s + wave::popc(g & wave::lanemask_lt());
      if (lane == k) next += wave::popc(g);
      todo &= ~g;
    }
    if (valid) perm[pos] = (uint32_t)i;
  }
}

// Persistent pass over the binned order. A wave serves 64 * run consecutive
// binned requests at a time (wave w takes spans w, w + W, ... of the W waves
// in the grid): lane l takes requests l * run .. l * run + run - 1 of the
// span. When they are all of one slice class (nearly always, the order being
// binned) the lane claims them with one bulk claim (malloc_run_at) and frees
// them with one free_run; otherwise it mallocs and frees them one by one.
// Lanes of one class are split into claim groups of 2^width_log2 lanes, each
// with its own slot hint.
constexpr uint32_t kMaxRun = 64;
// A lane holds at most this many bytes of one run at once: the persistent
// grid keeps thousands of waves on neighbouring spans of one bin, so runs of
// large requests would otherwise hold far more memory than the pool has.
constexpr uint64_t kRunBytes = 64 * 1024;
// A claim that finds memory short is retried this many times, with growing
// sleeps, before its requests count as misses: in a batch most of the
// memory is held by requests of the same launch that are about to be freed.
constexpr int kBatchRetries = 12;

// Tags of up to kTagChunk requests of one bulk claim, checked with their
// memory operations issued together (the per-thread form issues them from
// different threads): in store mode all tags are stored, then all read back;
// on free all are read, then all cleared. Atomic mode exchanges them.
constexpr uint32_t kTagChunk = 4;
__device__ inline uint32_t tags_in(uint64_t* const* p, const uint64_t* tag, uint32_t m, uint32_t verify) {
  uint32_t bad = 0;
  if (verify == GAL_VERIFY_ATOMIC) {
    uint64_t old[kTagChunk];
#pragma unroll
    for (uint32_t j = 0; j < kTagChunk; j++) old[j] = j < m ? atom::exch(p[j], tag[j]) : 0ull;
#pragma unroll
    for (uint32_t j = 0; j < kTagChunk; j++) bad += j < m && old[j] != 0ull;
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kTagChunk; j++)
      if (j < m) *(volatile uint64_t*)p[j] = tag[j];
#pragma unroll
    for (uint32_t j = 0; j < kTagChunk; j++) bad += j < m && *(volatile uint64_t*)p[j] != tag[j];
  }
  return bad;
}
__device__ inline uint32_t tags_out(uint64_t* const* p, const uint64_t* tag, uint32_t m, uint32_t verify) {
  uint32_t bad = 0;
  uint64_t got[kTagChunk];
  if (verify == GAL_VERIFY_ATOMIC) {
#pragma unroll
    for (uint32_t j = 0; j < kTagChunk; j++) got[j] = j < m ? atom::exch(p[j], 0ull) : 0ull;
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kTagChunk; j++) got[j] = j < m ? *(volatile uint64_t*)p[j] : 0ull;
#pragma unroll
    for (uint32_t j = 0; j < kTagChunk; j++)
      if (j < m) *(volatile uint64_t*)p[j] = 0ull;
  }
#pragma unroll
  for (uint32_t j = 0; j < kTagChunk; j++) bad += j < m && got[j] != tag[j];
  return bad;
}

__device__ inline void batch_backoff(int t) {
  if (t < 4)
    __builtin_amdgcn_s_sleep(16);
  else
    __builtin_amdgcn_s_sleep(127);
}

template <class A>
__global__ void __launch_bounds__(256) batch_kernel(A* a, const uint64_t* sizes, const uint32_t* perm, uint64_t n,
                                                    uint64_t** ptrs, uint32_t mode, uint32_t width_log2, uint32_t run,
                                                    uint64_t seed, uint32_t verify, run_counters* c) {
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint32_t lane = wave::lane_id();
  const uint32_t sub = width_log2 >= 6 ? 0u : lane >> width_log2;
  const uint32_t hint = A::cu_hint() + sub * (0x9E37u << 3);
  const uint64_t tid = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  const uint64_t span = 64ull * run;
  const uint64_t spans = (n + span - 1) / span;
  uint32_t nm = 0, nf = 0, nx = 0, nb = 0;  // this lane's counts
  for (uint64_t sp = global_wave(); sp < spans; sp += waves) {
    const uint64_t first = sp * span + (uint64_t)lane * run;
    const uint32_t k = first < n ? (uint32_t)(n - first < run ? n - first : run) : 0u;
    // one bulk claim when the lane's requests share a slice class: the
    // binned order is sorted by bin, so they do exactly when the first and
    // the last share a slice bin
    const uint64_t size0 = k ? sizes[perm[first]] : 0;
    const uint64_t size_last = k > 1 ? sizes[perm[first + k - 1]] : size0;
    const uint64_t stride = k ? A::run_stride(size0) : 0;
    const uint32_t bin0 = request_bin<A>(size0);
    const bool bulk = (mode & kBatchMalloc) && k > 1 && stride != 0 && bin0 < kBins - 1 &&
                      request_bin<A>(size_last) == bin0;
    if (bulk) {
      // sub-runs of at most kRunBytes
      const uint32_t cap = stride >= kRunBytes ? 1u : (uint32_t)(kRunBytes / stride);
#pragma unroll 1
      for (uint32_t r0 = 0; r0 < k; r0 += cap) {
        const uint32_t kk = k - r0 < cap ? k - r0 : cap;
        char* base = nullptr;
#pragma unroll 1
        for (int t = 0; t <= kBatchRetries && !base; t++) {
          if (t) batch_backoff(t);
          base = kk > 1 ? (char*)a->malloc_run_at(size0, kk, tid, hint, sub)
                        : (char*)a->malloc_at(size0, tid, hint, sub);
        }
        if (base) {
#pragma unroll 1
          for (uint32_t c0 = 0; c0 < kk; c0 += kTagChunk) {
            const uint32_t m = kk - c0 < kTagChunk ? kk - c0 : kTagChunk;
            uint32_t req[kTagChunk];
            uint64_t* p[kTagChunk];
            uint64_t tag[kTagChunk];
#pragma unroll
            for (uint32_t j = 0; j < kTagChunk; j++) {
              req[j] = j < m ? perm[first + r0 + c0 + j] : 0u;
              p[j] = (uint64_t*)(base + (uint64_t)(c0 + j) * stride);
              tag[j] = gw::payload_tag(seed, req[j]);
            }
            nb += tags_in(p, tag, m, verify);
            if (mode & kBatchFree) nb += tags_out(p, tag, m, verify);
            if (ptrs)
#pragma unroll
              for (uint32_t j = 0; j < kTagChunk; j++)
                if (j < m) ptrs[req[j]] = (mode & kBatchFree) ? nullptr : p[j];
          }
          nm += kk;
          if (mode & kBatchFree) {
            if (kk > 1)
              a->free_run(base, kk);
            else
              a->free(base);
            nf += kk;
          }
        } else {
          nx += kk;
          if (ptrs)
            for (uint32_t r = 0; r < kk; r++) ptrs[perm[first + r0 + r]] = nullptr;
        }
      }
    } else {
      // one request at a time
#pragma unroll 1
      for (uint32_t r = 0; r < k; r++) {
        const uint32_t req = perm[first + r];
        const uint64_t tag = gw::payload_tag(seed, req);
        uint64_t* p;
        if (mode & kBatchMalloc) {
          p = nullptr;
#pragma unroll 1
          for (int t = 0; t <= kBatchRetries && !p; t++) {
            if (t) batch_backoff(t);
            p = (uint64_t*)a->malloc_at(sizes[req], tid, hint, sub);
          }
          if (p) {
            nm++;
            nb += !tag_in(p, tag, verify);
          } else {
            nx++;
          }
        } else {
          p = ptrs[req];
        }
        if (p && (mode & kBatchFree)) {
          nb += !tag_out(p, tag, verify);
          a->free(p);
          nf++;
          p = nullptr;
        }
        if (ptrs) ptrs[req] = p;
      }
    }
  }
  // wave totals, one atomic per counter and wave
  const uint32_t tm = wave_inclusive_sum(nm), tf = wave_inclusive_sum(nf);
  const uint32_t tx = wave_inclusive_sum(nx), tb = wave_inclusive_sum(nb);
  if (lane == 63) {
    counter_stripe* k = stripe_of(c);
    if (tm) atom::add(&k->mallocs, (unsigned long long)tm);
    if (tf) atom::add(&k->frees, (unsigned long long)tf);
    if (tx) atom::add(&k->misses, (unsigned long long)tx);
    if (tb) atom::add(&k->tag_errors, (unsigned long long)tb);
  }
}

// the per-thread form: thread i mallocs request i, tags it, frees it
template <class A>
__global__ void batch_plain_kernel(A* a, const uint64_t* sizes, uint64_t n, uint64_t seed, uint32_t verify,
                                   run_counters* c) {
  const uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  bool got = false, bad = false, miss = false;
  if (i < n) {
    uint64_t* p = (uint64_t*)a->malloc(sizes[i]);
    if (p) {
      got = true;
      const uint64_t tag = gw::payload_tag(seed, i);
      bad = !tag_in(p, tag, verify);
      bad |= !tag_out(p, tag, verify);
      a->free(p);
    } else {
      miss = true;
    }
  }
  count(c, got, got, miss, bad);
}
