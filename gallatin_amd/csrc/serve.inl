// serve.inl — the serving loop (part of capi.hip).
//
// One persistent launch serves a stream of sequence requests. Every lane of
// the grid is a sequence slot with a small state machine:
//
//   idle     -> takes the next request (one ticket counter; a wave takes the
//               tickets of all its idle lanes with one atomic)
//   waiting  -> admitted once the request has arrived (device wall clock)
//               and the pool has room for all of its pages
//   growing  -> one page per loop iteration: malloc, write the page header
//               (link to the previous page + content words), read the
//               previous page's content back into the checksum
//   draining -> one page per loop iteration: read the content back, free
//   (finished: checksum and times written, back to idle)
//
// All growing lanes of a wave call malloc in the same call, so lanes of one
// page class claim as one group (one atomic); draining lanes free together.
// The pages of a sequence are chained through their headers, so the loop
// keeps no page table.
//
// Admission keeps   committed + need + headroom <= free segments * segment
// size, where committed is the page bytes of admitted sequences not yet
// allocated. A page takes at most its own bytes of fresh segment memory
// beyond what the headroom holds (blocks that pinned slots hold partly
// filled, segments attached to a class and not yet used up), and memory
// only returns to the free segments, so an admitted sequence always finds
// its pages: the loop cannot deadlock on memory. Each lane returns the bytes
// it allocated to `committed` every kFlushEvery iterations, so committed
// over-counts by at most that many pages per lane (conservative).

constexpr uint32_t kSeqIdle = 0, kSeqWaiting = 1, kSeqGrowing = 2, kSeqDraining = 3, kSeqExhausted = 4;
constexpr uint32_t kHistBuckets = 512;  // 16 per octave of wall-clock ticks
constexpr uint32_t kFlushEvery = 16;

struct serve_req {
  uint64_t arrival;  // wall-clock ticks after the start
  uint64_t id;
  uint32_t pages;
  uint32_t page_bytes;
};

struct serve_state {
  unsigned long long cursor;     // next ticket
  unsigned long long committed;  // admitted page bytes not yet allocated
  unsigned long long active;     // sequences admitted and not finished
  unsigned long long peak_active;
  unsigned long long peak_committed;
  unsigned long long page_mallocs, page_frees, page_retries;
  unsigned long long done, aborted, refused;
  unsigned long long t0;  // start of the loop (wall-clock ticks)
};

// 16 buckets per octave; exact below 32 ticks
__host__ __device__ inline uint32_t hist_bucket(uint64_t t) {
  if (t < 16) return (uint32_t)t;
  const int e = 63 - gu::clz64(t);  // >= 4
  const uint32_t b = (uint32_t)(e - 3) * 16u + ((uint32_t)(t >> (e - 4)) & 15u);
  return b < kHistBuckets ? b : kHistBuckets - 1;
}
// lowest tick count of bucket b and the bucket's width
inline void hist_range(uint32_t b, double& lo, double& width) {
  if (b < 16) {
    lo = b;
    width = 1;
    return;
  }
  const int shift = (int)(b / 16) - 1;
  lo = (double)((16u + b % 16u) * (1ull << shift));
  width = (double)(1ull << shift);
}

// sum over the 64 lanes (all lanes active)
__device__ inline uint64_t wave_sum64(uint64_t x) {
  const uint32_t lane = wave::lane_id();
  for (uint32_t d = 1; d < 64; d <<= 1) x += wave::shfl64(x, (int)(lane ^ d));
  return x;
}

__global__ void serve_start_kernel(serve_state* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) atom::store(&st->t0, (unsigned long long)wall_ticks());
}

template <class A>
__global__ void __launch_bounds__(256)
    serve_kernel(A* a, const serve_req* reqs, uint64_t n, serve_state* st, uint64_t* checksums,
                 uint64_t* admit_ticks, uint64_t* done_ticks, unsigned long long* hist_out, uint64_t headroom,
                 uint64_t deadline) {
  __shared__ uint32_t hist[kHistBuckets];
  for (uint32_t i = threadIdx.x; i < kHistBuckets; i += blockDim.x) hist[i] = 0;
  __syncthreads();

  const uint32_t lane = wave::lane_id();
  const uint64_t t0 = atom::load(&st->t0);
  uint32_t* free_segs = &a->control[ga::kCtlFreeSegs];

  uint32_t state = kSeqIdle;
  uint64_t ticket = 0;
  serve_req r{0, 0, 0, 0};
  uint32_t step = 0;            // pages allocated (growing) / left (draining)
  uint64_t* page = nullptr;     // the sequence's last page
  uint64_t h = 0;               // checksum
  uint64_t unflushed = 0;       // allocated bytes not yet taken off `committed`
  bool cut = false;             // aborted by the deadline
  uint32_t nm = 0, nf = 0, nr = 0, nd = 0, na = 0;

  for (uint32_t it = 0;; it++) {
    const uint64_t now = wall_ticks() - t0;
    const bool late = now > deadline;

    // ---- tickets for idle lanes
    const uint64_t idle = wave::ballot(state == kSeqIdle);
    if (idle) {
      const int leader = wave::first_lane(idle);
      uint64_t base = 0;
      if ((int)lane == leader) base = atom::add(&st->cursor, (unsigned long long)wave::popc(idle));
      base = wave::readlane64(base, leader);
      if (state == kSeqIdle) {
        ticket = base + wave::popc(idle & wave::lanemask_lt());
        if (ticket < n) {
          r = reqs[ticket];
          state = kSeqWaiting;
        } else {
          state = kSeqExhausted;
        }
      }
    }

    // ---- deadline: waiting requests are dropped, growing ones drain what they have
    if (late && state == kSeqWaiting) {
      checksums[ticket] = 0;
      admit_ticks[ticket] = ~0ull;
      done_ticks[ticket] = ~0ull;
      na++;
      state = kSeqIdle;
    }
    if (late && state == kSeqGrowing) {
      cut = true;
      unflushed += (uint64_t)(r.pages - step) * r.page_bytes;  // never allocated
      state = kSeqDraining;
    }

    // ---- admission of arrived requests, all of the wave's at once when they fit
    const bool arrived = state == kSeqWaiting && r.arrival <= now;
    const uint64_t want = wave::ballot(arrived);
    uint64_t granted = 0;
    if (want) {
      const uint64_t need = arrived ? (uint64_t)r.pages * r.page_bytes : 0ull;
      const uint64_t total = wave_sum64(need);
      const int leader = wave::first_lane(want);
      if ((int)lane == leader) {
        const uint64_t room = (uint64_t)atom::load(free_segs) * A::kBytesPerSegment;
        uint64_t c = atom::load(&st->committed);
        // all arrived lanes, else the leader alone
        for (int pass = 0; pass < 2 && !granted; pass++) {
          const uint64_t ask = pass == 0 ? total : need;
          if (pass == 1 && ask == total) break;
          if (c + ask + headroom > room) continue;
          c = atom::add(&st->committed, (unsigned long long)ask);
          if (c + ask + headroom <= room) {
            granted = pass == 0 ? want : (1ull << lane);
            __hip_atomic_fetch_max(&st->peak_committed, (unsigned long long)(c + ask), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
          } else {
            atom::sub(&st->committed, (unsigned long long)ask);
          }
        }
        if (!granted) atom::add(&st->refused, 1ull);
        if (granted) {
          const unsigned long long k = wave::popc(granted);
          const unsigned long long before = atom::add(&st->active, k);
          __hip_atomic_fetch_max(&st->peak_active, before + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      granted = wave::readlane64(granted, leader);
      if ((granted >> lane) & 1ull) {
        state = kSeqGrowing;
        step = 0;
        page = nullptr;
        h = r.id;
        cut = false;
        admit_ticks[ticket] = now;
      }
    }

    // ---- growing lanes: one page each
    const bool grow = state == kSeqGrowing;
    const uint64_t growing = wave::ballot(grow);
    if (growing) {
      const uint64_t ta = wall_ticks();
      uint64_t* p = nullptr;
      if (grow) p = (uint64_t*)a->malloc(r.page_bytes);
      const uint64_t tb = wall_ticks();
      if ((int)lane == wave::first_lane(growing)) atomicAdd(&hist[hist_bucket(tb - ta)], wave::popc(growing));
      if (grow) {
        if (p) {
          const uint32_t w = gw::page_words(r.page_bytes);
          p[0] = (uint64_t)page;
          for (uint32_t k = 1; k < w; k++) p[k] = gw::page_word(r.id, step, k);
          if (page)
            for (uint32_t k = 1; k < w; k++) h = gw::fold(h, page[k]);
          page = p;
          step++;
          nm++;
          unflushed += r.page_bytes;
          if (step == r.pages) state = kSeqDraining;
        } else {
          nr++;
        }
      }
    }

    // ---- draining lanes: one page each, newest first
    const bool drain = state == kSeqDraining && page != nullptr;
    const uint64_t draining = wave::ballot(drain);
    if (draining) {
      if (drain) {
        uint64_t* p = page;
        const uint32_t w = gw::page_words(r.page_bytes);
        for (uint32_t k = 1; k < w; k++) h = gw::fold(h, p[k]);
        page = (uint64_t*)p[0];
        a->free(p);
        nf++;
      }
    }
    // finished (also a cut sequence that never got a page)
    const bool finish = state == kSeqDraining && page == nullptr;
    const uint64_t finishing = wave::ballot(finish);
    if (finishing) {
      if (finish) {
        checksums[ticket] = cut ? 0ull : gw::fold(h, r.pages);
        done_ticks[ticket] = cut ? ~0ull : now;
        if (cut) {
          admit_ticks[ticket] = ~0ull;
          na++;
        } else {
          nd++;
        }
        state = kSeqIdle;
      }
      if ((int)lane == wave::first_lane(finishing))
        atom::sub(&st->active, (unsigned long long)wave::popc(finishing));
    }

    // ---- return allocated bytes to the admission budget, leave when done
    const bool all_done = wave::ballot(state != kSeqExhausted) == 0;
    if ((it % kFlushEvery) == kFlushEvery - 1 || all_done) {
      const uint64_t s = wave_sum64(unflushed);
      if (lane == 0 && s) atom::sub(&st->committed, (unsigned long long)s);
      unflushed = 0;
    }
    if (all_done) break;
    if (!growing && !draining && !granted) __builtin_amdgcn_s_sleep(32);
  }

  // per-wave totals
  const uint64_t tm = wave_sum64(nm), tf = wave_sum64(nf), tr = wave_sum64(nr);
  const uint64_t td = wave_sum64(nd), ta = wave_sum64(na);
  if (lane == 0) {
    if (tm) atom::add(&st->page_mallocs, (unsigned long long)tm);
    if (tf) atom::add(&st->page_frees, (unsigned long long)tf);
    if (tr) atom::add(&st->page_retries, (unsigned long long)tr);
    if (td) atom::add(&st->done, (unsigned long long)td);
    if (ta) atom::add(&st->aborted, (unsigned long long)ta);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kHistBuckets; i += blockDim.x)
    if (hist[i]) atom::add(&hist_out[i], (unsigned long long)hist[i]);
}
