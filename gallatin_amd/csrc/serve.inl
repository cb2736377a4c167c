// serve.inl — the serving loop (part of workloads.hip).
//
// One persistent launch serves a stream of sequence requests. Every lane of
// the grid is a sequence slot with a small state machine:
//
//   idle     -> takes the next request (one ticket counter per request bin;
//               a wave takes the tickets of its idle lanes together, with
//               one atomic, once enough of them are idle)
//   waiting  -> admitted once the request has arrived (device wall clock)
//               and the pool has room for all of its pages
//   growing  -> one page per loop iteration: malloc, write the page header
//               (link to the previous page + a content word), read the
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
// A wave whose admission was refused (the budget is full) skips admission
// for the next kAdmitBackoffMin iterations, doubling after each further
// refusal up to kAdmitBackoffMax and reset by a grant. Without it every wave
// holding a waiting request read the budget words each iteration, and its
// growing lanes waited for those reads: with 65536 slots over long
// sequences, where the budget is full most of the time, the loop ran at
// 4.7e9 page mallocs+frees/s instead of 8.1e9.
constexpr uint32_t kAdmitBackoffMin = 8;
constexpr uint32_t kAdmitBackoffMax = 64;
// Lanes holding a drained claim before the wave frees them together. The
// loop is bound by the memory accesses its waves have in flight, and each
// free call adds its own (fence, counters): 32 instead of 16 raised the
// short-sequence whole run by 4-6 % at the same page-malloc latency; 48 and
// 64 measured the same as 32.
constexpr uint32_t kFreeBatch = 32;
// Idle lanes take their tickets together: once kTicketBatch of them are
// idle, once the oldest has waited kTicketWait iterations, or when no lane
// of the wave is busy. Taking a ticket, admitting and claiming put dependent
// atomics and loads on the wave's critical path; with short sequences some
// lane finished in almost every iteration and every iteration paid for them.
constexpr uint32_t kTicketBatch = 16;
constexpr uint32_t kTicketWait = 16;

struct serve_req {
  uint64_t arrival;  // wall-clock ticks after the start
  uint64_t id;
  uint32_t pages;
  uint32_t page_bytes;
};

// The counters every wave updates while it runs get a 128-byte line each:
// returned atomics to one line are performed one after another.
struct alignas(128) serve_line {
  unsigned long long v;
};
struct serve_state {
  serve_line cursors[32];     // next ticket of each request bin
  serve_line committed;       // admitted page bytes not yet allocated
  serve_line peak_committed;  // largest value `committed` reached at an admission
  // written once per wave, at its end
  unsigned long long page_mallocs, page_frees, page_retries;
  unsigned long long done, aborted, refused;
  unsigned long long window_pages;  // pages written plus drained before the window's end
  unsigned long long first_dry;     // first time a wave found every request queue empty (~0: never)
  unsigned long long t0;  // start of the loop (wall-clock ticks)
};

// Requests are binned by the class their pages claim from: each bin is a
// queue in arrival order with a ticket counter of its own (the requests are
// stored bin after bin, so a ticket is the request's position and taking one
// is a single dependent load after the counter's add), and each wave
// serves one bin (waves dealt to bins in proportion to their work), so the
// growing lanes of a wave claim as one group and its draining lanes return
// pages of the same blocks. A wave whose bin has handed out every ticket
// moves to the next bin; once it has found every bin empty it leaves.
constexpr uint32_t kServeBins = 32;
struct serve_bins {
  uint32_t start[kServeBins];       // first position of each bin among the requests
  uint32_t count[kServeBins];       // requests of each bin
  uint32_t wave_start[kServeBins + 1];  // waves [wave_start[b], wave_start[b+1]) start on bin b
  uint32_t nbins;
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

// A page header (word 0 the link, word 1 the content) is read and written as
// one 16-byte access (pages are aligned to at least 16 bytes). Reading back
// 64 bytes of content per page instead (four 16-byte accesses) cost the
// short-sequence loop 12 % of its rate and the long one 21 %: its
// iterations wait on memory accesses in flight, not on bytes.
typedef uint64_t page_pair __attribute__((ext_vector_type(2)));

__device__ inline void write_page(uint64_t* p, uint64_t link, uint64_t id, uint32_t page) {
  *(page_pair*)p = page_pair{link, gw::page_word(id, page)};
}

// fold the content word of a header already loaded; returns the link
__device__ inline uint64_t fold_page(page_pair a, uint64_t& h) {
  h = gw::fold(h, a.y);
  return a.x;
}

__global__ void serve_start_kernel(serve_state* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    atom::store(&st->first_dry, ~0ull);
    atom::store(&st->t0, (unsigned long long)wall_ticks());
  }
}

// Pages of one sequence are claimed `run` at a time (1..64): a growing lane
// whose reserved pages are used up claims the next min(run, pages left)
// pages with one bulk claim (Gallatin::malloc_run_at: consecutive slices of
// one block, accounted as separate allocations), then takes one reserved
// page per step. The first page of each claim carries a mark in its link
// word; draining (newest page first) ends a claim's pages at that mark and
// frees them with one free_run. run == 1 is one malloc and one free per
// page. Pages of the block and segment paths are always claimed one by one.
constexpr uint32_t kServeMaxRun = 64;
constexpr uint64_t kRunStart = 1ull;  // link-word mark of a claim's first page (pages are 16-B aligned)

template <class A>
__global__ void __launch_bounds__(256)
    serve_kernel(A* a, const serve_req* reqs, uint64_t n, serve_state* st, uint64_t* checksums,
                 uint64_t* admit_ticks, uint64_t* done_ticks, unsigned long long* hist_out, uint64_t headroom,
                 uint64_t deadline, uint32_t run, serve_bins bins, uint64_t window) {
  __shared__ uint32_t hist[kHistBuckets];
  for (uint32_t i = threadIdx.x; i < kHistBuckets; i += blockDim.x) hist[i] = 0;
  __syncthreads();

  const uint32_t lane = wave::lane_id();
  const uint64_t t0 = atom::load(&st->t0);
  const uint64_t tid = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x;
  uint32_t* free_segs = ga::ctl_word(a->control, ga::kCtlFreeSegs);

  uint32_t state = kSeqIdle;
  uint64_t ticket = 0;
  // this wave's bin, and how many bins it has found out of tickets
  const uint64_t my_wave = global_wave();
  uint32_t bin = 0;
  while (bin + 1 < bins.nbins && my_wave >= bins.wave_start[bin + 1]) bin++;
  uint32_t tried = 0;
  serve_req r{0, 0, 0, 0};
  uint64_t stride = 0;          // bytes between the pages of one claim (0: one page per claim)
  uint32_t step = 0;            // pages written so far
  uint32_t claimed = 0;         // pages claimed so far (written + reserved)
  char* reserved = nullptr;     // next reserved page
  uint64_t* page = nullptr;     // the sequence's last page
  uint32_t drained = 0;         // pages drained since the last claim mark
  uint64_t* pend_p = nullptr;   // a drained claim waiting to be freed: its first page ...
  uint32_t pend_k = 0;          // ... and its page count (0: none)
  bool mark = false;            // the next page written is the first of a claim
  uint64_t h = 0;               // checksum
  uint64_t unflushed = 0;       // allocated bytes not yet taken off `committed`
  bool cut = false;             // aborted by the deadline
  uint32_t nm = 0, nf = 0, nr = 0, nd = 0, na = 0, nref = 0;
  uint32_t nw = 0;  // pages written and drained before the end of the measuring window

  uint32_t admit_wait = 0;                  // wave-uniform: iterations left before the next admission try
  uint32_t idle_age = 0;                    // wave-uniform: iterations the wave's idle lanes have waited
  uint32_t admit_backoff = kAdmitBackoffMin;  // wave-uniform: the wait after the next refusal
  for (uint32_t it = 0;; it++) {
    const uint64_t now = wall_ticks() - t0;
    const bool late = now > deadline;
    const bool in_window = now <= window;

    // Loads whose results are needed later in the iteration are issued
    // first, so that their latency overlaps the ticket atomic: the header of
    // the lane's current page (a growing lane folds its previous page, a
    // draining lane the page it frees next) and, while the wave has lanes to
    // admit, the admission budget.
    uint64_t* const top_page = page;
    const bool reading = (state == kSeqGrowing || state == kSeqDraining) && page != nullptr;
    page_pair q0 = {0, 0};
    if (reading) q0 = *(const page_pair*)page;
    // idle lanes that take tickets in this iteration
    uint64_t idle = wave::ballot(state == kSeqIdle);
    if (idle && wave::popc(idle) < kTicketBatch && idle_age < kTicketWait &&
        wave::ballot(state == kSeqWaiting || state == kSeqGrowing || state == kSeqDraining)) {
      idle_age++;
      idle = 0;
    }
    uint64_t room = 0, seen = 0;
    const bool admit_now = admit_wait == 0;
    if (admit_wait) admit_wait--;
    if (admit_now && (idle || wave::ballot(state == kSeqWaiting)) && lane == 0) {
      room = (uint64_t)atom::load(free_segs) * A::kBytesPerSegment;
      seen = atom::load(&st->committed.v);
    }

    // ---- tickets for idle lanes, from this wave's bin
    if (idle) {
      idle_age = 0;
      const int leader = wave::first_lane(idle);
      const uint32_t k = wave::popc(idle);
      uint64_t base = 0;
      if ((int)lane == leader) base = atom::add(&st->cursors[bin].v, (unsigned long long)k);
      base = wave::readlane64(base, leader);
      if (state == kSeqIdle) {
        const uint64_t t = base + wave::popc(idle & wave::lanemask_lt());
        if (t < bins.count[bin]) {
          ticket = bins.start[bin] + t;
          r = reqs[ticket];
          state = kSeqWaiting;
        }
      }
      if (base + k >= bins.count[bin]) {  // the bin is out of tickets
        tried++;
        bin = bin + 1 == bins.nbins ? 0 : bin + 1;
        if (tried >= bins.nbins && state == kSeqIdle) state = kSeqExhausted;
        if (tried == bins.nbins && (int)lane == leader)  // this wave found every queue empty
          (void)__hip_atomic_fetch_min(&st->first_dry, (unsigned long long)now, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
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
    const bool cutting = late && state == kSeqGrowing;
    if (cutting) {
      cut = true;
      unflushed += (uint64_t)(r.pages - claimed) * r.page_bytes;  // never claimed
      state = kSeqDraining;
    }
    // reserved pages of a cut sequence go back at once
    const bool spare = cutting && claimed > step;
    if (wave::ballot(spare)) {
      if (spare) {
        a->free_run(reserved, claimed - step);
        nf += claimed - step;
      }
    }

    // ---- admission of arrived requests, all of the wave's at once when they fit
    const bool arrived = admit_now && state == kSeqWaiting && r.arrival <= now;
    const uint64_t want = wave::ballot(arrived);
    uint64_t granted = 0;
    if (want) {
      const uint64_t need = arrived ? (uint64_t)r.pages * r.page_bytes : 0ull;
      const uint64_t total = wave_sum64(need);
      const int leader = wave::first_lane(want);
      room = wave::readlane64(room, 0);
      seen = wave::readlane64(seen, 0);
      if ((int)lane == leader) {
        // the add only when the ask looks like it fits
        // all arrived lanes, else the leader alone
        for (int pass = 0; pass < 2 && !granted; pass++) {
          const uint64_t ask = pass == 0 ? total : need;
          if (pass == 1 && ask == total) break;
          if (seen + ask + headroom > room) continue;
          const uint64_t c = atom::add(&st->committed.v, (unsigned long long)ask);
          if (c + ask + headroom <= room) {
            granted = pass == 0 ? want : (1ull << lane);
            (void)__hip_atomic_fetch_max(&st->peak_committed.v, (unsigned long long)(c + ask), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
          } else {
            atom::sub(&st->committed.v, (unsigned long long)ask);
          }
        }
        if (!granted) nref++;
      }
      granted = wave::readlane64(granted, leader);
      if (granted) {
        admit_backoff = kAdmitBackoffMin;
      } else {
        admit_wait = admit_backoff;
        admit_backoff = admit_backoff < kAdmitBackoffMax ? 2 * admit_backoff : kAdmitBackoffMax;
      }
      if ((granted >> lane) & 1ull) {
        state = kSeqGrowing;
        step = claimed = 0;
        page = nullptr;
        h = r.id;
        cut = false;
        stride = run > 1 ? A::run_stride(r.page_bytes) : 0;
        admit_ticks[ticket] = now;
      }
    }

    // ---- growing lanes: one page each, claimed `run` at a time
    const bool grow = state == kSeqGrowing;
    const bool claim = grow && claimed == step;
    const uint64_t claiming = wave::ballot(claim);
    if (claiming) {
      const uint64_t ta = wall_ticks();
      uint32_t k = 0;
      if (claim) {
        const uint32_t left = r.pages - step;
        k = stride ? (left < run ? left : run) : 1u;
        reserved = k > 1 ? (char*)a->malloc_run_at(r.page_bytes, k, tid, A::cu_hint(), 0)
                         : (char*)a->malloc(r.page_bytes);
      }
      const uint64_t tb = wall_ticks();
      const uint32_t pages_claimed = (uint32_t)wave_sum64(claim && reserved ? k : 0u);
      if (lane == 0 && pages_claimed) atomicAdd(&hist[hist_bucket(tb - ta)], pages_claimed);
      if (claim) {
        if (reserved) {
          mark = true;
          claimed += k;
          nm += k;
          unflushed += (uint64_t)k * r.page_bytes;
        } else {
          nr++;
        }
      }
    }
    if (grow && claimed > step) {
      uint64_t* p = (uint64_t*)reserved;
      reserved += stride;
      write_page(p, (uint64_t)page | (mark ? kRunStart : 0ull), r.id, step);
      mark = false;
      if (page) (void)fold_page(q0, h);
      page = p;
      step++;
      if (in_window) nw++;
      if (step == r.pages) state = kSeqDraining;
    }

    // ---- draining lanes: one page each, newest first; a claim's pages are
    // released together once its first page is reached. A release waits in
    // the lane's pending slot; the wave frees every pending claim in one call
    // (one release fence) once kFreeBatch lanes hold one, or a lane needs its
    // slot for the next release, or the wave leaves.
    const bool drain = state == kSeqDraining && page != nullptr && page == top_page && reading;
    const uint64_t draining = wave::ballot(drain);
    bool release = false;
    uint64_t* first = nullptr;
    if (draining) {
      if (drain) {
        first = page;
        const uint64_t link = fold_page(q0, h);
        if (in_window) nw++;
        page = (uint64_t*)(link & ~kRunStart);
        drained++;
        release = (link & kRunStart) != 0;
      }
      if (wave::ballot(release && pend_k != 0) || wave::popc(wave::ballot(pend_k != 0)) >= kFreeBatch) {
        if (pend_k) {
          if (pend_k > 1)
            a->free_run(pend_p, pend_k);
          else
            a->free(pend_p);
          nf += pend_k;
          pend_k = 0;
        }
      }
      if (release) {
        pend_p = first;
        pend_k = drained;
        drained = 0;
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
    }

    // ---- return allocated bytes to the admission budget, leave when done
    const bool all_done = wave::ballot(state != kSeqExhausted) == 0;
    if (all_done && wave::ballot(pend_k != 0)) {
      if (pend_k) {
        if (pend_k > 1)
          a->free_run(pend_p, pend_k);
        else
          a->free(pend_p);
        nf += pend_k;
        pend_k = 0;
      }
    }
    if ((it % kFlushEvery) == kFlushEvery - 1 || all_done) {
      const uint64_t s = wave_sum64(unflushed);
      if (lane == 0 && s) atom::sub(&st->committed.v, (unsigned long long)s);
      unflushed = 0;
    }
    if (all_done) break;
    if (!wave::ballot(grow) && !draining && !granted) __builtin_amdgcn_s_sleep(32);
  }

  // per-wave totals
  const uint64_t tm = wave_sum64(nm), tf = wave_sum64(nf), tr = wave_sum64(nr);
  const uint64_t td = wave_sum64(nd), ta = wave_sum64(na), tref = wave_sum64(nref), tw = wave_sum64(nw);
  if (lane == 0) {
    if (tm) atom::add(&st->page_mallocs, (unsigned long long)tm);
    if (tf) atom::add(&st->page_frees, (unsigned long long)tf);
    if (tr) atom::add(&st->page_retries, (unsigned long long)tr);
    if (td) atom::add(&st->done, (unsigned long long)td);
    if (ta) atom::add(&st->aborted, (unsigned long long)ta);
    if (tref) atom::add(&st->refused, (unsigned long long)tref);
    if (tw) atom::add(&st->window_pages, (unsigned long long)tw);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kHistBuckets; i += blockDim.x)
    if (hist[i]) atom::add(&hist_out[i], (unsigned long long)hist[i]);
}
