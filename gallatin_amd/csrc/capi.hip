// capi.hip — the C-ABI of the MI355X Gallatin allocator (include/gallatin_amd.h).
//
// Templates cannot cross a C ABI, so each supported (segment, min, max)
// triple is instantiated here once and reached through a table of function
// pointers. All device work runs on the handle's own stream. The benchmark
// and test workloads that drive the allocator live in a library of their
// own (workloads.hip, include/gallatin_workloads.h) built on this one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include <gallatin/allocators/global_allocator.hpp>

#include "../../include/gallatin_amd.h"

namespace ga = gallatin::allocators;
namespace gu = gallatin::utils;
namespace wave = gallatin::wave;
namespace atom = gallatin::atom;

namespace {
struct impl_ops;
}  // namespace

struct gal_allocator {
  const impl_ops* ops = nullptr;
  void* dev = nullptr;  // Gallatin<...>* in device memory
  gal_info_t info{};
  hipStream_t stream = nullptr;
  uint32_t* control = nullptr;  // device control words of the allocator
  // holds taken by objects built on this handle and not yet released
  // (gal_retain): they use its stream and allocator, so the handle outlives them
  int children = 0;
  // device address of the no-allocation scratch array on this handle's GPU
  uint64_t* scratch = nullptr;
};


namespace {

struct impl_ops {
  uint64_t seg, min, max;
  void* (*create)(uint64_t pool, uint64_t seed);
  void* (*create_host)(uint64_t pool, uint64_t seed);
  uint64_t (*allocated_size)(uint64_t size);
  void (*destroy)(void* dev);
  int (*describe)(gal_allocator* h);
  int (*snapshot)(gal_allocator* h, uint8_t* out);
  int (*replay)(gal_allocator* h, const gal_op* d_ops, uint64_t n, uint64_t* d_results, uint32_t* d_status);
  int (*group_replay)(gal_allocator* h, const gal_group_op* d_ops, uint64_t n, uint64_t* d_results,
                      uint32_t* d_status, uint64_t* d_freed);
  uint32_t (*path_of)(uint64_t size);  // 0 slice (incl. big slices), 1 whole block, 2 segments
  void (*print)(gal_allocator* h);
  int (*fill)(gal_allocator* h, gal_fill_report* out);
  int (*segment_words)(gal_allocator* h, std::vector<uint64_t>& words);
};

// ---------------------------------------------------------------- kernels
template <class A>
__global__ void replay_kernel(A* a, const gal_op* ops, uint64_t n, uint64_t* results, uint32_t* status) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  char* base = a->table.memory;
  for (uint64_t i = 0; i < n; i++) {
    const gal_op op = ops[i];
    if (op.kind == GAL_OP_MALLOC) {
      void* p = a->malloc_at(op.arg, op.tid, op.slot);
      results[i] = p ? (uint64_t)((char*)p - base) : GAL_NULL_OFFSET;
    } else {
      const uint64_t r = results[op.arg];
      if (r == GAL_NULL_OFFSET) {
        *status = 1u;
        return;
      }
      a->free(base + r);
      results[i] = 0;
    }
  }
}

// One wavefront runs the group ops in order: the lanes of op i's mask reach
// the malloc (or free) call together, with exactly that exec mask. Claims
// of GAL_OP_MALLOC use the reference's pinned-slot walk (no overflow-slot
// spreading); GAL_OP_MALLOC_SPREAD takes malloc's first look under load
// (the block has one wave, so the hint is the op's slot unmixed). Either way
// the results depend on the op stream only. Every lane writes its own result
// entry (GAL_NULL_OFFSET for lanes outside the op); a free reads the result
// of an earlier op of this wave, written by any of its lanes, after the
// workgroup-scope fence that ends each op.
//
// Before an op runs, its lanes check it: an unknown kind, an empty mask, or
// a free of a result that failed, stop the replay (*status = op index + 1).
// With `freed` (one u64 per op, zeroed: bit l set once result (op, l) has
// been freed) the wave also checks what the host checks otherwise: a free
// names a lane of an earlier malloc op, no result is freed twice, a free op
// of several lanes frees no segment run, a malloc op's sizes are nonzero
// and all on the slice path or all above it.
template <class A>
__global__ void __launch_bounds__(64) group_replay_kernel(A* a, const gal_group_op* ops, uint64_t n,
                                                          uint64_t* results, uint32_t* status, uint64_t* freed) {
  const uint32_t lane = wave::lane_id();
  char* base = a->table.memory;
  for (uint64_t i = 0; i < n; i++) {
    const gal_group_op* op = &ops[i];
    const uint64_t lanes = op->lanes;
    const uint32_t kind = op->kind;
    const bool in = (lanes >> lane) & 1ull;
    const uint64_t arg = in ? op->arg[lane] : 0;
    const bool is_malloc = kind == GAL_OP_MALLOC || kind == GAL_OP_MALLOC_SPREAD;
    bool bad = lanes == 0 || (!is_malloc && kind != GAL_OP_FREE);
    uint32_t path = 0;
    if (!bad && in && is_malloc) {
      bad = freed && arg == 0;
      path = A::request_path(arg);
    } else if (!bad && in) {
      if (freed) {
        const uint64_t src = arg / 64, src_lane = arg % 64;
        const uint32_t src_kind = arg >= i * 64 ? GAL_OP_FREE : ops[src].kind;
        bad = (src_kind != GAL_OP_MALLOC && src_kind != GAL_OP_MALLOC_SPREAD) || !((ops[src].lanes >> src_lane) & 1ull);
        if (!bad && (lanes & (lanes - 1)) != 0) bad = A::request_path(ops[src].arg[src_lane]) == 2u;
        if (!bad) bad = ((atom::fetch_or(&freed[src], 1ull << src_lane) >> src_lane) & 1ull) != 0;
      }
      if (!bad) bad = results[arg] == GAL_NULL_OFFSET;
    }
    if (freed && is_malloc) {
      const uint64_t large = wave::ballot(in && path != 0u), small = wave::ballot(in && path == 0u);
      bad = bad || (large != 0 && small != 0);
    }
    if (wave::ballot(bad)) {
      if (lane == 0) atom::store(status, (uint32_t)(i + 1));
      return;
    }
    if (!in) {
      results[i * 64 + lane] = GAL_NULL_OFFSET;
    } else if (is_malloc) {
      // the spread form as the oracle models it: first look, no block fetched ahead
      void* p = kind == GAL_OP_MALLOC_SPREAD ? a->template malloc_at<true, false>(arg, op->tid[lane], op->slot)
                                             : a->template malloc_at<false>(arg, op->tid[lane], op->slot);
      results[i * 64 + lane] = p ? (uint64_t)((char*)p - base) : GAL_NULL_OFFSET;
    } else {
      a->free(base + results[arg]);
      results[i * 64 + lane] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  }
}

// Single-lane replay through the global allocator: global_malloc uses the
// lane's CU slot hint (Gallatin::cu_hint) and thread id 0.
__global__ void global_replay_kernel(const gal_op* ops, uint64_t n, uint64_t* results, uint32_t* status,
                                     uint32_t* cu_id) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  *cu_id = ga::global_gallatin->cu_hint();
  char* base = ga::global_gallatin->table.memory;
  for (uint64_t i = 0; i < n; i++) {
    const gal_op op = ops[i];
    if (op.kind == GAL_OP_MALLOC) {
      void* p = ga::global_malloc(op.arg);
      results[i] = p ? (uint64_t)((char*)p - base) : GAL_NULL_OFFSET;
    } else {
      const uint64_t r = results[op.arg];
      if (r == GAL_NULL_OFFSET) {
        *status = 1u;
        return;
      }
      ga::global_free(base + r);
      results[i] = 0;
    }
  }
}

// The segment tree's gather rule (Gallatin::malloc_segment_allocation ->
// veb_tree::gather_bottom, reference veb.cuh:323-452) evaluated without
// claiming anything: can n contiguous segments be claimed from these item
// words right now? Same control flow as the claim: words from the top down,
// a run is a word's trailing ones extended through full words below and the
// top bits of the next lower one; it may not reach below word 0.
bool gather_would_succeed(const uint64_t* b, uint64_t words, uint64_t n) {
  if (n == 0 || words == 0) return false;
  uint64_t w = words - 1;
  while (w != 0) {
    uint64_t need = n;
    const int run = gu::trailing_ones(b[w]);
    if (run == 0) {
      w--;
      continue;
    }
    if ((uint64_t)run >= need) return true;
    need -= (uint64_t)run;
    bool ok = true;
    while (need > 64) {
      if (w == 0) return false;
      w--;
      if (b[w] != ~0ull) {
        ok = false;
        break;
      }
      need -= 64;
    }
    if (!ok) continue;
    if (w == 0) return false;
    w--;
    const uint64_t top = gu::bitmask(need) << (64 - need);
    if ((b[w] & top) == top) return true;
  }
  return false;
}

inline unsigned blocks_for(uint64_t n, unsigned threads) { return (unsigned)((n + threads - 1) / threads); }

// ---- scratch of the entry point that allocates nothing
// (gal_group_replay_noalloc): 16 words of status (the first u32 is the
// replay status), then one u64 per op whose bit l is set once result (op, l)
// has been freed. Every GPU holds its own copy until the process ends; its
// address on each device comes from hipGetSymbolAddress (the array's name is
// only a host-side stand-in).
constexpr uint64_t kScratchStatusWords = 16;
static __device__ __align__(256) uint64_t scratch[kScratchStatusWords + GAL_NOALLOC_MAX_OPS];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint64_t* g_scratch[kMaxDevices] = {};

// the scratch address on the current device, looked up once per device
int scratch_on_current_device(uint64_t** out) {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) return GAL_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GAL_ERR_HIP;
    g_scratch[device] = (uint64_t*)p;
  }
  *out = g_scratch[device];
  return GAL_OK;
}

__global__ void scratch_clear_kernel(uint64_t* p, uint64_t words) {
  for (uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// true when p is device memory (or memory the device can address as its own)
bool is_device_pointer(const void* p) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// ------------------------------------------------------ per-type impl
template <uint64_t S, uint64_t M, uint64_t B>
struct impl {
  using A = ga::Gallatin<S, M, B>;

  static void* create(uint64_t pool, uint64_t seed) { return (void*)A::generate_on_device(pool, seed, false); }
  static void* create_host(uint64_t pool, uint64_t seed) {
    return (void*)A::generate_on_device_host(pool, seed, false);
  }
  static uint64_t allocated_size(uint64_t size) { return A::allocated_size(size); }
  static void destroy(void* dev) { A::free_on_device((A*)dev); }

  static int describe(gal_allocator* h) {
    A host;
    if (!A::host_copy((const A*)h->dev, &host)) return GAL_ERR_HIP;
    gal_info_t& i = h->info;
    std::memset(&i, 0, sizeof(i));
    i.segment_size = S;
    i.min_size = M;
    i.max_size = B;
    i.num_segments = host.table.num_segments;
    i.blocks_per_segment = A::kBlocksPerSegment;
    i.num_trees = A::kNumTrees;
    i.veb_layers = (uint64_t)host.segment_tree.num_layers;
    i.total_slots = A::total_slots();
    i.pool_device_ptr = (uint64_t)host.table.memory;
    i.device_handle = (uint64_t)h->dev;
    uint64_t words = 8;
    for (int l = 0; l < host.segment_tree.num_layers; l++) words += host.segment_tree.words[l] * (1 + A::kNumTrees);
    const uint64_t nseg = i.num_segments, nblocks = nseg * A::kBlocksPerSegment;
    uint64_t bytes = words * 8 + ((nseg * 2 + 7) / 8) * 8 + nblocks * 12 + nseg * 16 + i.total_slots * 8;
    i.snapshot_bytes = ((bytes + 7) / 8) * 8;
    i.metadata_bytes =
        words * 8 + nseg * 4 + nblocks * (sizeof(ga::Block) + 4) + nseg * 16 + i.total_slots * ga::kSlotStride * 4 +
        (ga::kCtlWords + A::kNumTrees) * ga::kCtlStride * 4 + sizeof(A);
    h->control = host.control;
    return GAL_OK;
  }

  static int snapshot(gal_allocator* h, uint8_t* out) {
    A host;
    if (!A::host_copy((const A*)h->dev, &host)) return GAL_ERR_HIP;
    if (hipDeviceSynchronize() != hipSuccess) return GAL_ERR_HIP;
    const uint64_t nseg = host.table.num_segments, nblocks = nseg * A::kBlocksPerSegment;
    uint8_t* p = out;
    uint32_t locks = 0;
    if (hipMemcpy(&locks, ga::ctl_word(host.control, ga::kCtlLocks), 4, hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
    uint64_t header[8] = {nseg, A::kBlocksPerSegment, A::kNumTrees, (uint64_t)host.segment_tree.num_layers,
                          A::total_slots(), locks, 0, 0};
    std::memcpy(p, header, sizeof(header));
    p += sizeof(header);
    auto tree = [&](const ga::veb_tree& t) {
      for (int l = 0; l < t.num_layers; l++) {
        if (hipMemcpy(p, t.bits[l], t.words[l] * 8, hipMemcpyDeviceToHost) != hipSuccess) return false;
        p += t.words[l] * 8;
      }
      return true;
    };
    if (!tree(host.segment_tree)) return GAL_ERR_HIP;
    for (uint32_t c = 0; c < A::kNumTrees; c++)
      if (!tree(host.sub_trees[c])) return GAL_ERR_HIP;
    std::vector<uint32_t> ids(nseg);
    if (hipMemcpy(ids.data(), host.table.chunk_ids, nseg * 4, hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
    for (uint64_t s = 0; s < nseg; s++) {
      uint16_t v = (uint16_t)ids[s];
      std::memcpy(p, &v, 2);
      p += 2;
    }
    uint64_t pad = ((nseg * 2 + 7) / 8) * 8 - nseg * 2;
    std::memset(p, 0, pad);
    p += pad;
    // {malloc, free} per block in the reference's packing (class << 20 | count);
    // the padded records are copied whole (one contiguous copy) and packed here
    std::vector<ga::Block> rec(nblocks);
    if (hipMemcpy(rec.data(), host.table.blocks, nblocks * sizeof(ga::Block), hipMemcpyDeviceToHost) != hipSuccess)
      return GAL_ERR_HIP;
    for (uint64_t k = 0; k < nblocks; k++) {
      const uint64_t m = rec[k].malloc_counter;
      const uint32_t count = ga::counter_count(m) > 0xFFFFFu ? 0xFFFFFu : ga::counter_count(m);
      const uint32_t pair[2] = {(ga::counter_class(m) << ga::kSnapshotClassShift) | count, rec[k].free_counter};
      std::memcpy(p, pair, 8);
      p += 8;
    }
    if (hipMemcpy(p, host.table.queues, nblocks * 4, hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
    p += nblocks * 4;
    if (hipMemcpy(p, host.table.segs, nseg * 16, hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
    p += nseg * 16;
    {
      // each slot's entry and ahead word have 128-B lines of their own on the
      // device; packed in the image, the entries then the ahead words
      const uint32_t n = A::total_slots();
      std::vector<uint32_t> lines((size_t)n * ga::kSlotStride);
      if (hipMemcpy(lines.data(), host.slots, lines.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return GAL_ERR_HIP;
      for (uint32_t k = 0; k < n; k++) {
        std::memcpy(p + 4 * (size_t)k, &lines[(size_t)k * ga::kSlotStride], 4);
        std::memcpy(p + 4 * ((size_t)n + k), &lines[(size_t)k * ga::kSlotStride + ga::kAheadWord], 4);
      }
    }
    p += A::total_slots() * 8;
    uint64_t used = (uint64_t)(p - out);
    if (h->info.snapshot_bytes > used) std::memset(p, 0, h->info.snapshot_bytes - used);
    return GAL_OK;
  }

  static int replay(gal_allocator* h, const gal_op* d_ops, uint64_t n, uint64_t* d_results, uint32_t* d_status) {
    replay_kernel<A><<<1, 1, 0, h->stream>>>((A*)h->dev, d_ops, n, d_results, d_status);
    return hipGetLastError() == hipSuccess ? GAL_OK : GAL_ERR_HIP;
  }

  static int group_replay(gal_allocator* h, const gal_group_op* d_ops, uint64_t n, uint64_t* d_results,
                          uint32_t* d_status, uint64_t* d_freed) {
    group_replay_kernel<A><<<1, 64, 0, h->stream>>>((A*)h->dev, d_ops, n, d_results, d_status, d_freed);
    return hipGetLastError() == hipSuccess ? GAL_OK : GAL_ERR_HIP;
  }
  static uint32_t path_of(uint64_t size) { return A::request_path(size); }

  static void print(gal_allocator* h) { ((A*)h->dev)->print_info(); }
  static int fill(gal_allocator* h, gal_fill_report* out) {
    A host;
    if (!A::host_copy((const A*)h->dev, &host)) return GAL_ERR_HIP;
    std::memset(out, 0, sizeof(*out));
    out->num_trees = A::kNumTrees;
    auto items = [&](const ga::veb_tree& t, uint64_t& n) {
      const int b = t.num_layers - 1;
      std::vector<uint64_t> w(t.words[b]);
      if (hipMemcpy(w.data(), t.bits[b], w.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) return false;
      n = 0;
      for (uint64_t x : w) n += (uint64_t)__builtin_popcountll(x);
      return true;
    };
    if (!items(host.segment_tree, out->free_segments)) return GAL_ERR_HIP;
    out->max_segments = host.table.num_segments;
    for (uint32_t c = 0; c < A::kNumTrees; c++)
      if (!items(host.sub_trees[c], out->listed_segments[c])) return GAL_ERR_HIP;
    const uint64_t nseg = host.table.num_segments;
    std::vector<uint32_t> ids(nseg);
    std::vector<ga::segment_counters> segs(nseg);
    if (hipMemcpy(ids.data(), host.table.chunk_ids, nseg * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(segs.data(), host.table.segs, nseg * sizeof(ga::segment_counters), hipMemcpyDeviceToHost) !=
            hipSuccess)
      return GAL_ERR_HIP;
    for (uint64_t x = 0; x < nseg; x++) {
      const uint32_t c = ids[x];
      if (c >= A::kNumTrees || segs[x].active == -1) continue;
      const int64_t nb = (int64_t)A::table_type::blocks_in_segment(c);
      out->live_segments[c]++;
      out->blocks_available[c] += (uint64_t)segs[x].active;
      out->blocks_in_use[c] += (uint64_t)(nb - segs[x].active);
    }
    out->overhead_bytes = h->info.metadata_bytes;
    std::vector<int32_t> words(A::kNumTrees * ga::kCtlStride);
    if (hipMemcpy(words.data(), host.class_words, words.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
      return GAL_ERR_HIP;
    for (uint32_t c = 0; c < A::kNumTrees; c++) out->empty_pinned_slots[c] = (uint64_t)(int64_t)words[c * ga::kCtlStride];
    return GAL_OK;
  }
  // item words of the segment tree
  static int segment_words(gal_allocator* h, std::vector<uint64_t>& words) {
    A host;
    if (!A::host_copy((const A*)h->dev, &host)) return GAL_ERR_HIP;
    const ga::veb_tree& t = host.segment_tree;
    words.resize(t.words[t.num_layers - 1]);
    return hipMemcpy(words.data(), t.bits[t.num_layers - 1], words.size() * 8, hipMemcpyDeviceToHost) == hipSuccess
               ? GAL_OK
               : GAL_ERR_HIP;
  }

  static constexpr impl_ops ops() {
    return impl_ops{S,         M,       B,             &create,  &create_host, &allocated_size, &destroy, &describe,
                    &snapshot, &replay, &group_replay, &path_of, &print,       &fill,           &segment_words};
  }
};

// The instantiation table: (segment bytes, smallest slice, largest slice).
constexpr uint64_t MiB = 1ull << 20;
const impl_ops kImpls[] = {
    impl<16 * MiB, 16, 4096>::ops(),   // the reference's default and global allocator
    impl<1 * MiB, 16, 256>::ops(),     // small segments: exhaustion and recycling tests
    impl<4 * MiB, 128, 1024>::ops(),   // slab-hash tuning space (min 128 B)
    impl<16 * MiB, 128, 4096>::ops(),
    impl<64 * MiB, 128, 4096>::ops(),
    impl<64 * MiB, 16, 16384>::ops(),  // 16 KiB pages as single slices
};

const impl_ops* find_impl(uint64_t seg, uint64_t min, uint64_t max) {
  for (const auto& o : kImpls)
    if (o.seg == seg && o.min == min && o.max == max) return &o;
  return nullptr;
}

int read_errors(gal_allocator* h, uint64_t* internal, uint64_t* bad) {
  uint32_t ctl[ga::kCtlWords * ga::kCtlStride];
  if (hipMemcpy(ctl, h->control, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
  if (internal) *internal = *ga::ctl_word(ctl, ga::kCtlErrors);
  if (bad) *bad = *ga::ctl_word(ctl, ga::kCtlBadFrees);
  return GAL_OK;
}


// state of the global allocators made by gal_init_global_allocator(_host)
gal_allocator* g_global = nullptr;
gal_allocator* g_global_host = nullptr;

}  // namespace

extern "C" {

int gal_supported(uint64_t* out, int cap) {
  int n = (int)(sizeof(kImpls) / sizeof(kImpls[0]));
  for (int i = 0; i < n && i < cap && out; i++) {
    out[3 * i] = kImpls[i].seg;
    out[3 * i + 1] = kImpls[i].min;
    out[3 * i + 2] = kImpls[i].max;
  }
  return n;
}

// On failure the handle's own resources are released; the device allocator
// `dev` stays with the caller, which destroys it.
static int make_handle(const impl_ops* ops, void* dev, gal_allocator** out) {
  gal_allocator* h = new gal_allocator();
  h->ops = ops;
  h->dev = dev;
  auto fail = [&](int rc) {
    h->dev = nullptr;
    gal_destroy(h);
    return rc;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(GAL_ERR_HIP);
  int rc = ops->describe(h);
  if (rc != GAL_OK) return fail(rc);
  // set-up of the no-allocation entry point: the scratch array's address on
  // this handle's device (the current one)
  rc = scratch_on_current_device(&h->scratch);
  if (rc != GAL_OK) return fail(rc);
  *out = h;
  return GAL_OK;
}

int gal_create(uint64_t segment_size, uint64_t min_size, uint64_t max_size, uint64_t pool_bytes, uint64_t seed,
               gal_allocator** out) {
  if (!out) return GAL_ERR_INVALID;
  *out = nullptr;
  const impl_ops* ops = find_impl(segment_size, min_size, max_size);
  if (!ops) return GAL_ERR_UNSUPPORTED;
  if (pool_bytes < segment_size) return GAL_ERR_INVALID;
  void* dev = ops->create(pool_bytes, seed);
  if (!dev) return GAL_ERR_OOM;
  gal_allocator* h = nullptr;
  int rc = make_handle(ops, dev, &h);
  if (rc != GAL_OK) {
    ops->destroy(dev);
    return rc;
  }
  *out = h;
  return GAL_OK;
}

int gal_destroy(gal_allocator* h) {
  if (!h) return GAL_ERR_INVALID;
  if (h->children > 0) return GAL_ERR_STATE;  // an object built on it still holds it
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->dev) h->ops->destroy(h->dev);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return GAL_OK;
}

int gal_get_info(gal_allocator* h, gal_info_t* info) {
  if (!h || !info) return GAL_ERR_INVALID;
  *info = h->info;
  return GAL_OK;
}

int gal_print_stats(gal_allocator* h) {
  if (!h) return GAL_ERR_INVALID;
  (void)hipStreamSynchronize(h->stream);
  h->ops->print(h);
  fflush(stdout);
  return GAL_OK;
}

int gal_fill(gal_allocator* h, gal_fill_report* out) {
  if (!h || !out) return GAL_ERR_INVALID;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  return h->ops->fill(h, out);
}

int gal_snapshot(gal_allocator* h, void* buf, uint64_t cap, uint64_t* needed) {
  if (!h) return GAL_ERR_INVALID;
  if (needed) *needed = h->info.snapshot_bytes;
  if (!buf || cap < h->info.snapshot_bytes) return GAL_ERR_BUFFER;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  return h->ops->snapshot(h, (uint8_t*)buf);
}

int gal_errors(gal_allocator* h, uint64_t* internal_errors, uint64_t* bad_frees) {
  if (!h) return GAL_ERR_INVALID;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  return read_errors(h, internal_errors, bad_frees);
}

int gal_error_sites(gal_allocator* h, uint32_t* counts, uint32_t cap) {
  if (!h || !counts) return GAL_ERR_INVALID;
  uint32_t ctl[ga::kCtlWords * ga::kCtlStride];
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  if (hipMemcpy(ctl, h->control, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
  for (uint32_t i = 0; i < cap && i < ga::kErrSites; i++) counts[i] = *ga::ctl_word(ctl, ga::kCtlSiteBase + i);
  return GAL_OK;
}

int gal_stats(gal_allocator* h, uint32_t* counts, uint32_t cap) {
  if (!h || !counts) return GAL_ERR_INVALID;
  uint32_t ctl[ga::kCtlWords * ga::kCtlStride];
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  if (hipMemcpy(ctl, h->control, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
  for (uint32_t i = 0; i < cap && i < ga::kStatSites; i++) counts[i] = *ga::ctl_word(ctl, ga::kCtlStatBase + i);
  return GAL_OK;
}

int gal_replay(gal_allocator* h, const gal_op* ops, uint64_t n, uint64_t* results) {
  if (!h || (n && (!ops || !results))) return GAL_ERR_INVALID;
  if (n == 0) return GAL_OK;
  // a free must name an earlier malloc op
  for (uint64_t i = 0; i < n; i++) {
    if (ops[i].kind == GAL_OP_FREE) {
      if (ops[i].arg >= i || ops[ops[i].arg].kind != GAL_OP_MALLOC) return GAL_ERR_INVALID;
    } else if (ops[i].kind != GAL_OP_MALLOC) {
      return GAL_ERR_INVALID;
    }
  }
  gal_op* d_ops = nullptr;
  uint64_t* d_res = nullptr;
  uint32_t* d_status = nullptr;
  int rc = GAL_OK;
  if (hipMalloc(&d_ops, n * sizeof(gal_op)) != hipSuccess || hipMalloc(&d_res, n * 8) != hipSuccess ||
      hipMalloc(&d_status, 4) != hipSuccess) {
    rc = GAL_ERR_OOM;
  }
  if (rc == GAL_OK && (hipMemcpy(d_ops, ops, n * sizeof(gal_op), hipMemcpyHostToDevice) != hipSuccess ||
                       hipMemset(d_status, 0, 4) != hipSuccess))
    rc = GAL_ERR_HIP;
  if (rc == GAL_OK) rc = h->ops->replay(h, d_ops, n, d_res, d_status);
  if (rc == GAL_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc = GAL_ERR_HIP;
  uint32_t status = 0;
  if (rc == GAL_OK && (hipMemcpy(results, d_res, n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
                       hipMemcpy(&status, d_status, 4, hipMemcpyDeviceToHost) != hipSuccess))
    rc = GAL_ERR_HIP;
  if (rc == GAL_OK && status != 0) rc = GAL_ERR_INVALID;  // freed a failed malloc
  if (d_ops) (void)hipFree(d_ops);
  if (d_res) (void)hipFree(d_res);
  if (d_status) (void)hipFree(d_status);
  return rc;
}

int gal_group_replay(gal_allocator* h, const gal_group_op* ops, uint64_t n, uint64_t* results) {
  if (!h || (n && (!ops || !results))) return GAL_ERR_INVALID;
  if (n == 0) return GAL_OK;
  // Every op is checked before anything runs: a free names a lane of an
  // earlier malloc op, each result at most once; a malloc op's sizes take
  // one kind of path; a free op of several lanes frees no segment run.
  std::vector<uint8_t> freed(n * 64, 0);
  for (uint64_t i = 0; i < n; i++) {
    const gal_group_op& op = ops[i];
    if (op.lanes == 0) return GAL_ERR_INVALID;
    if (op.kind == GAL_OP_MALLOC || op.kind == GAL_OP_MALLOC_SPREAD) {
      bool large = false, small = false;
      for (uint64_t m = op.lanes; m; m &= m - 1) {
        const uint64_t size = op.arg[__builtin_ctzll(m)];
        if (size == 0) return GAL_ERR_INVALID;
        (h->ops->path_of(size) == 0 ? small : large) = true;
      }
      if (large && small) return GAL_ERR_INVALID;
    } else if (op.kind == GAL_OP_FREE) {
      const bool several = (op.lanes & (op.lanes - 1)) != 0;
      for (uint64_t m = op.lanes; m; m &= m - 1) {
        const uint64_t j = op.arg[__builtin_ctzll(m)];
        if (j >= i * 64 || freed[j]) return GAL_ERR_INVALID;
        const gal_group_op& src = ops[j / 64];
        if ((src.kind != GAL_OP_MALLOC && src.kind != GAL_OP_MALLOC_SPREAD) || !((src.lanes >> (j % 64)) & 1ull))
          return GAL_ERR_INVALID;
        if (several && h->ops->path_of(src.arg[j % 64]) == 2) return GAL_ERR_INVALID;
        freed[j] = 1;
      }
    } else {
      return GAL_ERR_INVALID;
    }
  }
  gal_group_op* d_ops = nullptr;
  uint64_t* d_res = nullptr;
  uint32_t* d_status = nullptr;
  int rc = GAL_OK;
  if (hipMalloc(&d_ops, n * sizeof(gal_group_op)) != hipSuccess || hipMalloc(&d_res, n * 64 * 8) != hipSuccess ||
      hipMalloc(&d_status, 4) != hipSuccess)
    rc = GAL_ERR_OOM;
  if (rc == GAL_OK && (hipMemcpy(d_ops, ops, n * sizeof(gal_group_op), hipMemcpyHostToDevice) != hipSuccess ||
                       hipMemset(d_res, 0xFF, n * 64 * 8) != hipSuccess || hipMemset(d_status, 0, 4) != hipSuccess))
    rc = GAL_ERR_HIP;
  if (rc == GAL_OK) rc = h->ops->group_replay(h, d_ops, n, d_res, d_status, nullptr);
  if (rc == GAL_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc = GAL_ERR_HIP;
  uint32_t status = 0;
  if (rc == GAL_OK && (hipMemcpy(results, d_res, n * 64 * 8, hipMemcpyDeviceToHost) != hipSuccess ||
                       hipMemcpy(&status, d_status, 4, hipMemcpyDeviceToHost) != hipSuccess))
    rc = GAL_ERR_HIP;
  if (rc == GAL_OK && status != 0) rc = GAL_ERR_INVALID;  // freed a failed malloc
  if (d_ops) (void)hipFree(d_ops);
  if (d_res) (void)hipFree(d_res);
  if (d_status) (void)hipFree(d_status);
  return rc;
}

int gal_group_replay_noalloc(gal_allocator* h, const gal_group_op* d_ops, uint64_t n, uint64_t* d_results) {
  if (!h || (n && (!d_ops || !d_results))) return GAL_ERR_INVALID;
  if (n > GAL_NOALLOC_MAX_OPS) return GAL_ERR_TOO_LARGE;
  if (!h->scratch) return GAL_ERR_STATE;
  if (n == 0) return GAL_OK;
  if (!is_device_pointer(d_ops) || !is_device_pointer(d_results)) return GAL_ERR_INVALID;
  // the status words and this call's part of the freed-result bitmap start at zero
  const uint64_t words = kScratchStatusWords + n;
  scratch_clear_kernel<<<blocks_for(words, 256), 256, 0, h->stream>>>(h->scratch, words);
  if (hipGetLastError() != hipSuccess) return GAL_ERR_HIP;
  int rc = h->ops->group_replay(h, d_ops, n, d_results, (uint32_t*)h->scratch, h->scratch + kScratchStatusWords);
  if (rc != GAL_OK) return rc;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  uint32_t status = 0;
  if (hipMemcpy(&status, h->scratch, sizeof(status), hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
  return status == 0 ? GAL_OK : GAL_ERR_INVALID;
}

int gal_stream(gal_allocator* h, void** stream) {
  if (!h || !stream) return GAL_ERR_INVALID;
  *stream = (void*)h->stream;
  return GAL_OK;
}

int gal_retain(gal_allocator* h) {
  if (!h) return GAL_ERR_INVALID;
  h->children++;
  return GAL_OK;
}

int gal_release(gal_allocator* h) {
  if (!h || h->children <= 0) return GAL_ERR_INVALID;
  h->children--;
  return GAL_OK;
}

int gal_read_pool(gal_allocator* h, uint64_t off, void* dst, uint64_t n) {
  if (!h || !dst) return GAL_ERR_INVALID;
  const uint64_t pool = h->info.num_segments * h->info.segment_size;
  if (off > pool || n > pool - off) return GAL_ERR_INVALID;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  return hipMemcpy(dst, (const char*)h->info.pool_device_ptr + off, n, hipMemcpyDeviceToHost) == hipSuccess
             ? GAL_OK
             : GAL_ERR_HIP;
}

int gal_gather_limit(const uint64_t* words, uint64_t nwords, uint64_t* segments) {
  if (!words || !segments || nwords == 0) return GAL_ERR_INVALID;
  uint64_t best = 0;
  for (uint64_t n = 1; n <= nwords * 64; n++)
    if (gather_would_succeed(words, nwords, n)) best = n;
  *segments = best;
  return GAL_OK;
}

int gal_largest_run(gal_allocator* h, uint64_t* segments) {
  if (!h || !segments) return GAL_ERR_INVALID;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  std::vector<uint64_t> words;
  const int rc = h->ops->segment_words(h, words);
  if (rc != GAL_OK) return rc;
  return gal_gather_limit(words.data(), words.size(), segments);
}

int gal_free_bytes(gal_allocator* h, uint64_t* bytes) {
  if (!h || !bytes) return GAL_ERR_INVALID;
  uint32_t segs = 0;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return GAL_ERR_HIP;
  if (hipMemcpy(&segs, ga::ctl_word(h->control, ga::kCtlFreeSegs), 4, hipMemcpyDeviceToHost) != hipSuccess) return GAL_ERR_HIP;
  *bytes = (uint64_t)segs * h->info.segment_size;
  return GAL_OK;
}

int gal_allocated_size(uint64_t segment_size, uint64_t min_size, uint64_t max_size, uint64_t size, uint64_t* bytes) {
  if (!bytes || size == 0) return GAL_ERR_INVALID;
  const impl_ops* ops = find_impl(segment_size, min_size, max_size);
  if (!ops) return GAL_ERR_UNSUPPORTED;
  *bytes = ops->allocated_size(size);
  return GAL_OK;
}

int gal_create_host(uint64_t segment_size, uint64_t min_size, uint64_t max_size, uint64_t pool_bytes, uint64_t seed,
                    gal_allocator** out) {
  if (!out) return GAL_ERR_INVALID;
  *out = nullptr;
  const impl_ops* ops = find_impl(segment_size, min_size, max_size);
  if (!ops) return GAL_ERR_UNSUPPORTED;
  if (pool_bytes < segment_size) return GAL_ERR_INVALID;
  void* dev = ops->create_host(pool_bytes, seed);
  if (!dev) return GAL_ERR_OOM;
  gal_allocator* h = nullptr;
  int rc = make_handle(ops, dev, &h);
  if (rc != GAL_OK) {
    ops->destroy(dev);
    return rc;
  }
  *out = h;
  return GAL_OK;
}

int gal_init_global_allocator(uint64_t num_bytes, uint64_t seed, int print_info) {
  if (g_global) return GAL_ERR_INVALID;
  if (!ga::init_global_allocator(num_bytes, seed, print_info != 0)) return GAL_ERR_OOM;
  const impl_ops* ops = find_impl(16 * MiB, 16, 4096);
  gal_allocator* h = nullptr;
  int rc = make_handle(ops, (void*)ga::get_global_allocator(), &h);
  if (rc != GAL_OK) {
    ga::free_global_allocator();
    return rc;
  }
  g_global = h;
  return GAL_OK;
}

int gal_free_global_allocator(void) {
  if (!g_global) return GAL_ERR_INVALID;
  if (g_global->children > 0) return GAL_ERR_STATE;
  ga::free_global_allocator();
  g_global->dev = nullptr;  // already released
  gal_destroy(g_global);
  g_global = nullptr;
  return GAL_OK;
}

int gal_print_global_stats(void) {
  if (!g_global) return GAL_ERR_INVALID;
  ga::print_global_stats();
  fflush(stdout);
  return GAL_OK;
}

int gal_init_global_allocator_host(uint64_t num_bytes, uint64_t seed, int print_info) {
  if (g_global_host) return GAL_ERR_INVALID;
  if (!ga::init_global_allocator_host(num_bytes, seed, print_info != 0)) return GAL_ERR_OOM;
  gal_allocator* h = nullptr;
  int rc = make_handle(find_impl(16 * MiB, 16, 4096), (void*)ga::get_global_allocator_host(), &h);
  if (rc != GAL_OK) {
    ga::free_global_allocator_host();
    return rc;
  }
  g_global_host = h;
  return GAL_OK;
}

int gal_free_global_allocator_host(void) {
  if (!g_global_host) return GAL_ERR_INVALID;
  if (g_global_host->children > 0) return GAL_ERR_STATE;
  ga::free_global_allocator_host();
  g_global_host->dev = nullptr;
  gal_destroy(g_global_host);
  g_global_host = nullptr;
  return GAL_OK;
}

int gal_print_global_stats_host(void) {
  if (!g_global_host) return GAL_ERR_INVALID;
  ga::print_global_stats_host();
  fflush(stdout);
  return GAL_OK;
}

int gal_init_global_allocator_combined(uint64_t num_bytes, uint64_t host_bytes, uint64_t seed, int print_info) {
  int rc = gal_init_global_allocator(num_bytes, seed, print_info);
  if (rc != GAL_OK) return rc;
  rc = gal_init_global_allocator_host(host_bytes, seed, print_info);
  if (rc != GAL_OK) gal_free_global_allocator();
  return rc;
}

int gal_free_global_allocator_combined(void) {
  if ((g_global && g_global->children > 0) || (g_global_host && g_global_host->children > 0)) return GAL_ERR_STATE;
  const int a = gal_free_global_allocator();
  const int b = gal_free_global_allocator_host();
  return a != GAL_OK ? a : b;
}

int gal_print_global_stats_combined(void) {
  if (!g_global || !g_global_host) return GAL_ERR_INVALID;
  ga::print_global_stats_combined();
  fflush(stdout);
  return GAL_OK;
}

int gal_global_handles(gal_allocator** device, gal_allocator** host) {
  if (device) *device = g_global;
  if (host) *host = g_global_host;
  return GAL_OK;
}

int gal_global_replay(const gal_op* ops, uint64_t n, uint64_t* results, uint32_t* cu_id) {
  if (!g_global || (n && (!ops || !results))) return GAL_ERR_INVALID;
  if (n == 0) return GAL_OK;
  for (uint64_t i = 0; i < n; i++) {
    if (ops[i].kind == GAL_OP_FREE) {
      if (ops[i].arg >= i || ops[ops[i].arg].kind != GAL_OP_MALLOC) return GAL_ERR_INVALID;
    } else if (ops[i].kind != GAL_OP_MALLOC || ops[i].arg == 0) {
      return GAL_ERR_INVALID;
    }
  }
  gal_op* d_ops = nullptr;
  uint64_t* d_res = nullptr;
  uint32_t* d_word = nullptr;  // status, CU id
  int rc = GAL_OK;
  if (hipMalloc(&d_ops, n * sizeof(gal_op)) != hipSuccess || hipMalloc(&d_res, n * 8) != hipSuccess ||
      hipMalloc(&d_word, 8) != hipSuccess)
    rc = GAL_ERR_OOM;
  if (rc == GAL_OK && (hipMemcpy(d_ops, ops, n * sizeof(gal_op), hipMemcpyHostToDevice) != hipSuccess ||
                       hipMemset(d_word, 0, 8) != hipSuccess))
    rc = GAL_ERR_HIP;
  if (rc == GAL_OK) {
    global_replay_kernel<<<1, 1, 0, g_global->stream>>>(d_ops, n, d_res, d_word, d_word + 1);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g_global->stream) != hipSuccess) rc = GAL_ERR_HIP;
  }
  uint32_t words[2] = {0, 0};
  if (rc == GAL_OK && (hipMemcpy(results, d_res, n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
                       hipMemcpy(words, d_word, 8, hipMemcpyDeviceToHost) != hipSuccess))
    rc = GAL_ERR_HIP;
  if (rc == GAL_OK && words[0] != 0) rc = GAL_ERR_INVALID;  // freed a failed malloc
  if (rc == GAL_OK && cu_id) *cu_id = words[1];
  if (d_ops) (void)hipFree(d_ops);
  if (d_res) (void)hipFree(d_res);
  if (d_word) (void)hipFree(d_word);
  return rc;
}

}  // extern "C"