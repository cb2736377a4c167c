// memory_table.hpp — blocks, per-segment counters and the recycle queues.
//
// Layout in HBM (one array each, all indexed by plain integers):
//   chunk_ids[num_segments]       u32: 0xFFFF free, c < T class c,
//                                 T+1+n first segment of an n-segment run
//   blocks[num_segments * BPS]    Block {malloc_counter, free_counter}, one
//                                 per 128-B line (8 B used)
//   queues[num_segments * BPS]    u32 recycled block id + 1 (0 = empty)
//   segs[num_segments]            segment_counters, 16 B
// BPS = SEG / (MIN * 4096) block slots per segment; a segment formatted for
// class c uses the first SEG / ((MIN << c) * 4096) of them.
// Same records as the reference's alloc_table (memory_table.cuh:229-1356) and
// Block (block.cuh:59-398); chunk ids widen to u32 so every update is a
// native 32-bit atomic, and queue entries are block indices, not pointers.
#pragma once

#include "device_utils.hpp"

namespace gallatin {
namespace allocators {

constexpr uint32_t kSlicesPerBlock = 4096;
// The snapshot image packs a block's malloc counter as (class << 20) | count,
// the reference's own encoding (block.cuh:215-224).
constexpr uint32_t kSnapshotClassShift = 20;
constexpr uint32_t kFreeChunk = 0xFFFFu;

// Each block's counters get a 128-byte line of their own. Returned atomics
// to one line are performed one after another (about 11 ns each on MI355X,
// whichever words of the line they target), while distinct lines proceed in
// parallel; the pinned blocks of a class are usually neighbours, so packed
// 8-byte records would put a whole class behind one line.
//
// malloc_counter holds the class in its high 32 bits and the slices handed
// out in its low 32 bits (the reference packs them as class << 20 | count in
// 32 bits). Every claim adds its group's slice total, also to a block that
// is already spent; with multi-slice requests one wave can add 4096 at once,
// so a few hundred late claims would carry a 20-bit count into the class
// bits and wrap it back into the valid range.
struct alignas(128) Block {
  uint64_t malloc_counter;  // (class << 32) | slices handed out
  uint32_t free_counter;    // slices returned (or written off) so far
  uint32_t unused;
};
static_assert(sizeof(Block) == 128, "one block per cache line");
__host__ __device__ constexpr uint64_t block_counter(uint32_t tree, uint32_t count) {
  return ((uint64_t)tree << 32) | count;
}
__host__ __device__ constexpr uint32_t counter_class(uint64_t m) { return (uint32_t)(m >> 32); }
__host__ __device__ constexpr uint32_t counter_count(uint64_t m) { return (uint32_t)m; }

struct segment_counters {
  int32_t active;        // block grants left (-1: none / segment not formatted)
  uint32_t queue_alloc;  // grants so far; >= nb means take from the queue
  uint32_t queue_free;   // returns so far (queue write position)
  uint32_t queue_final;  // returns whose queue write is complete
};

// device error counters (reference traps), kept in the control word array
enum control_word : uint32_t {
  kCtlLocks = 0,          // bit c: class c is attaching a segment
  kCtlErrors = 1,         // internal inconsistencies observed (sum of the sites below)
  kCtlBadFrees = 2,       // frees of pointers the allocator does not own
  kCtlBootStatus = 3,     // 0 booting, 1 ok, 2 failed
  kCtlSiteBase = 4,       // per-site counts of internal inconsistencies:
  kErrQueueEmpty = 0,     //   a grant found its recycle-queue slot empty
  kErrResetCount = 1,     //   a completed block's free count was not 4096 at reset
  kErrChunkReset = 2,     //   a released segment's class id changed under it
  kErrSetup = 3,          //   a free segment could not be formatted
  kErrQueueSpin = 4,      //   a queue write waited too long for its slot
  kErrFinalSpin = 5,      //   a queue write waited too long for earlier writes
  kErrFillClass = 6,      //   a whole-block grant carried another class
  kErrRunReset = 7,       //   a segment run's id changed under its free
  kErrFillCount = 8,      //   a whole-block grant found slices already claimed
  kErrInitCounter = 9,    //   a granted block's counter held the impossible value
  kErrGrantClass = 10,    //   a block was granted from a segment of another class
  kErrClaimClass = 11,    //   a slice was claimed from a segment of another class
  kErrFreeClass = 12,     //   a free hit a block whose counter carries another class
  kErrOverReturn = 13,    //   a segment got back more blocks than it granted
  kErrSites = 14,
  kCtlFreeSegs = 18,      // segments in the segment tree (free pool bytes / segment size); updated
                          // right after each claim from / return to the tree
  kCtlStatBase = 20,      // activity counters of the rare paths:
  kStatReplace = 0,       //   spent pinned blocks swapped out
  kStatRefill = 1,        //   empty pinned slots refilled
  kStatRequest = 2,       //   block requests (request_new_block_from_tree)
  kStatRequestRetry = 3,  //   block-request passes that lost a race
  kStatAttach = 4,        //   segments attached to a class
  kStatRelease = 5,       //   segments released to the segment tree
  kStatRecycle = 6,       //   blocks recycled after all slices came back
  kStatReplaceTime = 7,   //   wall-clock ticks / 16 spent in replacements
  kStatRefillTime = 8,    //   wall-clock ticks / 16 spent in refills
  kStatAhead = 9,         //   blocks fetched ahead for a slot
  kStatAheadTaken = 10,   //   replacements and refills served by a block fetched ahead
  kStatSites = 11,
  kCtlWords = 32
};
// Each control word has a 128-byte line of its own: the class locks, the
// free-segment gauge and the activity counters are all updated by the rare
// paths of many waves at once, and atomics to one line are performed one
// after another.
constexpr uint32_t kCtlStride = 32;  // u32 words per control word's line
__host__ __device__ inline uint32_t* ctl_word(uint32_t* control, uint32_t w) { return control + w * kCtlStride; }

#if defined(__HIPCC__)
__device__ inline void count_error_at(uint32_t* control, uint32_t site) {
  atom::add(ctl_word(control, kCtlErrors), 1u);
  atom::add(ctl_word(control, kCtlSiteBase + site), 1u);
}
__device__ inline void count_stat(uint32_t* control, uint32_t stat, uint32_t n = 1u) {
  atom::add(ctl_word(control, kCtlStatBase + stat), n);
}
#endif

template <uint64_t SEG, uint64_t MIN>
struct alloc_table {
  static constexpr uint64_t blocks_per_segment = SEG / (MIN * 4096);
  static constexpr uint64_t max_allocations_per_segment = SEG / MIN;

  uint32_t* chunk_ids;
  Block* blocks;
  uint32_t* queues;
  segment_counters* segs;
  char* memory;
  uint64_t num_segments;

  __host__ __device__ static constexpr uint64_t tree_alloc_size(uint32_t tree) { return MIN << tree; }
  __host__ __device__ static constexpr uint64_t blocks_in_segment(uint32_t tree) {
    return SEG / ((MIN << tree) * 4096);
  }

#if defined(__HIPCC__)
  __device__ uint32_t read_tree_id(uint64_t segment) const { return atom::load(&chunk_ids[segment]); }
  __device__ bool set_tree_id(uint64_t segment, uint32_t id) {
    return atom::cas(&chunk_ids[segment], kFreeChunk, id) == kFreeChunk;
  }
  __device__ bool reset_tree_id(uint64_t segment, uint32_t id) {
    return atom::cas(&chunk_ids[segment], id, kFreeChunk) == id;
  }

  // format a free segment for class `tree` (reference memory_table.cuh:746-818)
  __device__ bool setup_segment(uint64_t segment, uint32_t tree) {
    if (!set_tree_id(segment, tree)) return false;
    const uint64_t nb = blocks_in_segment(tree);
    uint32_t* q = queues + segment * blocks_per_segment;
    for (uint64_t i = 0; i < nb; i++) atom::store(&q[i], 0u);
    segment_counters* s = &segs[segment];
    atom::exch(&s->queue_alloc, 0u);
    atom::exch(&s->queue_free, 0u);
    atom::exch(&s->queue_final, 0u);
    // grants open only after the queue state above is in place
    atom::order();
    atom::exch(&s->active, (int32_t)nb - 1);
    return true;
  }

  // Grant one block of `segment` to class `tree` (reference
  // memory_table.cuh:915-993): fresh blocks in order first, then recycled
  // blocks in the order they were returned. Returns the global block id, or
  // -1 when the segment has nothing to grant or changed class. `last` is set
  // when this grant took the segment's final block.
  __device__ int64_t get_block(uint64_t segment, uint32_t tree, bool& last, uint32_t* control) {
    last = false;
    segment_counters* s = &segs[segment];
    // Take one grant without ever driving the count below -1: an undo
    // increment after a failed decrement could land after setup_segment has
    // reset the count for the segment's next use and over-admit it.
    int32_t active_count = atom::load(&s->active);
    while (active_count >= 0) {
      const int32_t seen = atom::cas(&s->active, active_count, active_count - 1);
      if (seen == active_count) break;
      active_count = seen;
    }
    if (active_count < 0) return -1;
    if (read_tree_id(segment) != tree) {
      atom::add(&s->active, 1);
      return -1;
    }
    const uint64_t nb = blocks_in_segment(tree);
    int32_t queue_pos = (int32_t)atom::add(&s->queue_alloc, 1u);
    int64_t block;
    if ((uint64_t)queue_pos < nb) {
      block = (int64_t)(segment * blocks_per_segment + (uint64_t)queue_pos);
    } else {
      uint32_t v = atom::exch(&queues[segment * blocks_per_segment + (uint64_t)queue_pos % nb], 0u);
      if (v == 0u) {  // the reference traps here; count it instead
        count_error_at(control, kErrQueueEmpty);
        return -1;
      }
      block = (int64_t)v - 1;
    }
    // Block::init_malloc. The old value is read so that the reset is
    // performed before the block can be published anywhere.
    const uint64_t old_counter = atom::exch(&blocks[block].malloc_counter, block_counter(tree, 0));
    if (old_counter == ~0ull) count_error_at(control, kErrInitCounter);
    if (read_tree_id(segment) != tree) count_error_at(control, kErrGrantClass);
    if (active_count == 0) last = true;
    return block;
  }

  // Enqueue a fully returned block; returns its queue slot (reference
  // memory_table.cuh:1158-1222). Slot writes complete in reservation order,
  // which is what lets get_block trust the queue entry it is sent to.
  __device__ uint32_t reserve_segment_slot(uint64_t block, uint64_t segment, uint64_t nb, uint32_t* control) {
    segment_counters* s = &segs[segment];
    uint32_t pos = atom::add(&s->queue_free, 1u);
    uint32_t live = (uint32_t)(pos % nb);
    // The slot still holds the entry of `pos - nb` when the grant that owns
    // that entry has taken its queue position but not yet read the slot: an
    // exchange would overwrite (lose) that block. Wait for the read instead.
    // The reader is never in this wave (a wave runs one allocator path at a
    // time and get_block reads its slot without waiting on anything).
    uint32_t* q = &queues[segment * blocks_per_segment + live];
    for (uint32_t spin = 0; atom::cas(q, 0u, (uint32_t)(block + 1)) != 0u; spin++) {
      if (spin > (1u << 24)) {
        count_error_at(control, kErrQueueSpin);
        atom::exch(q, (uint32_t)(block + 1));
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    // Earlier reservations are held by other waves (a wave completes its own
    // before it can take another), so this wait always ends; the bound only
    // turns a broken invariant into a counted error instead of a hang.
    for (uint32_t spin = 0; atom::cas(&s->queue_final, pos, pos + 1) != pos; spin++) {
      if (spin > (1u << 24)) {
        count_error_at(control, kErrFinalSpin);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    return live;
  }

  // reference memory_table.cuh:1226-1241; true when this return makes every
  // block of the segment free and this lane won the right to release it
  __device__ bool finish_freeing_block(uint64_t segment, uint64_t nb, uint32_t* control) {
    atom::order();  // the queue entry is visible before the grant is
    int32_t return_id = atom::add(&segs[segment].active, 1);
    if ((int64_t)return_id > (int64_t)nb - 2) count_error_at(control, kErrOverReturn);
    if ((int64_t)return_id == (int64_t)nb - 2) {
      if (atom::cas(&segs[segment].active, (int32_t)nb - 1, (int32_t)-1) == (int32_t)nb - 1) return true;
    }
    return false;
  }
#endif  // __HIPCC__
};

}  // namespace allocators
}  // namespace gallatin
