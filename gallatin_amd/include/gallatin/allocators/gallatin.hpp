// gallatin.hpp — Gallatin<bytes_per_segment, smallest, biggest> for MI355X.
//
// Drop-in for the reference's allocator (include/gallatin/allocators/
// gallatin.cuh:292-2035): host construction with generate_on_device /
// free_on_device / print_info, and in-kernel malloc / free callable from any
// lane under any divergence.
//
// A request of `size` bytes goes one of three ways (reference :1186-1276):
//  * slice:   size <= biggest          -> one slice of class ceil(log2(size))-log2(smallest)
//             size <= smallest*2048    -> 2^k contiguous slices of the top class
//  * block:   size <= biggest*4096     -> a whole 4096-slice block of class
//                                         ceil(log2(size)) - log2(smallest*4096)
//  * segment: larger                   -> ceil(size / bytes_per_segment) contiguous segments
// Slices are bump-claimed from per-class "pinned" blocks. Lanes of one
// wavefront asking for the same class form a group: the group's last lane
// makes ONE atomic add on the pinned block's malloc counter for the whole
// group, and each lane takes its exclusive prefix of the group's slice count.
// Frees are grouped the same way by block: one atomic add per (wave, block).
// A block whose 4096 slices have all come back goes to its segment's recycle
// queue; a segment whose blocks have all come back goes to the segment tree.
#pragma once

#include "device_utils.hpp"
#include "memory_table.hpp"
#include "veb.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

namespace gallatin {
namespace allocators {

// Retry bounds. The reference's values (gallatin.cuh:108-112) where a retry
// can see new state; out-of-memory retries are capped lower because with one
// lane (or no concurrent frees) a retry repeats an identical failure.
constexpr int kMallocLoopAttempts = 10;     // GALLATIN_MALLOC_LOOP_ATTEMPTS
constexpr int kSliceAttempts = 500;         // GALLATIN_MAX_ATTEMPTS
constexpr int kBusyPasses = 1 << 16;        // safety bound on passes that wait for a busy slot
constexpr int kRequestBlockAttempts = 1000; // REQUEST_BLOCK_MAX_ATTEMPTS
constexpr int kRequestBlockPasses = 1 << 20; // safety bound on lost-race retries
constexpr int kRequestBlockOomAttempts = 2; // failures with no segment anywhere
constexpr int kBlockAttempts = 4;           // GALLATIN_MALLOC_BLOCK_ATTEMPTS (500) on OOM
constexpr int kSegmentAttempts = 4;         // GALLATIN_MALLOC_SEGMENT_ATTEMPTS (500) on OOM
constexpr int kPinnedWalk = 30;             // SHARED_BLOCK_COUNTER_CUTOFF
constexpr int kDirectAttachPass = 2;        // passes a block request waits on a held class lock ...
constexpr uint64_t kDirectAttachBlocks = 4; // ... before attaching its own segment (classes with <= 4 blocks per segment)
constexpr uint32_t kPinnedSlotsClass0 = 128;
// a pinned slot holds 0 (empty: a replacement ran out of memory), a block id
// + 1, or kSlotBusy while one lane replaces or refills it
constexpr uint32_t kSlotBusy = 0xFFFFFFFFu;
constexpr uint32_t kPinnedSlotsMin = 4;     // MIN_PINNED_CUTOFF
// Classes with fewer than kSlotsUnderLoad pinned slots get overflow slots up
// to that count, which a single lane never uses. Under load (more than one
// lane of a wave looking) a group starts at one of its class's usable slots,
// pinned or overflow, and a group that finds every pinned slot busy looks at
// the overflow ones: the claims of a class then spread over many block
// counters instead of queueing on a few (same-address atomics on one line
// run at about 1e8/s, and every wave has a lane of every class in a mixed
// workload), and replacements of one class overlap. A class's usable slots
// hold at most 1/kSlotShare of the pool in blocks (at least its pinned slots),
// so that a small pool is not tied up in partly used blocks.
constexpr uint32_t kSlotsUnderLoad = 64;
// u32 words between slot entries: two 128-B lines per slot, the slot's entry
// (word 0) and the block fetched ahead for it (word kAheadWord)
constexpr uint32_t kSlotStride = 64;
constexpr uint32_t kAheadWord = 32;
// Under load, the group whose claim crosses this slice of a slot's block
// fetches the slot's next block ahead of time (see fetch_ahead).
constexpr uint32_t kAheadAt = 2048;
// ORed into the class argument of replace_block: only fetch the slot's next
// block (the slot's block is not spent). A flag bit rather than a second
// call site or one more argument, either of which made the claim path spill
// a register.
constexpr uint32_t kFetchOnly = 1u << 31;
constexpr uint64_t kSlotShare = 16;

// Where the pool lives (reference memory_table.cuh:61). Metadata always lives
// in device memory; only the pool moves. host_only: pinned host memory mapped
// into the GPU's address space (hipHostMalloc, mapped); allocations are then
// reached over the host link. managed is not provided (HIP managed memory
// needs XNACK to migrate pages).
enum Gallatin_memory_type { device_only, host_only, managed };

__host__ __device__ constexpr int ilog2(uint64_t x) { return x <= 1 ? 0 : 1 + ilog2(x >> 1); }
__host__ __device__ constexpr bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

// pinned slots of class c: 128 >> c, at least 4 (reference
// shared_block_storage.cuh:251-274)
__host__ __device__ constexpr uint32_t pinned_slots_of(uint32_t c) {
  return (kPinnedSlotsClass0 >> (c > 31 ? 31 : c)) < kPinnedSlotsMin ? kPinnedSlotsMin : (kPinnedSlotsClass0 >> c);
}
__host__ __device__ constexpr uint32_t overflow_slots_of(uint32_t c) {
  return pinned_slots_of(c) < kSlotsUnderLoad ? kSlotsUnderLoad - pinned_slots_of(c) : 0u;
}
// slot region of class c: its pinned slots, then its overflow slots
__host__ __device__ constexpr uint32_t slot_region_of(uint32_t c) { return pinned_slots_of(c) + overflow_slots_of(c); }
// slots of class c usable under load, for a pool of pool_bytes whose class-c
// blocks are block_bytes each
__host__ __device__ constexpr uint32_t usable_slots_of(uint32_t c, uint64_t pool_bytes, uint64_t block_bytes) {
  const uint64_t budget = pool_bytes / kSlotShare / block_bytes;
  const uint64_t lo = pinned_slots_of(c), hi = slot_region_of(c);
  return (uint32_t)(budget < lo ? lo : budget > hi ? hi : budget);
}

template <uint64_t bytes_per_segment, uint64_t smallest, uint64_t biggest>
struct Gallatin {
  using my_type = Gallatin<bytes_per_segment, smallest, biggest>;
  using table_type = alloc_table<bytes_per_segment, smallest>;

  static_assert(is_pow2(bytes_per_segment) && is_pow2(smallest) && is_pow2(biggest), "powers of two");
  static_assert(smallest <= biggest, "smallest <= biggest");
  static_assert(bytes_per_segment >= biggest * 4096, "a segment holds at least one block of the top class");

  static constexpr uint64_t kBytesPerSegment = bytes_per_segment;
  static constexpr uint32_t kNumTrees = (uint32_t)(ilog2(biggest) - ilog2(smallest) + 1);
  static constexpr int kSmallestBits = ilog2(smallest);
  static constexpr int kSmallestTreeBits = ilog2(smallest * 4096);
  static constexpr uint64_t kBlocksPerSegment = table_type::blocks_per_segment;
  static_assert(kNumTrees <= 30, "the class lock bits fit one u32");

  static constexpr uint32_t total_slots() {
    uint32_t t = 0;
    for (uint32_t c = 0; c < kNumTrees; c++) t += slot_region_of(c);
    return t;
  }

  // ---- device-resident state (immutable after construction except through
  // the pointers; read with ordinary loads)
  veb_tree segment_tree;          // free segments
  veb_tree sub_trees[kNumTrees];  // segments formatted for each class
  table_type table;
  // pinned blocks, class-major (slot i of class c is entry slot_offset[c] + i):
  // block id + 1. Each entry has a 128-B line of its own: every claim reads
  // its group's slot while replacements swap slots with atomics, and atomics
  // to one line are performed one after another (packed, the 64 slots of a
  // class shared two lines, and a filling round of the mixed workload took
  // 1.4 ms instead of 1.0 ms).
  uint32_t* slots;
  uint32_t slot_offset[kNumTrees];
  uint32_t usable_slots[kNumTrees];  // usable_slots_of each class for this pool
  uint32_t* control;              // control_word array
  // per class, one 128-B line; word 0: how many of the class's pinned slots
  // hold 0 (a replacement or refill found no memory). Changed only on those
  // rare paths, lowered before and raised after the slot itself changes, so
  // it never exceeds the true count (read as signed: it can dip below 0).
  int32_t* class_words;
  uint64_t seed;
  // device allocations, for free_on_device
  void* owned[6];
  uint32_t memory_kind;  // Gallatin_memory_type of the pool

#if defined(__HIPCC__)
  // ------------------------------------------------------------ size classes
  __host__ __device__ static constexpr uint32_t num_trees() { return kNumTrees; }
  __device__ static uint32_t get_tree_id_from_size(uint64_t size) {
    if (size < smallest) return 0;
    return (uint32_t)(uint16_t)(utils::get_first_bit_bigger(size) - kSmallestBits);
  }

  __device__ uint32_t* lock_word() const { return ctl_word(control, kCtlLocks); }
  __device__ bool acquire_tree_lock(uint32_t tree) {
    uint32_t m = 1u << tree;
    return (atom::fetch_or(lock_word(), m) & m) == 0;
  }
  __device__ void release_tree_lock(uint32_t tree) { atom::fetch_and(lock_word(), ~(1u << tree)); }
  __device__ void count_error(uint32_t site) { count_error_at(control, site); }

  __device__ uint32_t* slot_ptr(uint32_t tree, uint32_t index) const {
    return &slots[(uint64_t)(slot_offset[tree] + index) * kSlotStride];
  }
  // the block fetched ahead for a slot: 0 none, block id + 1, or kSlotBusy
  // while one lane fetches it
  __device__ uint32_t* ahead_ptr(uint32_t tree, uint32_t index) const {
    return &slots[(uint64_t)(slot_offset[tree] + index) * kSlotStride + kAheadWord];
  }
  __device__ int32_t* empty_pinned_word(uint32_t tree) const { return class_words + tree * kCtlStride; }

  // ------------------------------------------------- segments and blocks
  // Nothing for malloc_first to claim in the segment tree. The free-segment
  // gauge answers first (one load): a segment leaves the tree before the
  // gauge is lowered and joins it before the gauge is raised, so while the
  // gauge counts one the tree has one or is about to, and a claim attempt
  // may follow (it fails cleanly if the segment went). Only a gauge of 0 is
  // confirmed by the walk over the tree (a segment may have joined it an
  // instant before the gauge). With one lane the gauge is exact and the
  // answer is the walk's. In a burst of block requests the walk, run by
  // every request pass, took a third of a request's time.
  __device__ bool no_free_segment() const {
    return atom::load(ctl_word(control, kCtlFreeSegs)) == 0u && segment_tree.nothing_to_claim();
  }

  // attach the lowest free segment to class `tree` (reference :1465-1510)
  __device__ __noinline__ int64_t gather_new_segment(uint32_t tree) {
    uint64_t segment = segment_tree.malloc_first();
    if (segment == utils::kFail) return -1;
    atom::sub(ctl_word(control, kCtlFreeSegs), 1u);
    if (!table.setup_segment(segment, tree)) {
      count_error(kErrSetup);
      return -1;
    }
    count_stat(control, kStatAttach);
    atom::order();
    sub_trees[tree].insert_force_update(segment);
    atom::order();
    return (int64_t)segment;
  }

  // The reference's eager step after a segment's last block is granted
  // (:1595-1617): attach a fresh segment to the class. The grant's caller
  // runs it once the block is published, so that a pinned slot is not held
  // busy for it; with one lane the state after the call is the same.
  __device__ __noinline__ void gather_after_last(uint32_t tree) {
    if (acquire_tree_lock(tree)) {
      gather_new_segment(tree);
      release_tree_lock(tree);
    }
    atom::order();
  }

  // find a block for class `tree`, attaching a segment when the class has
  // none with free blocks (reference :1525-1631); tid seeds the hashed
  // search. `last` reports that the block was its segment's last, after
  // which the caller runs gather_after_last.
  __device__ __noinline__ int64_t request_new_block_from_tree(uint32_t tree, uint64_t tid, bool& last) {
    last = false;
    int oom = 0;
    // Like the reference, only passes that fail for lack of memory count
    // toward giving up; passes that lose a race (the class lock is held, the
    // segment found is exhausted) retry, up to a large safety bound.
    int attempts = 0;
    count_stat(control, kStatRequest);
    for (int pass = 0; attempts < kRequestBlockAttempts && pass < kRequestBlockPasses; pass++) {
      if (pass > 0) count_stat(control, kStatRequestRetry);
      uint64_t segment = sub_trees[tree].find_random_valid_index(tid);
      if (segment == utils::kFail) {
        // out of memory is confirmed without the class lock: gather_new_segment
        // would fail the same way (its malloc_first sees exactly this)
        if (no_free_segment()) {
          attempts++;
          if (++oom >= kRequestBlockOomAttempts) return -1;
          continue;
        }
        const bool locked = acquire_tree_lock(tree);
        // A class with few blocks per segment needs a fresh segment for
        // nearly every block; after a pass spent waiting for the lane that
        // holds the class lock, a request attaches a segment of its own
        // instead of queueing behind it (it over-attaches at most that
        // class's few blocks). A single lane always gets the lock.
        const bool direct = !locked && pass >= kDirectAttachPass &&
                            table_type::blocks_in_segment(tree) <= kDirectAttachBlocks;
        if (locked || direct) {
          int64_t got = gather_new_segment(tree);
          if (locked) release_tree_lock(tree);
          if (got == -1) {
            attempts++;
            if (++oom >= kRequestBlockOomAttempts) return -1;
            continue;
          }
          segment = (uint64_t)got;
        } else {
          wave::backoff(pass);  // another lane is attaching a segment
          continue;
        }
      }
      atom::order();
      bool was_last = false;
      int64_t block = table.get_block(segment, tree, was_last, control);
      if (was_last) {
        sub_trees[tree].remove(segment);
        if (block < 0) {
          gather_after_last(tree);
        } else {
          last = true;
          return block;
        }
      }
      if (block >= 0) return block;
    }
    return -1;
  }

  // return a fully freed block to its segment (reference :1637-1726)
  __device__ __noinline__ void return_block(uint64_t block, uint64_t segment, uint32_t tree) {
    // everything this wave did to the block (its counter reset above all) is
    // performed before the block becomes grantable again
    atom::order();
    const uint64_t nb = table_type::blocks_in_segment(tree);
    const uint32_t slot = table.reserve_segment_slot(block, segment, nb, control);
    // 1.0*slot/nb >= 0.1 && 1.0*(slot-1)/nb < 0.1 for power-of-two nb
    // (reference :1646): re-list the segment once a tenth of it is back
    if (slot >= 1 && 10ull * slot > nb && 10ull * (slot - 1) < nb) {
      sub_trees[tree].insert_force_update(segment);
      atom::order();
    }
    const bool release = table.finish_freeing_block(segment, nb, control);
    atom::order();
    if (release) {
      count_stat(control, kStatRelease);
      sub_trees[tree].remove(segment);
      atom::order();
      if (!table.reset_tree_id(segment, tree)) count_error(kErrChunkReset);
      atom::order();
      segment_tree.insert_force_update(segment);
      atom::add(ctl_word(control, kCtlFreeSegs), 1u);
    }
  }

  __device__ void free_block(uint64_t block) {
    const uint64_t segment = block / kBlocksPerSegment;
    return_block(block, segment, table.read_tree_id(segment));
  }

  // account n returned slices of `block` (one lane); the return that
  // completes 4096 recycles the block (reference block.cuh:87-98,
  // gallatin.cuh:1790-1799)
  __device__ void free_slices(uint64_t block, uint32_t n) {
    const uint32_t old = atom::add(&table.blocks[block].free_counter, n);
    if (old + n == kSlicesPerBlock) recycle_block(block);
  }

  // reset the free count of a block whose 4096 slices are all back and hand
  // the block to its segment (reference block.cuh:200-212 + free_block)
  __device__ __noinline__ void recycle_block(uint64_t block) {
    // the reset's old value is read: the reset must be performed before the
    // block is queued, or a late reset would wipe its next owner's count
    const uint32_t prev = atom::exch(&table.blocks[block].free_counter, 0u);
    if (prev != kSlicesPerBlock) count_error(kErrResetCount);
    count_stat(control, kStatRecycle);
    free_block(block);
  }

  // Blocks fetched ahead (no counterpart in the reference, whose slot is
  // busy for the whole block search of a replacement). Under load, the group
  // whose claim crosses slice kAheadAt of a slot's block fetches the slot's
  // next block into the slot's ahead word while the slot keeps serving; the
  // replacement that finds the block spent publishes that block (the slot is
  // busy for a few memory round trips instead of a search) and fetches the
  // next one. An ahead word changes 0 -> busy (the one lane that fetches) ->
  // block + 1 or 0 (that lane) -> 0 (the lane whose compare-and-swap takes
  // it), so a block fetched ahead is published once. A replacement that
  // finds no memory takes the block fetched ahead for another slot of the
  // class before it leaves its slot empty, so those blocks are spent before
  // the class runs dry. One lane never fetches ahead, so single-lane results
  // are the reference's (their ahead words stay 0).
  //
  // Take the block `seen` in slot (tree, index)'s ahead word; 0 if it is not
  // a block or another lane took it first.
  __device__ uint32_t take_ahead(uint32_t tree, uint32_t index, uint32_t seen) {
    if (seen == 0u || seen == kSlotBusy || atom::cas(ahead_ptr(tree, index), seen, 0u) != seen) return 0u;
    count_stat(control, kStatAheadTaken);
    return seen;
  }

  // A block fetched ahead for any usable slot of the class, taken; 0 if none.
  // For a slot about to be left empty: the class has no other memory.
  __device__ uint32_t take_any_ahead(uint32_t tree) {
    const uint32_t n = usable_slots[tree];
    for (uint32_t base = 0; base < n; base += 8) {
      uint32_t seen[8];
#pragma unroll
      for (uint32_t k = 0; k < 8; k++) seen[k] = base + k < n ? atom::load(ahead_ptr(tree, base + k)) : 0u;
      for (uint32_t k = 0; k < 8; k++) {
        const uint32_t got = take_ahead(tree, base + k, seen[k]);
        if (got != 0u) return got;
      }
    }
    return 0u;
  }

  // Fetch the next block of slot (tree, index) ahead of time, unless one is
  // ready or being fetched. A slot that ran empty meanwhile (its replacement
  // found no block before this one arrived) gets the block directly.
  __device__ __noinline__ void fetch_ahead(uint32_t tree, uint32_t index, uint64_t tid) {
    // No free segment and nothing listed in the class's tree: the search
    // would fail.
    const uint64_t top = sub_trees[tree].top_word();
    if (atom::load(ctl_word(control, kCtlFreeSegs)) == 0u && sub_trees[tree].top_looks_empty(top)) return;
    uint32_t* a = ahead_ptr(tree, index);
    if (atom::cas(a, 0u, kSlotBusy) != 0u) return;
    count_stat(control, kStatAhead);
    bool last = false;
    const int64_t fresh = request_new_block_from_tree(tree, tid, last);
    atom::order();  // the block's counter reset precedes its publication
    bool placed = false;
    uint32_t* s = slot_ptr(tree, index);
    if (fresh >= 0 && atom::load(s) == 0u) {
      const bool pinned = index < pinned_slots_of(tree);
      if (pinned) atom::sub(empty_pinned_word(tree), 1);
      placed = atom::cas(s, 0u, (uint32_t)(fresh + 1)) == 0u;
      if (!placed && pinned) atom::add(empty_pinned_word(tree), 1);
    }
    atom::exch(a, fresh < 0 || placed ? 0u : (uint32_t)(fresh + 1));
    if (last) gather_after_last(tree);
  }

  // Publish the result of a slot's block search (the slot is held busy):
  // the block found, else a block fetched ahead for a slot of the class (one
  // may have arrived during the search), else 0 (empty). Returns whether the
  // slot got a block.
  __device__ bool publish_searched(uint32_t tree, uint32_t index, int64_t fresh) {
    const uint32_t entry = fresh >= 0 ? (uint32_t)(fresh + 1) : take_any_ahead(tree);
    atom::exch(slot_ptr(tree, index), entry);
    if (entry == 0u && index < pinned_slots_of(tree)) atom::add(empty_pinned_word(tree), 1);
    return entry != 0u;
  }

  // Swap a spent pinned block for a fresh one (reference :1083-1138). The
  // slot is marked busy while the new block is found, so that other groups
  // wait for it instead of racing to refill it; it is left empty when memory
  // runs out. A block fetched ahead for the slot is published without a
  // search, and the slot's next block is fetched. With kFetchOnly in
  // `tree_flags` it only fetches the slot's next block (the claim that
  // crossed slice kAheadAt, under load).
  __device__ __noinline__ bool replace_block(uint32_t tree_flags, uint32_t index, uint64_t block, uint64_t tid) {
    const uint32_t tree = tree_flags & ~kFetchOnly;
    if (tree_flags & kFetchOnly) {
      fetch_ahead(tree, index, tid);
      return true;
    }
    uint32_t* s = slot_ptr(tree, index);
    if (atom::cas(s, (uint32_t)(block + 1), kSlotBusy) != (uint32_t)(block + 1)) return true;
    // The slot may hold a later incarnation of the same block: between the
    // claim that found it spent and this swap, the block can be freed,
    // recycled, granted again and published in this very slot. Only a block
    // that is spent (4096 slices handed out) or serves another class is
    // replaced; a live one is put back. With one lane the block is always
    // the spent one. The ahead word is read in the same round trip.
    const uint64_t now = atom::load(&table.blocks[block].malloc_counter);
    const uint32_t ahead = atom::load(ahead_ptr(tree, index));
    if (counter_class(now) == tree && counter_count(now) < kSlicesPerBlock) {
      atom::exch(s, (uint32_t)(block + 1));
      return true;
    }
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    count_stat(control, kStatReplace);
    const uint32_t taken = take_ahead(tree, index, ahead);
    if (taken != 0u) {
      atom::exch(s, taken);
      count_stat(control, kStatReplaceTime, (uint32_t)((__builtin_amdgcn_s_memrealtime() - t0) >> 4));
      fetch_ahead(tree, index, tid);  // the slot's next block
      return true;
    }
    atom::order();
    bool last = false;
    const int64_t fresh = request_new_block_from_tree(tree, tid, last);
    atom::order();  // the block's counter reset precedes its publication
    const bool got = publish_searched(tree, index, fresh);
    count_stat(control, kStatReplaceTime, (uint32_t)((__builtin_amdgcn_s_memrealtime() - t0) >> 4));
    if (last) gather_after_last(tree);
    return got;
  }

  // refill an empty slot (left so by an out-of-memory replace); false when
  // memory is still short
  __device__ __noinline__ bool refill_slot(uint32_t tree, uint32_t index, uint64_t tid) {
    uint32_t* s = slot_ptr(tree, index);
    const bool pinned = index < pinned_slots_of(tree);
    if (pinned) atom::sub(empty_pinned_word(tree), 1);
    if (atom::cas(s, 0u, kSlotBusy) != 0u) {  // another group is on it
      if (pinned) atom::add(empty_pinned_word(tree), 1);
      return true;
    }
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    count_stat(control, kStatRefill);
    const uint32_t taken = take_ahead(tree, index, atom::load(ahead_ptr(tree, index)));
    if (taken != 0u) {
      atom::exch(s, taken);
      return true;
    }
    bool last = false;
    const int64_t fresh = request_new_block_from_tree(tree, tid, last);
    atom::order();
    const bool got = publish_searched(tree, index, fresh);
    count_stat(control, kStatRefillTime, (uint32_t)((__builtin_amdgcn_s_memrealtime() - t0) >> 4));
    if (last) gather_after_last(tree);
    return got;
  }

  // nothing a block request seeded with `tid` could find: its hashed search
  // of the class's segments fails and no free segment is left (with one
  // lane, exactly when request_new_block_from_tree fails, which changes no
  // state). The search, not an emptiness test, decides: a class whose only
  // listed segment is segment 0 is invisible to hashed starts above 0
  // (reference veb.cuh:960-975), and a stale summary bit can make an empty
  // tree look occupied; either way a refill could only fail, and every wave
  // of the class would queue for it.
  __device__ bool class_out_of_memory(uint32_t tree, uint64_t tid) const {
    return no_free_segment() && sub_trees[tree].find_random_valid_index(tid) == utils::kFail;
  }

  // The pinned slot a group uses, by the reference's walk
  // (shared_block_storage.cuh:119-145): start at hint % n and step past slots
  // that are empty or busy, at most 30 steps, ending after a full cycle once a
  // busy slot has been seen (then reporting busy: the caller tries the
  // overflow slots; a single lane never sees a busy slot); returns the slot's
  // index and the entry read there. The lanes of `group` load consecutive
  // positions of the walk at once (lane of rank r: position base + r) and
  // the first position where the sequential walk stops is found with ballots
  // (an entry holding a block; the walk's last position; a busy entry seen
  // and a full cycle walked), so every lane of the group gets the sequential
  // walk's result. A group of one lane walks one position at a time. Called
  // by all lanes of `looking`; each group walks its own class.
  __device__ void pinned_slot_group(uint32_t tree, uint32_t hint, uint64_t group, uint64_t looking,
                                    uint32_t& index, uint32_t& entry) const {
    const uint32_t n = pinned_slots_of(tree);
    const uint32_t original = hint % n;
    const uint32_t cap = (uint32_t)kPinnedWalk;
    const uint32_t last = original >= 1 && n - 1 < cap ? n - 1 : cap;  // the walk's last position
    const uint32_t g = wave::popc(group);
    const uint32_t rank = wave::popc(group & wave::lanemask_lt());
    const uint64_t at_or_below = wave::lanemask_lt() | (1ull << wave::lane_id());
    bool saw_busy = false;
    bool found = ((looking >> wave::lane_id()) & 1ull) == 0;  // lanes outside `looking` only vote
    index = original;
    entry = 0u;
    for (uint32_t base = 0; wave::ballot(!found) & looking; base += g) {
      const uint32_t k = base + rank;
      const bool in = !found && k <= last;
      const uint32_t pos = (original + k) % n;
      const uint32_t e = in ? atom::load(slot_ptr(tree, pos)) : 0u;
      const uint64_t busy_m = wave::ballot(in && e == kSlotBusy) & group;
      const bool busy_so_far = saw_busy || (busy_m & at_or_below) != 0;
      const bool stop = in && ((e != 0u && e != kSlotBusy) || k == last || (busy_so_far && k + 1 >= n));
      const uint64_t stop_m = wave::ballot(stop) & group;
      const int src = stop_m ? wave::first_lane(stop_m) : (int)wave::lane_id();
      const uint32_t s_pos = wave::shfl(pos, src);
      const uint32_t s_e = wave::shfl(e, src);
      const uint32_t s_busy = wave::shfl((uint32_t)busy_so_far, src);
      if (!found && stop_m) {
        index = s_pos;
        entry = (s_busy && (s_e == 0u || s_e == kSlotBusy)) ? kSlotBusy : s_e;
        found = true;
      }
      saw_busy = saw_busy || busy_m != 0;
    }
  }

  // ------------------------------------------------------- slice claims
  // One claim pass for the lanes of `looking` (the caller's exec mask),
  // reference gallatin.cuh:865-1078. Lanes asking for the same class form a
  // group; ALL groups of the wave claim at once: one slot-load instruction,
  // one atomic instruction in which each group's last lane adds the group's
  // slice total to its own block's malloc counter, and a lane shuffle that
  // hands each lane its group's old count. Rare per-group work (refilling an
  // empty slot, replacing a spent block, recycling a written-off block,
  // returning slices drawn from a stale block) runs one group at a time.
  // `subgroup` splits a class's lanes further (lanes with equal class and
  // subgroup form one group); malloc() passes 0, so a group is every lane of
  // the class. `allocs` is the number of allocations a lane's `count` slices
  // stand for: 1 for malloc, the run length for a bulk claim. `spread`
  // false keeps every group on the reference's pinned-slot walk (no start at
  // an overflow slot), for deterministic multi-lane replays. `ahead` lets a
  // group of a multi-lane claim fetch its slot's next block ahead of time
  // (fetch_ahead); the deterministic replays of the spread form run without.
  template <bool spread = true, bool ahead = spread>
  __device__ __forceinline__ void claim_pass(uint64_t looking, uint32_t tree, uint32_t count, uint32_t allocs,
                                             uint64_t tid, uint32_t hint, uint32_t subgroup, uint32_t attempt,
                                             uint64_t& result, bool& done, bool& oom, bool& waited) {
    const uint32_t me = wave::lane_id();
    const uint64_t group = wave::match_groups(looking, (uint64_t)tree | ((uint64_t)subgroup << 32));
    const int leader = wave::first_lane(group);
    const int last = wave::last_lane(group);
    const bool is_leader = (int)me == leader;

    uint32_t index = 0u, entry = 0u;
    // Under load the hint (cu_hint() for malloc) is mixed with the wave's
    // index in its block: the waves of one block run on one CU, and with the
    // CU id alone all of them would start every class's claim on the same
    // slot. The mix adds multiples of 8, keeping the hint's XCD bits. A
    // single lane keeps the hint as given.
    const bool multi = (looking & (looking - 1)) != 0;
    const uint32_t wave_in_block = (uint32_t)(threadIdx.x >> 6);
    const uint32_t lead_hint = spread && multi && wave_in_block != 0
                                   ? wave::shfl(hint, leader) + wave_in_block * (0x9E3779B1u << 3)
                                   : wave::shfl(hint, leader);
    const uint32_t usable = usable_slots[tree];
    // Under load a group's first pass starts at slot hint % usable; an empty
    // slot there is filled unless the class is out of memory. Otherwise (and
    // always for a single lane) the reference's walk over the pinned slots.
    bool placed = false;
    bool dry = false;    // the leader found the class out of memory in this pass
    bool spare = false;  // the leader saw free segments in this pass (a refill can get a block)
    bool look_refill = false;  // the group refills the empty slot its first look landed on
    //
    // A class whose usable slots are its pinned slots (the two smallest
    // classes) takes the same first look at slot hint % pinned, where the
    // walk would start: a block there is the walk's own result, and when the
    // class is out of memory with every pinned slot holding 0 the group fails
    // at once (the walk would end on an empty slot whose refill fails).
    const uint32_t pinned = pinned_slots_of(tree);
    if (spread && attempt == 0 && (looking & (looking - 1)) != 0) {
      const bool spreads = usable > pinned;
      const uint32_t si = lead_hint % (spreads ? usable : pinned);
      // read with the slot, in the same round trip: the class's tree summary,
      // the free-segment gauge and the count of empty pinned slots
      const uint64_t class_top = sub_trees[tree].top_word();
      const uint32_t free_segments = atom::load(ctl_word(control, kCtlFreeSegs));
      const int32_t empty_pinned = atom::load(empty_pinned_word(tree));
      const uint32_t se = atom::load(slot_ptr(tree, si));
      // The class's tree empty (its summaries are exact: it changes only by
      // insert_force_update/remove) and no free segment left is out of
      // memory without the search: the search would find nothing either.
      const bool tree_dry = sub_trees[tree].top_looks_empty(class_top) && free_segments == 0u;
      // Out of memory, and every pinned slot holds 0 as well: a walk over the
      // pinned slots would find no block.
      const bool all_empty = empty_pinned >= (int32_t)pinned;
      if (se != 0u && se != kSlotBusy) {
        index = si;
        entry = se;
        placed = true;
      } else if (se == 0u && spreads) {
        // free segments left: a refill attaches one if the class has no
        // block to grant, so the class is not out of memory (no search)
        spare = free_segments != 0u;
        dry = tree_dry || (!spare && class_out_of_memory(tree, tid));
        if (!dry || all_empty) {
          index = si;
          entry = 0u;
          placed = true;
          look_refill = !dry;
        }
      } else if (se == 0u && tree_dry && all_empty) {
        dry = true;
        index = si;
        entry = 0u;
        placed = true;
      }
    }
    placed = wave::shfl((uint32_t)placed, leader) != 0u;
    dry = wave::shfl((uint32_t)dry, leader) != 0u;
    spare = wave::shfl((uint32_t)spare, leader) != 0u;
    look_refill = wave::shfl((uint32_t)look_refill, leader) != 0u;
    if (placed) {
      index = wave::shfl(index, leader);
      entry = wave::shfl(entry, leader);
    }
    if (wave::ballot(!placed) & looking) {
      uint32_t wi = 0u, we = 0u;
      pinned_slot_group(tree, lead_hint, group, looking & ~wave::ballot(placed), wi, we);
      if (!placed) {
        index = wi;
        entry = we;
      }
    }
    // Every pinned slot of the class busy: the group's lanes each look at
    // one overflow slot, from a start that moves with the hint and the pass,
    // and take the first holding a block, else the first empty one (which
    // the group then fills), else stay busy.
    if (wave::ballot(entry == kSlotBusy && usable > pinned_slots_of(tree)) & looking) {
      const uint32_t k = usable - pinned_slots_of(tree);
      const bool want = entry == kSlotBusy && k > 0;
      const uint32_t rank = wave::popc(group & wave::lanemask_lt());
      const bool look = want && rank < k;
      uint32_t my_index = 0, my_entry = kSlotBusy;
      if (look) {
        const uint32_t start = (lead_hint * 0x9E3779B1u + attempt * 0x85EBCA6Bu) % k;
        my_index = pinned_slots_of(tree) + (start + rank) % k;
        my_entry = atom::load(slot_ptr(tree, my_index));
      }
      const uint64_t with_block = wave::ballot(look && my_entry != 0u && my_entry != kSlotBusy) & group;
      const uint64_t empty_one = wave::ballot(look && my_entry == 0u) & group;
      const uint64_t pick = with_block ? with_block : empty_one;
      const int src = (want && pick) ? wave::first_lane(pick) : (int)me;
      const uint32_t picked_index = wave::shfl(my_index, src);
      const uint32_t picked_entry = wave::shfl(my_entry, src);
      if (want && pick) {
        index = picked_index;
        entry = picked_entry;
      }
    }
    const bool busy = entry == kSlotBusy;  // being replaced: wait for it
    const bool has_block = entry != 0u && !busy;
    const uint64_t block = has_block ? entry - 1u : 0;

    // exclusive scan of the slice counts inside each group, in lane order
    uint32_t prefix, total;
    if ((wave::ballot(count != 1u) & looking) == 0) {
      prefix = wave::popc(group & wave::lanemask_lt());
      total = wave::popc(group);
    } else {
      wave::group_scan(group, count, 13, prefix, total);
    }
    // each group's last lane holds the inclusive total and does the one add
    // (reference block.cuh:245-264)
    uint64_t merged = 0;
    if ((int)me == last && has_block) merged = atom::add(&table.blocks[block].malloc_counter, (uint64_t)total);
    merged = wave::shfl64(merged, last);

    const uint32_t c0 = counter_count(merged);
    const uint64_t allocation = (uint64_t)c0 + prefix;
    const bool should_replace = has_block && allocation <= 4095 && allocation + count > 4095;
    const bool should_fetch = spread && ahead && multi && has_block && allocation <= kAheadAt && allocation + count > kAheadAt;
    const bool valid = has_block && allocation + count <= kSlicesPerBlock;
    const uint64_t valid_lanes = wave::ballot(valid) & group;
    // Slices of the claimed range that lie inside the block but belong to no
    // successful lane (the extra slices of multi-slice requests and a
    // straddling request's tail) are written off as already freed
    // (reference block.cuh:267-281): in-range slices minus successful lanes.
    const uint32_t in_range =
        c0 >= kSlicesPerBlock ? 0u : ((c0 + total < kSlicesPerBlock ? c0 + total : kSlicesPerBlock) - c0);
    // allocations the successful lanes hold: one per lane, or a lane's run
    // length for the batched bulk claims
    uint32_t held;
    if ((wave::ballot(allocs != 1u) & looking) == 0) {
      held = wave::popc(valid_lanes);
    } else {
      uint32_t unused;
      wave::group_scan(group, valid ? allocs : 0u, 7, unused, held);
    }
    const uint32_t written_off = has_block ? in_range - held : 0u;
    bool completed = false;
    if (is_leader && written_off > 0) {
      const uint32_t old = atom::add(&table.blocks[block].free_counter, written_off);
      completed = (old + written_off == kSlicesPerBlock);
    }
    // a stale pinned read: the block now serves another class; what the valid
    // lanes drew is given back (reference gallatin.cuh:1037-1056)
    const bool stale = has_block && counter_class(merged) != tree;
    const uint32_t give_back = stale ? held : 0u;
    // The group whose range crosses slice 4095 replaces the block (reference
    // :992, :1015-1026). A group that finds the block already spent, or
    // serving another class, replaces it too: a claim made through a stale
    // slot read can exhaust a block that was recycled into another slot, and
    // its replace attempt then targets the old slot, leaving the new one
    // pinned to a spent block. With one lane this never happens (the crossing
    // lane always swaps its own slot), so single-lane results are unchanged.
    const bool replace = ((wave::ballot(should_replace) & group) != 0) || (has_block && c0 > 4095) || stale;

    bool refilled = true;
    const bool empty = entry == 0u;
    // the group whose range crosses slice kAheadAt fetches the slot's next block
    const bool fetch = ((wave::ballot(should_fetch) & group) != 0) && !replace && !stale;
    const uint64_t rare =
        wave::ballot(is_leader && (empty || replace || completed || give_back > 0 || fetch)) & looking;
    for (uint64_t todo = rare; todo; todo &= todo - 1) {
      if ((int)me == wave::first_lane(todo)) {
        const uint64_t lead_tid = tid;
        if (empty) {
          // `dry` / `spare`: this pass's first look already answered the test
          refilled = !dry && (spare || !class_out_of_memory(tree, lead_tid)) && refill_slot(tree, index, lead_tid);
        } else {
          if (replace || fetch) replace_block(tree | (replace ? 0u : kFetchOnly), index, block, lead_tid);
          if (completed) recycle_block(block);  // the write-off finished the block
          if (give_back > 0) free_slices(block, give_back);
        }
      }
    }
    if (wave::ballot(busy) & looking) wave::backoff(8);
    waited = busy;
    // A refill that finds no block is the class out of memory, unless the
    // slot was the first look's: blocks there can still be in the pinned
    // slots, which the next pass walks (it takes no first look).
    if (empty && wave::shfl((uint32_t)refilled, leader) == 0u && !look_refill) {  // class out of memory
      done = true;
      oom = true;
      result = utils::kFail;
    }
    if (valid && !stale) {
      result = allocation + block * kSlicesPerBlock;
      done = true;
    }
  }

  // Slice claims for every lane of the wave that calls it. The loop is
  // uniform (driven by a lane mask, never by a lane leaving it): the AMDGPU
  // structurizer runs a loop's exit block once, after the loop, for all lanes
  // that left, so per-group work must not sit in one. `oom` is set when the
  // failure is a confirmed out-of-memory, which a further call would repeat.
  template <bool spread = true, bool ahead = spread>
  __device__ __forceinline__ uint64_t malloc_slice_allocation(uint32_t tree, uint32_t count, uint32_t allocs,
                                                              uint64_t tid, uint32_t hint, uint32_t subgroup,
                                                              bool& oom) {
    uint64_t result = utils::kFail;
    bool done = false;
    oom = false;
    uint64_t looking = wave::active_mask();
    // A pass spent waiting for a busy slot is not a failed attempt: the
    // reference's attempt budget counts claims that failed, and a single
    // lane never waits. Waiting is bounded separately.
    uint32_t tries = 0;
    for (int pass = 0; pass < kBusyPasses && looking; pass++) {
      bool waited = false;
      if ((looking >> wave::lane_id()) & 1ull)
        claim_pass<spread, ahead>(looking, tree, count, allocs, tid, hint, subgroup, (uint32_t)pass, result, done, oom,
                           waited);
      if (!done && !waited) tries++;
      looking &= ~wave::ballot(done || tries >= (uint32_t)kSliceAttempts);
    }
    return result;
  }

  // whole-block request (reference :828-861 + Block::malloc_fill_block)
  __device__ __noinline__ uint64_t malloc_block_allocation(uint32_t tree, uint64_t tid) {
    bool last = false;
    int64_t block = request_new_block_from_tree(tree, tid, last);
    if (block < 0) return utils::kFail;
    Block* b = &table.blocks[block];
    const uint64_t old_merged = atom::add(&b->malloc_counter, (uint64_t)kSlicesPerBlock);
    const uint32_t old_count = counter_count(old_merged);
    if (counter_class(old_merged) != tree) count_error(kErrFillClass);
    atom::add(&b->free_counter, 4095u - old_count);
    if (last) gather_after_last(tree);
    if (old_count != 0) {
      count_error(kErrFillCount);
      free_slices((uint64_t)block, 1u);
      return utils::kFail;
    }
    return (uint64_t)block * kSlicesPerBlock;
  }

  // contiguous segments (reference :790-826). The calling lanes (`lanes`,
  // all active, the call uniform) find the word the claim starts at
  // together (veb_tree::gather_plan), which reads the tree one word per lane
  // per round trip; one lane then claims the run from that word with the
  // gather rule (veb_tree::gather_multiple). When the tree as read holds no
  // run of n free segments the request fails at once (`oom`, a further
  // attempt would repeat it).
  //
  // The reference serialises these claims on a spin lock (bit T of its lock
  // word). Here they run concurrently: every bit of a run is taken with an
  // atomic and a claim that finds part of its run gone puts back exactly
  // what it took and goes on searching (gather_bottom), so two claims never
  // share a segment; the lock added nothing to that, and with it every
  // claim of the GPU waited for the one before it (the block and segment
  // churn ran at 7e5 mallocs+frees/s). What concurrency changes is which run
  // a claim gets: the highest run when its search ran, not at one
  // serialised instant. With one lane nothing runs between the search and
  // the claim, and the result is the reference's.
  __device__ __noinline__ uint64_t malloc_segment_allocation(uint64_t n, uint64_t lanes, bool& oom) {
    const uint64_t start = segment_tree.gather_plan(n, lanes);
    oom = start == utils::kFail;
    if (oom) return utils::kFail;
    const int leader = wave::first_lane(lanes);
    uint64_t first = utils::kFail;
    if ((int)wave::lane_id() == leader) {
      first = segment_tree.gather_multiple(n, start);
      if (first != utils::kFail) {
        atom::sub(ctl_word(control, kCtlFreeSegs), (uint32_t)n);
        if (!table.set_tree_id(first, (uint32_t)(kNumTrees + 1 + n))) count_error(kErrSetup);
      }
    }
    first = wave::readlane64(first, leader);
    atom::order();
    return first == utils::kFail ? utils::kFail : first * kBlocksPerSegment * kSlicesPerBlock;
  }

  // memory_table.cuh:1072-1086
  __device__ void* offset_to_allocation(uint64_t offset, uint32_t tree) const {
    const uint64_t per = table_type::max_allocations_per_segment;
    return table.memory + (offset / per) * bytes_per_segment + (offset % per) * table_type::tree_alloc_size(tree);
  }

  // Block and segment requests. A segment request runs with every calling
  // lane helping to search the segment tree, one request of the calling
  // lanes at a time. A block request runs in its own lane: one at a time in
  // lane order as well when `spread` is false (the order the deterministic
  // multi-lane replays model), all of the wave's block requests together
  // under load (their lanes run the block search side by side: with one at
  // a time a wave of block requests took as long as all of them in a row).
  // Kept out of line: the slice path inlined into a kernel then saves no
  // registers for it.
  template <bool spread>
  __device__ __noinline__ void* malloc_large(uint64_t size, uint64_t tid, int block_tree) {
    const bool is_block = (uint32_t)block_tree < kNumTrees;
    const uint64_t n = is_block ? 0 : (size - 1) / bytes_per_segment + 1;
    const uint64_t callers = wave::active_mask();
    void* out = nullptr;
    uint64_t pending = callers;
    if (spread) {
      if (is_block) {
        uint64_t offset = utils::kFail;
        for (int a = 0; a < kBlockAttempts && offset == utils::kFail; a++)
          offset = malloc_block_allocation((uint32_t)block_tree, tid);
        if (offset != utils::kFail) out = offset_to_allocation(offset, (uint32_t)block_tree);
      }
      pending &= ~wave::ballot(is_block);
    }
    for (; pending; pending &= pending - 1) {
      const int who = wave::first_lane(pending);
      if (wave::readlane((uint32_t)is_block, who)) {
        if ((int)wave::lane_id() == who) {
          uint64_t offset = utils::kFail;
          for (int a = 0; a < kBlockAttempts && offset == utils::kFail; a++)
            offset = malloc_block_allocation((uint32_t)block_tree, tid);
          if (offset != utils::kFail) out = offset_to_allocation(offset, (uint32_t)block_tree);
        }
      } else {
        const uint64_t runs = wave::readlane64(n, who);
        uint64_t offset = utils::kFail;
        bool oom = false;
        if (runs + kNumTrees + 1 < 0xFFFFull)
          for (int a = 0; a < kSegmentAttempts && offset == utils::kFail && !oom; a++)
            offset = malloc_segment_allocation(runs, callers, oom);
        if ((int)wave::lane_id() == who && offset != utils::kFail) out = offset_to_allocation(offset, 0);
      }
    }
    return out;
  }

  // --------------------------------------------------------------- malloc
  // malloc with the slot selector and hash id given explicitly; malloc()
  // passes the hardware CU id and the global thread id. `subgroup` (0 for
  // malloc) splits the lanes of one class into separate claim groups: the
  // coalescing width of the batched kernels. `spread` and `ahead` as in
  // claim_pass.
  template <bool spread = true, bool ahead = spread>
  __device__ __forceinline__ void* malloc_at(uint64_t size, uint64_t tid, uint32_t hint, uint32_t subgroup = 0) {
    uint32_t tree = get_tree_id_from_size(size);
    uint32_t count = 1;
    if (tree >= kNumTrees) {
      const int block_tree = utils::get_first_bit_bigger(size) - kSmallestTreeBits;
      if (block_tree >= 0) return malloc_large<spread>(size, tid, block_tree);
      count = 1u << (tree - (kNumTrees - 1));
      tree = kNumTrees - 1;
    }
    uint64_t offset = utils::kFail;
    bool oom = false;
    for (int attempt = 0; attempt < kMallocLoopAttempts && offset == utils::kFail && !oom; attempt++)
      offset = malloc_slice_allocation<spread, ahead>(tree, count, 1u, tid, hint, subgroup, oom);
    if (offset == utils::kFail) return nullptr;
    return offset_to_allocation(offset, tree);
  }

  // The path malloc takes for a request: 0 slices (one, or 2^k top-class
  // slices), 1 a whole block, 2 a run of segments (reference :1186-1276).
  __host__ __device__ static uint32_t request_path(uint64_t size) {
    const int bits = utils::get_first_bit_bigger(size < smallest ? smallest : size);
    if (bits - kSmallestBits < (int)kNumTrees) return 0u;
    const int block_tree = bits - kSmallestTreeBits;
    if (block_tree < 0) return 0u;
    return block_tree < (int)kNumTrees ? 1u : 2u;
  }

  // Bytes between consecutive allocations of a run of `size`-byte requests
  // (0: the size takes the block or segment path, which has no runs).
  __host__ __device__ static uint64_t run_stride(uint64_t size) {
    const uint32_t tree = size < smallest ? 0u : (uint32_t)(utils::get_first_bit_bigger(size) - kSmallestBits);
    if (tree < kNumTrees) return smallest << tree;
    if (utils::get_first_bit_bigger(size) - kSmallestTreeBits >= 0) return 0;
    return (uint64_t)biggest << (tree - (kNumTrees - 1));
  }

  // Bulk claim for the batched kernels (no counterpart in the reference,
  // whose callers malloc one request per lane): k requests of one size class
  // served from consecutive slices of one block with the lane's share of one
  // group add, exactly as a k-times larger multi-slice request is claimed
  // (reference gallatin.cuh:865-1078), but accounted as k allocations, each
  // freed on its own (or all at once with free_run). Returns the first
  // allocation; allocation r is at + r * run_stride(size). nullptr if the
  // size has no runs or k * slices-per-request exceeds a block.
  __device__ __forceinline__ void* malloc_run_at(uint64_t size, uint32_t k, uint64_t tid, uint32_t hint,
                                                 uint32_t subgroup) {
    uint32_t tree = get_tree_id_from_size(size);
    uint32_t per = 1;
    bool ok = k >= 1 && k <= 64;
    if (tree >= kNumTrees) {
      if (utils::get_first_bit_bigger(size) - kSmallestTreeBits >= 0) ok = false;
      per = ok ? 1u << (tree - (kNumTrees - 1)) : 1u;
      tree = kNumTrees - 1;
    }
    if (per * k > kSlicesPerBlock) ok = false;
    uint64_t offset = utils::kFail;
    if (ok) {
      bool oom = false;
      for (int attempt = 0; attempt < kMallocLoopAttempts && offset == utils::kFail && !oom; attempt++)
        offset = malloc_slice_allocation(tree, per * k, k, tid, hint, subgroup, oom);
    }
    if (offset == utils::kFail) return nullptr;
    return offset_to_allocation(offset, tree);
  }

  // The slot hint of malloc (the reference's SM id, gallatin.cuh:898): the
  // calling wave's CU with its XCD in the low three bits. On MI355X __smid()
  // holds the XCD above the shader-engine and CU ids, so hint % slots alone
  // would send the CUs of all eight XCDs to the same slots; with the XCD
  // low, the waves of one XCD start on slots of their own (slot = XCD mod 8
  // for the classes' 32, 64 and 128 usable slots), and a pinned block's
  // counter and payload lines are worked from one XCD (mixed workload
  // 5.7e9 -> 6.1e9 mallocs+frees/s).
  __device__ static uint32_t cu_hint() {
    const uint32_t sm = __smid();
    return (sm >> 6) | ((sm & 63u) << 3);
  }

  // reference gallatin.cuh:1186
  __device__ __forceinline__ void* malloc(uint64_t size) {
    const uint64_t tid = (uint64_t)(uint32_t)(threadIdx.x + blockIdx.x * blockDim.x);
    return malloc_at(size, tid, cu_hint());
  }

  // ----------------------------------------------------------------- free
  // Frees of the lanes of a wave, grouped by block: one release fence for the
  // wave (its stores are performed before any of its slices can be handed
  // out again), then one atomic instruction in which each group's first lane
  // adds the group's size to its block's free counter. A block whose count
  // completes is recycled, one at a time. `release` false skips the fence,
  // for callers whose stores to the freed memory are already released.
  __device__ __forceinline__ void free_offset(uint64_t offset, uint32_t allocs = 1u, bool release = true) {
    const uint64_t block = offset / kSlicesPerBlock;
    const uint64_t active = wave::active_mask();
    const uint64_t group = wave::match_groups(active, block);
    const bool is_leader = (int)wave::lane_id() == wave::first_lane(group);
    uint32_t n;
    if ((wave::ballot(allocs != 1u) & active) == 0) {
      n = wave::popc(group);
    } else {
      uint32_t unused;
      wave::group_scan(group, allocs, 7, unused, n);
    }
    if (release) atom::fence_release();
    bool complete = false;
    if (is_leader) {
      const uint32_t old = atom::add(&table.blocks[block].free_counter, n);
      complete = (old + n == kSlicesPerBlock);
    }
    for (uint64_t todo = wave::ballot(complete); todo; todo &= todo - 1)
      if ((int)wave::lane_id() == wave::first_lane(todo)) recycle_block(block);
  }

  // A run of segments (reference gallatin.cuh:1283-1302): release the id
  // first so that no claimer can find the segment in the tree while the old
  // id is still set. Only the lane whose reset succeeds returns the segments
  // and adds them to the free-segment gauge: a second free of the same run
  // (two lanes at once, or a double free) would otherwise count it twice.
  __device__ __noinline__ void free_segment_run(uint64_t segment, uint32_t tree) {
    if (!table.reset_tree_id(segment, tree)) {
      count_error(kErrRunReset);
      return;
    }
    atom::fence_release();
    segment_tree.return_multiple(segment, tree - kNumTrees - 1);
    atom::add(ctl_word(control, kCtlFreeSegs), tree - kNumTrees - 1);
  }

  // Free all k allocations of a run from malloc_run_at at once: one share
  // of one group add, as k separate frees would add one each.
  __device__ __forceinline__ void free_run(void* first, uint32_t k) {
    const uint64_t byte = (uint64_t)((char*)first - table.memory);
    const uint64_t segment = byte / bytes_per_segment;
    const uint32_t tree = byte < table.num_segments * bytes_per_segment ? table.read_tree_id(segment) : kFreeChunk;
    if (tree >= kNumTrees || k == 0 || k > 64) {
      atom::add(ctl_word(control, kCtlBadFrees), 1u);
      return;
    }
    const uint64_t offset = ((byte - segment * bytes_per_segment) >> (kSmallestBits + tree)) +
                            segment * table_type::max_allocations_per_segment;
    free_offset(offset, k);
  }

  // reference gallatin.cuh:1378-1461. The 8 XCD L2 caches of MI355X are not
  // coherent with each other: a free first writes back the calling wave's
  // stores (an agent-scope release, one L2 writeback per call), so that a
  // dirty line of the freed memory cannot be written back over the data of
  // its next owner on another XCD.
  __device__ __forceinline__ void free(void* allocation) { free_impl(allocation, true); }

  // free() for memory whose stores are already released: nothing the calling
  // wave stored to it since its last agent-scope release fence (or since its
  // kernel began: a kernel boundary releases everything) is still pending.
  // Typical caller: a kernel that frees what an earlier kernel allocated and
  // wrote, reading it at most. Skips the L2 writeback of free(); has no
  // counterpart in the reference, whose L2 is coherent.
  __device__ __forceinline__ void free_released(void* allocation) { free_impl(allocation, false); }

  __device__ __forceinline__ void free_impl(void* allocation, bool release) {
    const uint64_t byte = (uint64_t)((char*)allocation - table.memory);
    const uint64_t segment = byte / bytes_per_segment;
    if (byte >= table.num_segments * bytes_per_segment) {
      atom::add(ctl_word(control, kCtlBadFrees), 1u);
      return;
    }
    const uint32_t tree = table.read_tree_id(segment);
    if (tree == kFreeChunk || tree == kNumTrees || tree == kNumTrees + 1) {
      atom::add(ctl_word(control, kCtlBadFrees), 1u);
      return;
    }
    if (tree > kNumTrees) {
      free_segment_run(segment, tree);
      return;
    }
    const uint64_t offset = ((byte - segment * bytes_per_segment) >> (kSmallestBits + tree)) +
                            segment * table_type::max_allocations_per_segment;
    free_offset(offset, 1u, release);
  }

  __device__ bool owns_allocation(void* allocation) const {
    return (uint64_t)((char*)allocation - table.memory) < table.num_segments * bytes_per_segment;
  }

  // ------------------------------------------------------------ boot
  // fill every pinned slot in class-major order with hash id 0
  // (reference boot_shared_block_container_one_thread, gallatin.cuh:196-221)
  __device__ bool boot_pinned_blocks() {
    for (uint32_t tree = 0; tree < kNumTrees; tree++) {
      for (uint32_t s = 0; s < pinned_slots_of(tree); s++) {
        bool last = false;
        int64_t block = request_new_block_from_tree(tree, 0, last);
        if (block < 0) return false;
        if (table.read_tree_id((uint64_t)block / kBlocksPerSegment) != tree) return false;
        if (atom::cas(slot_ptr(tree, s), 0u, (uint32_t)(block + 1)) != 0u) return false;
        if (last) gather_after_last(tree);
      }
    }
    return true;
  }
#endif  // __HIPCC__

  // ------------------------------------------------------------ host side
  __host__ static uint64_t layer_words(uint64_t universe, int layers, uint64_t* words, uint64_t* universes) {
    uint64_t u = universe, total = 0;
    for (int i = layers - 1; i >= 0; i--) {
      universes[i] = u;
      words[i] = (u - 1) / 64 + 1;
      total += words[i];
      u = (u - 1) / 64 + 1;
    }
    return total;
  }

  // Bytes a request of `size` is granted (the path malloc takes for it):
  // a slice of its class, 2^k top-class slices, a whole block, or whole
  // segments (reference gallatin.cuh:1152-1183, which reports the next
  // power of two for every request below a segment).
  __host__ __device__ static uint64_t allocated_size(uint64_t size) {
    if (size < smallest) size = smallest;
    const int bits = utils::get_first_bit_bigger(size);
    const int tree = bits - kSmallestBits;
    if (tree < (int)kNumTrees) return smallest << tree;
    if (bits - kSmallestTreeBits < 0) return (uint64_t)biggest << (tree - (int)(kNumTrees - 1));
    if (bits - kSmallestTreeBits < (int)kNumTrees) return 1ull << bits;
    return ((size - 1) / bytes_per_segment + 1) * bytes_per_segment;
  }

  // the reference's names for the same figures (gallatin.cuh:1152, 1901-1910)
  __host__ __device__ static uint64_t get_allocated_size(uint64_t size) { return allocated_size(size); }
  __host__ __device__ static uint64_t get_blocks_per_segment(uint16_t tree) {
    return table_type::blocks_in_segment(tree);
  }
  __host__ __device__ static uint64_t get_max_allocations_per_segment() {
    return table_type::max_allocations_per_segment;
  }

  // reference gallatin.cuh:324-421. Returns a device pointer, nullptr on failure.
  __host__ static my_type* generate_on_device(uint64_t max_bytes, uint64_t seed, bool print_info = true,
                                              bool running_calloc = false) {
    return generate(max_bytes, seed, print_info, device_only);
  }
  // the same over a pool of mapped host memory (reference gallatin.cuh:427-630)
  __host__ static my_type* generate_on_device_host(uint64_t max_bytes, uint64_t seed, bool print_info = true,
                                                   bool running_calloc = false) {
    return generate(max_bytes, seed, print_info, host_only);
  }
  __host__ static my_type* generate(uint64_t max_bytes, uint64_t seed, bool print_info, Gallatin_memory_type kind);
  __host__ static void free_on_device(my_type* dev_version);
  // called on the device pointer, as in the reference (:1855)
  __host__ void print_info();
  __host__ static bool host_copy(const my_type* dev, my_type* out) {
    return hipMemcpy(out, dev, sizeof(my_type), hipMemcpyDeviceToHost) == hipSuccess;
  }
};

#if defined(__HIPCC__)
// ---- construction kernels
static __global__ void gallatin_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
static __global__ void gallatin_init_blocks(Block* b, uint64_t n) {
  for (uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    b[i] = Block{4097ull, 0u, 0u};  // Block::init, reference block.cuh:64-68
}
static __global__ void gallatin_init_segments(segment_counters* s, uint64_t n) {
  for (uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    s[i] = segment_counters{-1, 0u, 0u, 0u};
}
// item layer and summaries all set for the first `universe` items of each layer
static __global__ void gallatin_init_full_layer(uint64_t* w, uint64_t words, uint64_t universe) {
  for (uint64_t i = threadIdx.x + (uint64_t)blockIdx.x * blockDim.x; i < words; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo = i * 64;
    uint64_t n = universe > lo ? universe - lo : 0;
    w[i] = utils::bitmask(n > 64 ? 64 : n);
  }
}
template <class A>
__global__ void gallatin_boot_kernel(A* alloc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    bool ok = alloc->boot_pinned_blocks();
    atom::store(ctl_word(alloc->control, kCtlBootStatus), ok ? 1u : 2u);
  }
}
#endif

namespace detail {
__host__ inline unsigned grid_for(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  return (unsigned)(g == 0 ? 1 : (g > 65535 ? 65535 : g));
}
}  // namespace detail

template <uint64_t S, uint64_t M, uint64_t B>
__host__ Gallatin<S, M, B>* Gallatin<S, M, B>::generate(uint64_t max_bytes, uint64_t seed, bool print_info,
                                                       Gallatin_memory_type kind) {
  // calloc (running_calloc) needs device-side launches; not supported
  if (kind != device_only && kind != host_only) return nullptr;
  const uint64_t nseg = max_bytes / S;
  if (nseg == 0) return nullptr;
  const int layers = veb_tree::layers_for(nseg);
  if (layers > kMaxVebLayers) return nullptr;

  my_type host;
  std::memset(&host, 0, sizeof(host));
  host.seed = seed;

  uint64_t words[kMaxVebLayers] = {0}, universes[kMaxVebLayers] = {0};
  const uint64_t tree_words = layer_words(nseg, layers, words, universes);
  const uint64_t nblocks = nseg * kBlocksPerSegment;

  void* veb = nullptr;
  void* table_mem = nullptr;
  void* pool = nullptr;       // device address of the pool
  void* pool_host = nullptr;  // host address of a host_only pool
  void* dev = nullptr;
  const size_t veb_bytes = (size_t)tree_words * 8 * (1 + kNumTrees);
  // chunk ids | blocks | queues | segment counters | slots | control
  const size_t off_blocks = ((nseg * 4 + 255) / 256) * 256;
  const size_t off_queues = off_blocks + nblocks * sizeof(Block);
  const size_t off_segs = off_queues + ((nblocks * 4 + 255) / 256) * 256;
  const size_t off_slots = off_segs + nseg * sizeof(segment_counters);
  const size_t off_ctl = off_slots + ((total_slots() * kSlotStride * 4 + 255) / 256) * 256;
  const size_t off_class = off_ctl + kCtlWords * kCtlStride * 4;
  const size_t table_bytes = off_class + kNumTrees * kCtlStride * 4;

  auto fail = [&]() -> my_type* {
    if (veb) (void)hipFree(veb);
    if (table_mem) (void)hipFree(table_mem);
    if (pool_host) (void)hipHostFree(pool_host);
    else if (pool) (void)hipFree(pool);
    if (dev) (void)hipFree(dev);
    return nullptr;
  };
  if (hipMalloc(&veb, veb_bytes) != hipSuccess) return fail();
  if (hipMalloc(&table_mem, table_bytes) != hipSuccess) return fail();
  if (kind == host_only) {
    if (hipHostMalloc(&pool_host, nseg * S, hipHostMallocMapped) != hipSuccess) return fail();
    std::memset(pool_host, 0, nseg * S);
    if (hipHostGetDevicePointer(&pool, pool_host, 0) != hipSuccess) return fail();
  } else if (hipMalloc(&pool, nseg * S) != hipSuccess) {
    return fail();
  }
  if (hipMalloc(&dev, sizeof(my_type)) != hipSuccess) return fail();

  // vEB trees: segment tree full, class trees empty
  uint64_t* vw = (uint64_t*)veb;
  if (hipMemset(veb, 0, veb_bytes) != hipSuccess) return fail();
  auto place = [&](veb_tree& t, uint64_t* base, uint64_t tree_seed) {
    t.num_layers = layers;
    t.seed = tree_seed;
    t.total_universe = nseg;
    uint64_t off = 0;
    for (int i = 0; i < layers; i++) {
      t.bits[i] = base + off;
      t.words[i] = words[i];
      t.universe[i] = universes[i];
      off += words[i];
    }
  };
  place(host.segment_tree, vw, seed);
  for (int i = 0; i < layers; i++)
    gallatin_init_full_layer<<<detail::grid_for(words[i]), 256>>>(host.segment_tree.bits[i], words[i], universes[i]);
  for (uint32_t c = 0; c < kNumTrees; c++) place(host.sub_trees[c], vw + tree_words * (1 + c), seed + c);

  char* tm = (char*)table_mem;
  host.table.chunk_ids = (uint32_t*)tm;
  host.table.blocks = (Block*)(tm + off_blocks);
  host.table.queues = (uint32_t*)(tm + off_queues);
  host.table.segs = (segment_counters*)(tm + off_segs);
  host.table.memory = (char*)pool;
  host.table.num_segments = nseg;
  host.slots = (uint32_t*)(tm + off_slots);
  host.control = (uint32_t*)(tm + off_ctl);
  host.class_words = (int32_t*)(tm + off_class);
  uint32_t off = 0;
  for (uint32_t c = 0; c < kNumTrees; c++) {
    host.slot_offset[c] = off;
    off += slot_region_of(c);
    host.usable_slots[c] = usable_slots_of(c, nseg * S, table_type::tree_alloc_size(c) * kSlicesPerBlock);
  }
  if (hipMemset(table_mem, 0, table_bytes) != hipSuccess) return fail();
  if (nseg > 0xFFFFFFFFull) return fail();
  const uint32_t free_segs = (uint32_t)nseg;
  if (hipMemcpy(ctl_word(host.control, kCtlFreeSegs), &free_segs, 4, hipMemcpyHostToDevice) != hipSuccess) return fail();
  gallatin_fill_u32<<<detail::grid_for(nseg), 256>>>(host.table.chunk_ids, nseg, kFreeChunk);
  gallatin_init_blocks<<<detail::grid_for(nblocks), 256>>>(host.table.blocks, nblocks);
  gallatin_init_segments<<<detail::grid_for(nseg), 256>>>(host.table.segs, nseg);
  if (kind == device_only && hipMemset(pool, 0, nseg * S) != hipSuccess) return fail();

  host.owned[0] = veb;
  host.owned[1] = table_mem;
  host.owned[2] = kind == host_only ? pool_host : pool;
  host.owned[3] = dev;
  host.memory_kind = (uint32_t)kind;
  if (hipMemcpy(dev, &host, sizeof(my_type), hipMemcpyHostToDevice) != hipSuccess) return fail();
  gallatin_boot_kernel<my_type><<<1, 1>>>((my_type*)dev);
  if (hipDeviceSynchronize() != hipSuccess) return fail();
  uint32_t status = 0;
  if (hipMemcpy(&status, ctl_word(host.control, kCtlBootStatus), 4, hipMemcpyDeviceToHost) != hipSuccess || status != 1u)
    return fail();
  if (print_info)
    printf("Booted Gallatin with %u trees in range %llu-%llu and %f GB of memory %llu segments\n", kNumTrees,
           (unsigned long long)M, (unsigned long long)B, 1.0 * (double)max_bytes / 1024 / 1024 / 1024,
           (unsigned long long)nseg);
  return (my_type*)dev;
}

template <uint64_t S, uint64_t M, uint64_t B>
__host__ void Gallatin<S, M, B>::free_on_device(my_type* dev_version) {
  if (!dev_version) return;
  my_type host;
  if (!host_copy(dev_version, &host)) return;
  (void)hipDeviceSynchronize();
  for (int i = 0; i < 2; i++)
    if (host.owned[i]) (void)hipFree(host.owned[i]);
  if (host.owned[2]) {
    if (host.memory_kind == host_only)
      (void)hipHostFree(host.owned[2]);
    else
      (void)hipFree(host.owned[2]);
  }
  (void)hipFree(dev_version);
}

template <uint64_t S, uint64_t M, uint64_t B>
__host__ void Gallatin<S, M, B>::print_info() {
  my_type host;
  if (!host_copy(this, &host)) return;
  (void)hipDeviceSynchronize();
  auto fill = [&](const veb_tree& t) {
    int b = t.num_layers - 1;
    std::vector<uint64_t> w(t.words[b]);
    if (hipMemcpy(w.data(), t.bits[b], w.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) return (uint64_t)0;
    uint64_t n = 0;
    for (uint64_t x : w) n += (uint64_t)__builtin_popcountll(x);
    return n;
  };
  const uint64_t nseg = host.table.num_segments;
  printf("\n\033[1;32mGallatin usage stats:\033[1;0m\n");
  printf("Segment tree sees %llu/%llu segments available\n", (unsigned long long)fill(host.segment_tree),
         (unsigned long long)nseg);
  for (uint32_t c = 0; c < kNumTrees; c++)
    printf("Tree %u: size %llu, owns %llu/%llu\n", c, (unsigned long long)table_type::tree_alloc_size(c),
           (unsigned long long)fill(host.sub_trees[c]), (unsigned long long)nseg);
  // per-class live blocks from the segment counters (reference :1976-2011)
  std::vector<uint32_t> ids(nseg);
  std::vector<segment_counters> segs(nseg);
  if (hipMemcpy(ids.data(), host.table.chunk_ids, nseg * 4, hipMemcpyDeviceToHost) != hipSuccess) return;
  if (hipMemcpy(segs.data(), host.table.segs, nseg * sizeof(segment_counters), hipMemcpyDeviceToHost) != hipSuccess)
    return;
  printf("\n\033[1;32mTree usage stats:\033[1;0m\n");
  for (uint32_t c = 0; c < kNumTrees; c++) {
    uint64_t count = 0;
    int64_t avail = 0, in_use = 0;
    const int64_t nb = (int64_t)table_type::blocks_in_segment(c);
    for (uint64_t s = 0; s < nseg; s++) {
      if (ids[s] != c || segs[s].active == -1) continue;
      count++;
      avail += segs[s].active;
      in_use += nb - segs[s].active;
    }
    printf("Tree %u: %llu live blocks | avg available %f / %lld | avg in use %f / %lld\n", c,
           (unsigned long long)count, count ? 1.0 * (double)avail / (double)count : 0.0, (long long)nb,
           count ? 1.0 * (double)in_use / (double)count : 0.0, (long long)nb);
  }
  uint32_t ctl[kCtlWords * kCtlStride];
  if (hipMemcpy(ctl, host.control, sizeof(ctl), hipMemcpyDeviceToHost) == hipSuccess)
    printf("Allocator errors %u, rejected frees %u\n", *ctl_word(ctl, kCtlErrors), *ctl_word(ctl, kCtlBadFrees));
  uint64_t tree_words = 0;
  for (int l = 0; l < host.segment_tree.num_layers; l++) tree_words += host.segment_tree.words[l];
  const uint64_t overhead = sizeof(my_type) + tree_words * 8 * (1 + kNumTrees) +
                            nseg * (4 + sizeof(segment_counters)) + nseg * kBlocksPerSegment * (sizeof(Block) + 4) +
                            (uint64_t)total_slots() * kSlotStride * 4 + (kCtlWords + kNumTrees) * kCtlStride * 4;
  printf("Allocator is using %llu bytes of overhead\n", (unsigned long long)overhead);
}

}  // namespace allocators
}  // namespace gallatin
