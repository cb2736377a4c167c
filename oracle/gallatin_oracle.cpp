/*
 * gallatin_oracle.cpp — sequential CPU oracle of the Gallatin allocator.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE. This file restates, one lane at a
 * time and without concurrency, the algorithm of the reference allocator
 * (chillenzer/gallatin, include/gallatin/allocators/). Every function
 * cites the reference lines it follows. Atomic read-modify-writes become plain
 * read-modify-writes because only one lane ever runs.
 *
 * Decisions that differ from a literal reading of the reference (the GPU
 * product makes the same decisions, so single-lane results agree):
 *  Q4  layer::gather_multiple (veb.cuh:323-452) can extend a run below word 0
 *      (index underflow) and its final `rollback(0, ...)` (veb.cuh:447) ORs
 *      never-claimed bits into word 0. Fixed: extension below word 0 rolls
 *      back exactly the claimed bits and fails; the trailing rollback is gone
 *      (it is a no-op on every path that does not corrupt).
 *  Q5  free() treats chunk id 0xFFFF (and the never-assigned ids T, T+1) as
 *      an invalid free instead of a huge segment free (gallatin.cuh:1394).
 *  Q6  a pinned slot left empty by an out-of-memory replace_block
 *      (gallatin.cuh:1089-1116) is refilled by the next group that lands on
 *      it, instead of staying empty for the life of the allocator. (The GPU
 *      marks a slot busy while one lane replaces or refills it, so no refill
 *      can race a replacement; a single lane never observes the marker.)
 *  Q13 a group whose leftover correction (block.cuh:267-281) completes a
 *      block's free count (possible only when the block's last slices go to
 *      a multi-slice request that straddles slice 4095) returns the block to
 *      its segment, as a free that completes the count does
 *      (gallatin.cuh:1792-1798); the reference leaks it.
 *  Retry loops whose repetition would observe an unchanged state (a single
 *  lane holds no lock across calls and nothing else runs) are not repeated:
 *  1000 identical failed iterations of request_new_block_from_tree and one
 *  give the same answer.
 */
#include "gallatin_oracle.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

namespace orc {

constexpr uint64_t FAIL = ~0ull;

// veb.cuh:66-67 BITMASK(n): n == 64 -> all ones; n == 0 -> 0.
static inline uint64_t bitmask(uint64_t n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }
static inline int popc(uint64_t x) { return __builtin_popcountll(x); }
// CUDA __ffsll: 1-based index of the lowest set bit, 0 for x == 0.
static inline int ffs64(uint64_t x) { return x ? __builtin_ctzll(x) + 1 : 0; }
// CUDA __clzll(0) == 64.
static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// alloc_utils.cuh:382-398 get_first_bit_bigger: ceil(log2(x)) for x >= 1.
static inline int first_bit_bigger(uint64_t x) { return 63 - clz64(x) + (popc(x) != 1); }

// alloc_utils.cuh:440-453 __cfcll: number of trailing one bits.
static inline int cfcll(uint64_t bits) {
  int p = popc(bits);
  bool valid = (p == 0 || p == 64);
  return p * valid + (ffs64(~bits) - 1) * (!valid);
}

// murmurhash.cuh:20-61 MurmurHash64A over one 8-byte key (len == 8, no tail).
static inline uint64_t murmur64a(uint64_t key, uint64_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ull;
  const int r = 47;
  uint64_t h = seed ^ (8ull * m);
  uint64_t k = key;
  k *= m;
  k ^= k >> r;
  k *= m;
  h ^= k;
  h *= m;
  h ^= h >> r;
  h *= m;
  h ^= h >> r;
  return h;
}

static inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------- vEB layer
// veb.cuh:109-526 struct layer.
struct Layer {
  uint64_t universe = 0;
  uint64_t num_blocks = 0;
  std::vector<uint64_t> bits;

  // veb.cuh:162-202 generate_on_device_nowait + init_bits (veb.cuh:72-84):
  // every item of the universe starts set.
  explicit Layer(uint64_t items) : universe(items), num_blocks((items - 1) / 64 + 1), bits(num_blocks, 0) {
    for (uint64_t t = 0; t < items; t++) bits[t / 64] |= 1ull << (t % 64);
  }

  // veb.cuh:242-248 insert/remove return the old word (atomicOr/atomicAnd).
  uint64_t insert(uint64_t high, int low) {
    uint64_t old = bits[high];
    bits[high] = old | (1ull << low);
    return old;
  }
  uint64_t remove(uint64_t high, int low) {
    uint64_t old = bits[high];
    bits[high] = old & ~(1ull << low);
    return old;
  }
  uint64_t fetch_and(uint64_t word, uint64_t mask) {
    uint64_t old = bits[word];
    bits[word] = old & mask;
    return old;
  }
  void fetch_or(uint64_t word, uint64_t mask) { bits[word] |= mask; }

  // veb.cuh:250-275 find_next: first set bit strictly above low (any if -1).
  int find_next(uint64_t high, int low) const {
    if (high >= universe) return -1;
    if (high >= num_blocks) return -1;  // unreachable for in-range queries
    if (low == -1) return ffs64(bits[high]) - 1;
    return ffs64(bits[high] & ~bitmask((uint64_t)(low + 1))) - 1;
  }

  // veb.cuh:280-286
  bool query(uint64_t high, int low) const { return (bits[high] >> low) & 1ull; }

  // veb.cuh:305-321 rollback: re-set bits_to_set bits starting at bit 0 of
  // start_block, 64 per word.
  void rollback(uint64_t start_block, uint64_t bits_to_set) {
    while (bits_to_set > 64) {
      fetch_or(start_block, bitmask(64));
      bits_to_set -= 64;
      start_block++;
    }
    fetch_or(start_block, bitmask(bits_to_set));
  }

  // veb.cuh:323-452 gather_multiple, with fix Q4 (see file header).
  uint64_t gather_multiple(uint64_t num_contiguous) {
    uint64_t start_block = num_blocks - 1;
    uint64_t num_needed = num_contiguous;

    while (start_block != 0) {
      num_contiguous = num_needed;
      int contig_start = cfcll(bits[start_block]);
      if (contig_start == 0) {
        start_block -= 1;
        continue;
      }
      if ((uint64_t)contig_start >= num_contiguous) {
        // the top num_contiguous bits of the low run of ones
        uint64_t acq_bitmask = bitmask(num_contiguous) << (contig_start - num_contiguous);
        uint64_t acquired = fetch_and(start_block, ~acq_bitmask);
        if ((uint64_t)popc(acquired & acq_bitmask) == num_contiguous)
          return start_block * 64 + (contig_start - num_contiguous);
        fetch_or(start_block, acq_bitmask & acquired);
        continue;
      }
      uint64_t acquired = fetch_and(start_block, ~bitmask(contig_start));
      bool succeeded = true;
      if (popc(acquired & bitmask(contig_start)) == contig_start) {
        num_contiguous -= contig_start;
        while (num_contiguous > 64) {
          if (start_block == 0) {  // Q4: the run would extend below word 0
            rollback(start_block, num_needed - num_contiguous);
            return FAIL;
          }
          start_block--;
          acquired = fetch_and(start_block, 0ull);
          if (popc(acquired) != 64) {
            fetch_or(start_block, acquired);
            rollback(start_block + 1, num_needed - num_contiguous);
            succeeded = false;
            break;
          }
          num_contiguous -= 64;
        }
        if (!succeeded) continue;
        if (start_block == 0) {  // Q4
          rollback(start_block, num_needed - num_contiguous);
          return FAIL;
        }
        start_block--;
        uint64_t final_bitmask = bitmask(num_contiguous) << (64 - num_contiguous);
        uint64_t final_bits = fetch_and(start_block, ~final_bitmask);
        if ((uint64_t)popc(final_bits & final_bitmask) == num_contiguous)
          return start_block * 64 + 64 - num_contiguous;
        fetch_or(start_block, final_bitmask & final_bits);
        rollback(start_block + 1, num_needed - num_contiguous);
        continue;
      } else {
        fetch_or(start_block, acquired & bitmask(contig_start));
        continue;
      }
    }
    return FAIL;  // Q4: reference's rollback(0, ...) (veb.cuh:447) removed
  }

  // veb.cuh:495-523 return_multiple: set a contiguous band of bits.
  void return_multiple(uint64_t start_index, uint64_t num_indices) {
    uint64_t block_index = start_index / 64;
    uint64_t block_start = start_index % 64;
    while (num_indices + block_start > 64) {
      uint64_t num_added = 64 - block_start;
      fetch_or(block_index, bitmask(num_added) << block_start);
      block_start = 0;
      block_index += 1;
      num_indices -= num_added;
    }
    fetch_or(block_index, bitmask(num_indices) << block_start);
  }
};

// ----------------------------------------------------------------- vEB tree
// veb.cuh:528-1228 struct veb_tree. layers[0] is the top summary,
// layers[num_layers-1] holds one bit per item.
struct Veb {
  uint64_t seed = 0;
  uint64_t total_universe = 0;
  int num_layers = 0;
  std::vector<Layer> layers;

  // veb.cuh:595-649 generate_on_device_nowait.
  Veb(uint64_t universe, uint64_t ext_seed) : seed(ext_seed), total_universe(universe) {
    int max_height = 64 - clz64(universe);
    num_layers = (max_height - 1) / 6 + 1;
    std::vector<Layer> bottom_up;
    uint64_t u = universe;
    for (int i = 0; i < num_layers; i++) {
      bottom_up.emplace_back(u);
      u = (u - 1) / 64 + 1;
    }
    for (int i = num_layers - 1; i >= 0; i--) layers.push_back(bottom_up[i]);
  }

  // gallatin.cuh:136-145 boot_segment_trees removes every item; the end state
  // of that kernel is every word of every layer zero.
  void clear_all() {
    for (auto& l : layers) std::fill(l.bits.begin(), l.bits.end(), 0ull);
  }

  // The GPU's out-of-memory shortcut reads: the top layer's first word is
  // empty and it is the only word of its layer or summarises a lower layer
  // (veb.hpp top_looks_empty); the number of items (the free-segment gauge
  // for the segment tree).
  bool top_looks_empty() const { return layers[0].bits[0] == 0ull && (num_layers > 1 || layers[0].num_blocks == 1); }
  uint64_t count_items() const {
    uint64_t n = 0;
    for (uint64_t w : layers[num_layers - 1].bits) n += (uint64_t)popc(w);
    return n;
  }

  // veb.cuh:697-712
  static bool float_up(int& layer, uint64_t& high, int& low) {
    layer -= 1;
    low = (int)(high & bitmask(6));
    high = high >> 6;
    return layer >= 0;
  }
  bool float_down(int& layer, uint64_t& high, int& low) const {
    layer += 1;
    high = (high << 6) + (uint64_t)low;
    low = -1;
    return layer < num_layers;
  }

  // veb.cuh:737-755
  bool remove(uint64_t v) {
    uint64_t high = v >> 6;
    int low = (int)(v & bitmask(6));
    int layer = num_layers - 1;
    uint64_t old = layers[layer].remove(high, low);
    if (!(old & (1ull << low))) return false;
    while (popc(old) == 1 && float_up(layer, high, low)) old = layers[layer].remove(high, low);
    return true;
  }

  // veb.cuh:775-798
  bool insert_force_update(uint64_t v) {
    uint64_t high = v >> 6;
    int low = (int)(v & bitmask(6));
    int layer = num_layers - 1;
    uint64_t old = layers[layer].insert(high, low);
    if (old & (1ull << low)) return false;
    while (popc(old) == 0 && float_up(layer, high, low)) old = layers[layer].insert(high, low);
    return true;
  }

  // veb.cuh:824-829
  bool query(uint64_t v) const { return layers[num_layers - 1].query(v >> 6, (int)(v & 63)); }

  // veb.cuh:878-916 successor_thorough: first item strictly above q visible
  // through the summaries; stale summary bits are skipped by re-floating.
  uint64_t successor_thorough(uint64_t q) const {
    uint64_t high = q >> 6;
    int low = (int)(q & bitmask(6));
    int layer = num_layers - 1;
    while (true) {
      while (true) {
        int found = layers[layer].find_next(high, low);
        if (found == -1) {
          if (layer == 0) return FAIL;
          float_up(layer, high, low);
          continue;
        }
        break;
      }
      while (layer != num_layers - 1) {
        low = layers[layer].find_next(high, low);
        if (low == -1) break;
        float_down(layer, high, low);
      }
      low = layers[layer].find_next(high, low);
      if (low != -1) return (high << 6) + (uint64_t)low;
    }
  }

  // veb.cuh:918-937
  uint64_t lock_offset(uint64_t start) {
    if (query(start) && remove(start)) return start;
    while (true) {
      start = successor_thorough(start);
      if (start == FAIL) return start;
      if (remove(start)) return start;
    }
  }

  // veb.cuh:946-977; tid stands for threadIdx.x + blockIdx.x * blockDim.x.
  uint64_t find_random_valid_index(uint64_t tid) const {
    uint64_t hash1 = murmur64a(tid, seed);
    uint64_t hash2 = murmur64a(tid, hash1);
    for (int attempts = 0; attempts < 15; attempts++) {  // VEB_MAX_ATTEMPTS, veb.cuh:59
      uint64_t index_to_start = (hash1 + (uint64_t)attempts * hash2) % total_universe;
      if (query(index_to_start)) return index_to_start;
      uint64_t index = successor_thorough(index_to_start);
      if (index != FAIL) return index;
    }
    return successor_thorough(0);
  }

  // veb.cuh:979-991
  uint64_t malloc_first() {
    for (int attempts = 0; attempts < 15; attempts++) {
      uint64_t offset = lock_offset(0);
      if (offset != FAIL) return offset;
    }
    return FAIL;
  }

  // veb.cuh:1100-1126: clear a summary bit whose lower word is empty.
  void maybe_remove_layer_bits(int layer, uint64_t start_bit, uint64_t bits_to_check) {
    for (uint64_t wb = start_bit; wb < start_bit + bits_to_check; wb++) {
      if (layers[layer + 1].bits[wb] == 0) layers[layer].remove(wb / 64, (int)(wb % 64));
    }
  }

  // veb.cuh:1128-1154: set a summary bit only if its lower word is full.
  void maybe_return_layer_bits(int layer, uint64_t start_bit, uint64_t bits_to_check) {
    for (uint64_t wb = start_bit; wb < start_bit + bits_to_check; wb++) {
      if (popc(layers[layer + 1].bits[wb]) == 64) layers[layer].insert(wb / 64, (int)(wb % 64));
    }
  }

  // veb.cuh:1160-1196
  uint64_t gather_multiple(uint64_t num_contiguous) {
    uint64_t found = layers[num_layers - 1].gather_multiple(num_contiguous);
    uint64_t original_found = found;
    if (found == FAIL) return FAIL;
    for (int i = num_layers - 2; i >= 0; i--) {
      uint64_t tread_high = (found + num_contiguous - 1) / 64 + 1;
      uint64_t tread_low = found / 64;
      found = tread_low;
      num_contiguous = tread_high - tread_low;
      maybe_remove_layer_bits(i, found, num_contiguous);
    }
    return original_found;
  }

  // veb.cuh:1198-1218
  void return_multiple(uint64_t start_bit, uint64_t num_contiguous) {
    layers[num_layers - 1].return_multiple(start_bit, num_contiguous);
    for (int i = num_layers - 2; i >= 0; i--) {
      uint64_t tread_high = (start_bit + num_contiguous - 1) / 64 + 1;
      uint64_t tread_low = start_bit / 64;
      start_bit = tread_low;
      num_contiguous = tread_high - tread_low;
      maybe_return_layer_bits(i, start_bit, num_contiguous);
    }
  }
};

// -------------------------------------------------------------------- Block
// block.cuh:59-398. malloc_counter = (tree << 20) | count; free_counter counts
// returned slices up to 4096.
constexpr uint32_t TREE_OFFSET = 20;  // GALLATIN_BLOCK_TREE_OFFSET, block.cuh:53

struct Block {
  uint32_t malloc_counter = 4097u;  // init(), block.cuh:64-68
  uint32_t free_counter = 0u;
};

static inline uint32_t clip_count(uint32_t merged) { return merged & (uint32_t)bitmask(TREE_OFFSET); }  // block.cuh:391
static inline bool check_valid(uint32_t merged, uint32_t tree) { return (merged >> TREE_OFFSET) == tree; }  // block.cuh:343

// -------------------------------------------------------------- memory table
// memory_table.cuh:229-1356 alloc_table<SEG, MIN>.
struct Table {
  uint64_t seg_bytes = 0, min_size = 0, num_segments = 0, bps = 0;
  std::vector<uint16_t> chunk_ids;  // 0xFFFF free
  std::vector<Block> blocks;        // num_segments * bps
  std::vector<uint32_t> queues;     // block id + 1, 0 = nullptr
  std::vector<int32_t> active;      // active_counts
  std::vector<uint32_t> q_alloc, q_free, q_final;  // queue_counters, queue_free_counters, final_queue_free_counters

  // memory_table.cuh:471-647 (device_only pool) + gallatin_init_counters_kernel :173-201
  Table(uint64_t seg, uint64_t min, uint64_t nseg)
      : seg_bytes(seg), min_size(min), num_segments(nseg), bps(seg / (min * 4096)),
        chunk_ids(nseg, 0xFFFF), blocks(nseg * bps), queues(nseg * bps, 0u), active(nseg, -1),
        q_alloc(nseg, 0u), q_free(nseg, 0u), q_final(nseg, 0u) {}

  uint64_t tree_alloc_size(uint32_t tree) const { return min_size << tree; }  // :1046-1049
  uint64_t blocks_per_segment(uint32_t tree) const { return seg_bytes / (tree_alloc_size(tree) * 4096); }  // :1057-1061
  uint64_t max_allocations_per_segment() const { return seg_bytes / min_size; }  // :1065-1070

  // :823-827 CAS 0xFFFF -> id
  bool set_tree_id(uint64_t seg, uint16_t id) {
    if (chunk_ids[seg] != 0xFFFF) return false;
    chunk_ids[seg] = id;
    return true;
  }
  // :847-851 CAS id -> 0xFFFF
  bool reset_tree_id(uint64_t seg, uint16_t id) {
    if (chunk_ids[seg] != id) return false;
    chunk_ids[seg] = 0xFFFF;
    return true;
  }

  // :746-818
  bool setup_segment(uint64_t seg, uint32_t tree) {
    if (!set_tree_id(seg, (uint16_t)tree)) return false;
    uint64_t nb = blocks_per_segment(tree);
    for (uint64_t i = 0; i < nb; i++) queues[seg * bps + i] = 0u;
    q_alloc[seg] = 0u;
    q_free[seg] = 0u;
    q_final[seg] = 0u;
    active[seg] = (int32_t)nb - 1;
    return true;
  }

  // :915-993 get_block; returns a global block id or -1 (nullptr).
  int64_t get_block(uint64_t seg, uint32_t tree, bool& last, uint64_t& faults) {
    last = false;
    int32_t active_count = active[seg]--;  // get_slot_in_segment, :866-868
    if (active_count < 0) {
      active[seg]++;  // return_slot_to_segment
      return -1;
    }
    if (chunk_ids[seg] != tree) {
      active[seg]++;
      return -1;
    }
    uint64_t nb = blocks_per_segment(tree);
    int32_t queue_pos = (int32_t)(q_alloc[seg]++);
    int64_t block;
    if ((uint64_t)queue_pos < nb) {
      block = (int64_t)(seg * bps + (uint64_t)queue_pos);
    } else {
      uint64_t wrapped = (uint64_t)queue_pos % nb;
      uint32_t v = queues[seg * bps + wrapped];  // atomicExch with nullptr
      queues[seg * bps + wrapped] = 0u;
      if (v == 0u) {  // the reference traps here (:977-982)
        faults++;
        return -1;
      }
      block = (int64_t)v - 1;
    }
    blocks[block].malloc_counter = tree << TREE_OFFSET;  // Block::init_malloc, block.cuh:215-224
    if (active_count == 0) last = true;
    return block;
  }

  // :1158-1222 reserve_segment_slot (+ finalize_free_queue :906-910)
  uint32_t reserve_segment_slot(uint64_t block, uint64_t seg, uint64_t nb, uint64_t& faults) {
    uint32_t pos = q_free[seg]++;
    uint32_t live = (uint32_t)(pos % nb);
    // with one lane the slot's previous entry has always been taken: a segment
    // with all nb blocks queued has no block left to return
    if (queues[seg * bps + live] != 0u) faults++;
    queues[seg * bps + live] = (uint32_t)(block + 1);
    if (q_final[seg] != pos) faults++;  // the CAS spin would never end
    q_final[seg] = pos + 1;
    return live;
  }

  // :1226-1241 finish_freeing_block
  bool finish_freeing_block(uint64_t seg, uint64_t nb) {
    int32_t return_id = active[seg]++;
    if ((int64_t)return_id == (int64_t)nb - 2) {  // all_blocks_free, :876-880
      if (active[seg] == (int32_t)nb - 1) {
        active[seg] = -1;
        return true;
      }
    }
    return false;
  }
};

// gallatin.cuh:1646: re-list a segment in its class tree when the slot a
// returned block takes crosses a tenth of the segment: 1.0*slot/nb >= .1 &&
// 1.0*(slot-1)/nb < .1 (REREGISTER_CUTOFF, :123). nb is a power of two, so
// no slot/nb equals 0.1 and the test is exactly 10*slot > nb &&
// 10*(slot-1) < nb; slot == 0 makes the unsigned slot-1 wrap, and the second
// clause is false.
static inline bool reregister_now(uint32_t slot, uint64_t nb) {
  return slot >= 1 && 10ull * slot > nb && 10ull * (slot - 1) < nb;
}

// ---------------------------------------------------------------- allocator
struct Alloc {
  uint64_t SEG = 0, MIN = 0, MAX = 0, seed = 0, pool_bytes = 0;
  uint32_t num_trees = 0;
  int smallest_bits = 0;
  uint32_t locks = 0;
  Veb segment_tree;
  std::vector<Veb> sub_trees;
  Table table;
  std::vector<std::vector<uint32_t>> slots;     // pinned blocks: block id + 1
  std::vector<std::vector<uint32_t>> overflow;  // overflow slots of each class (block id + 1)
  uint64_t faults = 0;

  Alloc(uint64_t seg, uint64_t min, uint64_t max, uint64_t nseg, uint64_t sd)
      : SEG(seg), MIN(min), MAX(max), seed(sd), segment_tree(nseg, sd), table(seg, min, nseg) {}

  // gallatin.cuh:1513-1519 (try-lock bit `tree` of `locks`)
  bool acquire_tree_lock(uint32_t tree) {
    uint32_t mask = 1u << tree;
    bool ok = (locks & mask) == 0;
    locks |= mask;
    return ok;
  }
  void release_tree_lock(uint32_t tree) { locks &= ~(1u << tree); }

  // shared_block_storage.cuh:119-134 get_valid_block_index; `slot` stands for smid.
  int valid_block_index(uint32_t tree, uint32_t slot) const {
    const auto& s = slots[tree];
    int n = (int)s.size();
    int my = (int)(slot % (uint32_t)n);
    int original = my;
    int counter = 0;
    while (s[my] == 0u && my != (original - 1)) {
      my = (my + 1) % n;
      counter += 1;
      if (counter >= 30) break;  // SHARED_BLOCK_COUNTER_CUTOFF, :31
    }
    return my;
  }

  // gallatin.cuh:1465-1510
  int64_t gather_new_segment(uint32_t tree) {
    uint64_t seg = segment_tree.malloc_first();
    if (seg == FAIL) return -1;
    if (!table.setup_segment(seg, tree)) {
      faults++;  // reference traps
      return -1;
    }
    sub_trees[tree].insert_force_update(seg);
    return (int64_t)seg;
  }

  // gallatin.cuh:1525-1631; returns a global block id or -1.
  int64_t request_new_block_from_tree(uint32_t tree, uint64_t tid) {
    for (int attempts = 0; attempts < 1000;) {  // REQUEST_BLOCK_MAX_ATTEMPTS
      uint64_t segment = sub_trees[tree].find_random_valid_index(tid);
      if (segment == FAIL) {
        if (!acquire_tree_lock(tree)) {  // cannot happen with one lane
          faults++;
          return -1;
        }
        int64_t success = gather_new_segment(tree);
        release_tree_lock(tree);
        if (success == -1) {
          // nothing changed: every further iteration fails the same way
          return -1;
        }
        segment = (uint64_t)success;
      }
      bool last = false;
      int64_t block = table.get_block(segment, tree, last, faults);
      if (last) {
        sub_trees[tree].remove(segment);
        if (acquire_tree_lock(tree)) {
          gather_new_segment(tree);
          release_tree_lock(tree);
        }
      }
      if (block >= 0) return block;
      // get_block refused a segment its sub-tree listed: only a concurrent
      // run can produce that state.
      faults++;
      return -1;
    }
    return -1;
  }

  // gallatin.cuh:1637-1726 return_block
  void return_block(uint64_t block, uint64_t segment, uint32_t tree) {
    uint64_t nb = table.blocks_per_segment(tree);
    uint32_t reserved_slot = table.reserve_segment_slot(block, segment, nb, faults);
    if (reregister_now(reserved_slot, nb)) {
      sub_trees[tree].insert_force_update(segment);
    }
    bool need_to_deregister = table.finish_freeing_block(segment, nb);
    if (need_to_deregister) {
      sub_trees[tree].remove(segment);
      if (!table.reset_tree_id(segment, (uint16_t)tree)) faults++;  // reference traps
      segment_tree.insert_force_update(segment);
    }
  }

  // gallatin.cuh:1732-1766 free_block (device pool, no calloc)
  void free_block(uint64_t block) {
    uint64_t segment = block / table.bps;
    uint32_t tree = table.chunk_ids[segment];
    return_block(block, segment, tree);
  }

  // gallatin.cuh:1774-1827 free_offset for a team of one lane.
  void free_offset(uint64_t offset) {
    uint64_t block_id = offset / 4096;
    Block& b = table.blocks[block_id];
    uint32_t old = b.free_counter;
    b.free_counter = old + 1;  // block_free_multiple(1), block.cuh:87-98
    if (old + 1 == 4096) {
      b.free_counter = 0;  // reset_free, block.cuh:200-212
      free_block(block_id);
    }
  }

  // slot `index` of class `tree`: its pinned slots first, then its overflow slots
  uint32_t& slot_ref(uint32_t tree, uint64_t index) {
    const uint64_t pinned = slots[tree].size();
    return index < pinned ? slots[tree][index] : overflow[tree][index - pinned];
  }

  // Slots of class `tree` a group may start at under load: its pinned slots
  // and overflow slots holding at most 1/16 of the pool in blocks, at least
  // its pinned slots (the GPU's usable_slots_of, kSlotShare = 16).
  uint64_t usable_slots(uint32_t tree) const {
    const uint64_t budget = table.num_segments * SEG / 16 / (table.tree_alloc_size(tree) * 4096);
    const uint64_t lo = slots[tree].size(), hi = lo + overflow[tree].size();
    return budget < lo ? lo : budget > hi ? hi : budget;
  }

  // gallatin.cuh:1083-1138 replace_block (+ Q6: no trap on a racing refill)
  bool replace_block(uint32_t tree, int slot_index, uint64_t block, uint64_t tid) {
    uint32_t& s = slot_ref(tree, (uint64_t)slot_index);
    if (s != (uint32_t)(block + 1)) return true;  // swap_out_block CAS failed
    s = 0u;
    int64_t nb = request_new_block_from_tree(tree, tid);
    if (nb < 0) return false;
    if (s != 0u) {  // swap_out_nullptr CAS failed
      free_block((uint64_t)nb);
      return false;
    }
    s = (uint32_t)(nb + 1);
    return true;
  }

  // gallatin.cuh:865-1078 malloc_slice_allocation for a group of one lane:
  // the exclusive scan is 0 and the group sum is the lane's own count.
  uint64_t malloc_slice_allocation(uint32_t tree, uint32_t alloc_count, uint64_t tid, uint32_t slot) {
    for (int num_attempts = 0; num_attempts < 500; num_attempts++) {  // GALLATIN_MAX_ATTEMPTS
      int index = valid_block_index(tree, slot);
      uint32_t entry = slots[tree][index];
      if (entry == 0u) {
        // Q6: refill the empty slot.
        int64_t fresh = request_new_block_from_tree(tree, tid);
        if (fresh < 0) return FAIL;
        if (slots[tree][index] == 0u) {
          slots[tree][index] = (uint32_t)(fresh + 1);
        } else {
          free_block((uint64_t)fresh);
        }
        continue;
      }
      uint64_t global_block_id = entry - 1;
      Block& b = table.blocks[global_block_id];
      // block_malloc_tree_multi_size, block.cuh:245-264
      uint32_t merged_count = b.malloc_counter;
      b.malloc_counter = merged_count + alloc_count;
      // extract_count_multi_size, block.cuh:373-388
      uint64_t allocation = (uint64_t)clip_count(merged_count);
      bool should_replace = (allocation <= 4095 && (allocation + alloc_count) > 4095);
      bool start_valid = (allocation <= 4095);
      bool end_valid = (allocation + alloc_count <= 4096);
      uint32_t leftover = (uint32_t)((start_valid && end_valid) * (alloc_count - 1) +
                                     (start_valid && !end_valid) * (4096 - allocation));
      // block_correct_frees, block.cuh:267-281
      bool completed_by_leftover = false;
      if (leftover > 0) {
        uint32_t old = b.free_counter;
        b.free_counter = old + leftover;
        completed_by_leftover = (old + leftover == 4096);
      }
      if (allocation + alloc_count > 4096) allocation = FAIL;
      if (should_replace) replace_block(tree, index, global_block_id, tid);
      if (completed_by_leftover) {  // Q13
        table.blocks[global_block_id].free_counter = 0;
        free_block(global_block_id);
      }
      if (allocation != FAIL) {
        if (!check_valid(merged_count, tree)) {
          free_offset(allocation + global_block_id * 4096);
        } else {
          return allocation + global_block_id * 4096;
        }
      }
    }
    return FAIL;
  }

  // gallatin.cuh:828-861 malloc_block_allocation
  uint64_t malloc_block_allocation(uint32_t tree, uint64_t tid) {
    int64_t block = request_new_block_from_tree(tree, tid);
    if (block < 0) return FAIL;
    Block& b = table.blocks[block];
    // malloc_fill_block, block.cuh:288-340
    uint32_t old_merged = b.malloc_counter;
    b.malloc_counter = old_merged + 4096;
    uint32_t old_count = clip_count(old_merged);
    if (!check_valid(old_merged, tree)) faults++;  // reference traps
    b.free_counter += 4095 - old_count;
    if (old_count != 0) {
      free_offset((uint64_t)block * 4096);
      return FAIL;
    }
    return (uint64_t)block * 4096;
  }

  // gallatin.cuh:790-826 malloc_segment_allocation
  uint64_t malloc_segment_allocation(uint64_t n) {
    acquire_tree_lock(num_trees);
    uint64_t alloc_index = segment_tree.gather_multiple(n);
    if (alloc_index == FAIL) {
      release_tree_lock(num_trees);
      return FAIL;
    }
    if (!table.set_tree_id(alloc_index, (uint16_t)(num_trees + 1 + n))) faults++;  // reference traps
    release_tree_lock(num_trees);
    return alloc_index * table.bps * 4096;
  }

  // memory_table.cuh:1072-1086 offset_to_allocation, as a byte offset.
  uint64_t offset_to_bytes(uint64_t offset, uint32_t tree) const {
    uint64_t per = table.max_allocations_per_segment();
    return (offset / per) * SEG + (offset % per) * table.tree_alloc_size(tree);
  }

  // gallatin.cuh:1141-1147
  uint32_t tree_id_from_size(uint64_t size) const {
    if (size < MIN) return 0;
    return (uint16_t)(first_bit_bigger(size) - smallest_bits);
  }

  // gallatin.cuh:1186-1276 malloc
  uint64_t malloc(uint64_t size, uint64_t tid, uint32_t slot) {
    uint32_t tree_id = tree_id_from_size(size);
    uint32_t alloc_count = 1;
    if (tree_id >= num_trees) {
      int smallest_tree_bits = first_bit_bigger(MIN * 4096);
      int block_tree = first_bit_bigger(size) - smallest_tree_bits;
      if (block_tree < 0) {
        alloc_count = (uint32_t)(1ull << (tree_id - (num_trees - 1)));
        tree_id = num_trees - 1;
      } else if ((uint32_t)block_tree < num_trees) {
        // GALLATIN_MALLOC_BLOCK_ATTEMPTS retries repeat an unchanged failure
        uint64_t offset = malloc_block_allocation((uint32_t)block_tree, tid);
        if (offset == FAIL) return FAIL;
        return offset_to_bytes(offset, (uint32_t)block_tree);
      } else {
        uint64_t n = (size - 1) / SEG + 1;
        if (n + num_trees + 1 >= 0xFFFFull) return FAIL;  // chunk id would overflow u16
        uint64_t offset = malloc_segment_allocation(n);
        if (offset == FAIL) return FAIL;
        return offset_to_bytes(offset, 0);
      }
    }
    uint64_t offset = FAIL;
    for (int attempt = 0; attempt < 10 && offset == FAIL; attempt++)  // GALLATIN_MALLOC_LOOP_ATTEMPTS
      offset = malloc_slice_allocation(tree_id, alloc_count, tid, slot);
    if (offset == FAIL) return FAIL;
    return offset_to_bytes(offset, tree_id);
  }

  // gallatin.cuh:1378-1461 free (+ Q5)
  int free(uint64_t byte_offset) {
    if (byte_offset >= table.num_segments * SEG) return GAL_ERR_INVALID;
    uint64_t segment = byte_offset / SEG;
    uint32_t tree_id = table.chunk_ids[segment];
    if (tree_id == 0xFFFF || tree_id == num_trees || tree_id == num_trees + 1) return GAL_ERR_INVALID;
    if (tree_id > num_trees) {
      // submit_segment_for_free, gallatin.cuh:1283-1302
      uint64_t size = tree_id - num_trees - 1;
      segment_tree.return_multiple(segment, size);
      if (!table.reset_tree_id(segment, (uint16_t)tree_id)) faults++;
      return GAL_OK;
    }
    // memory_table.cuh:1090-1127 allocation_to_offset
    uint64_t segment_byte_offset = byte_offset - segment * SEG;
    uint64_t offset = segment_byte_offset / table.tree_alloc_size(tree_id) +
                      segment * table.max_allocations_per_segment();
    free_offset(offset);
    return GAL_OK;
  }

  // ---- one warp's call site executed by several lanes together.
  //
  // The reference forms its teams with cg::coalesced_threads() and
  // labeled_partition by size class (gallatin.cuh:881-885): every active lane
  // asking for one class is in one team, whose thread ranks follow lane
  // order. Per team and attempt: rank 0 picks the pinned slot
  // (get_valid_block_index, with the warp's smid) and shares the block
  // (:887-894); an exclusive scan of the lanes' slice counts in rank order
  // (:982); the team's last rank adds the team total to the block's malloc
  // counter and broadcasts the old value (block.cuh:245-264); each lane's
  // allocation is old count + its prefix (block.cuh:373-388); the leftovers
  // of fully valid multi-slice lanes (count - 1) and of the lane that
  // straddles slice 4095 (4096 - allocation) are summed and added to the free
  // counter once (block.cuh:267-281); if any lane's range covers slice 4095,
  // rank 0 replaces the block (:1015-1026). Lanes that got no slice loop
  // again and re-form teams among themselves.
  //
  // Teams of different classes touch different slots and blocks, so their
  // claims commute; their rare work (refill, replace, recycle) changes the
  // shared trees and runs team by team in the order of the teams' first
  // lanes. The lanes of one call site run in lockstep, so a lane's attempt
  // budget (10 calls of up to 500 passes, :1263-1268, :878) advances with
  // every other lane still looking. A team that finds its slot empty refills
  // it (Q6) and claims in the next pass; one whose refill finds no memory
  // fails for all its lanes.

  struct Team {
    uint64_t lanes = 0;
    uint32_t tree = 0;
    int index = 0;
    uint32_t entry = 0;
    uint64_t block = 0;
    uint32_t c0 = 0;
    bool empty = false, replace = false, completed = false, stale = false;
    uint32_t give_back = 0;
    // the under-load first look: the class found out of memory; the team
    // refills the empty slot the look landed on (a failed refill then sends
    // it to the pinned walk in the next pass instead of failing)
    bool dry = false, look_refill = false;
  };

  // The GPU's first look under load (more than one lane of the wave
  // claiming, the first pass of a malloc call): the team starts at slot
  // hint % usable of its class, pinned or overflow, instead of the pinned
  // walk. A slot holding a block is taken. An empty slot is taken (and
  // refilled) unless the class is out of memory: its tree empty with no free
  // segment left, or no free segment left and the leader's hashed search
  // finding nothing; when it is, and every pinned slot holds 0 as well, the
  // team fails at once. A class whose usable slots are its pinned slots
  // looks at slot hint % pinned, where the walk starts, and only the failure
  // test applies to an empty slot there. Returns whether the look placed
  // the team (t.index set).
  bool first_look(Team& t, uint32_t hint, uint64_t leader_tid) {
    const uint64_t pinned = slots[t.tree].size();
    const uint64_t usable = usable_slots(t.tree);
    const bool spreads = usable > pinned;
    const uint64_t si = hint % (spreads ? usable : pinned);
    const uint64_t free_segments = segment_tree.count_items();
    const bool tree_dry = sub_trees[t.tree].top_looks_empty() && free_segments == 0;
    uint64_t empty_pinned = 0;
    for (uint32_t e : slots[t.tree]) empty_pinned += e == 0u;
    const bool all_empty = empty_pinned >= pinned;
    const uint32_t se = slot_ref(t.tree, si);
    if (se != 0u) {
      t.index = (int)si;
      return true;
    }
    if (spreads) {
      const bool spare = free_segments != 0;
      t.dry = tree_dry || (!spare && sub_trees[t.tree].find_random_valid_index(leader_tid) == FAIL);
      if (!t.dry || all_empty) {
        t.index = (int)si;
        t.look_refill = !t.dry;
        return true;
      }
      return false;
    }
    if (tree_dry && all_empty) {
      t.dry = true;
      t.index = (int)si;
      return true;
    }
    return false;
  }

  // one pass of the slice claim over the lanes of `looking`; `look`: the
  // teams take the first look under load (see first_look)
  void team_pass(uint64_t looking, const uint32_t* tree, const uint32_t* count, const uint64_t* tid, uint32_t slot,
                 bool look, uint64_t* result, bool* done, bool* oom) {
    std::vector<Team> teams;
    for (uint64_t todo = looking; todo;) {
      const int first = ffs64(todo) - 1;
      Team t;
      t.tree = tree[first];
      for (uint64_t m = todo; m; m &= m - 1) {
        const int l = ffs64(m) - 1;
        if (tree[l] == t.tree) t.lanes |= 1ull << l;
      }
      todo &= ~t.lanes;
      teams.push_back(t);
    }
    // claims: each team on its own class's slot and block
    for (Team& t : teams) {
      if (!(look && first_look(t, slot, tid[ffs64(t.lanes) - 1]))) t.index = valid_block_index(t.tree, slot);
      t.entry = slot_ref(t.tree, (uint64_t)t.index);
      t.empty = t.entry == 0u;
      if (t.empty) continue;
      t.block = t.entry - 1;
      Block& b = table.blocks[t.block];
      uint32_t total = 0;
      for (uint64_t m = t.lanes; m; m &= m - 1) total += count[ffs64(m) - 1];
      const uint32_t merged = b.malloc_counter;
      b.malloc_counter = merged + total;
      t.c0 = clip_count(merged);
      t.stale = !check_valid(merged, t.tree);
      uint32_t prefix = 0, held = 0;
      bool crosses = false;
      for (uint64_t m = t.lanes; m; m &= m - 1) {
        const int l = ffs64(m) - 1;
        const uint64_t allocation = (uint64_t)t.c0 + prefix;
        crosses |= allocation <= 4095 && allocation + count[l] > 4095;
        if (allocation + count[l] <= 4096) {
          held++;
          if (!t.stale) {
            result[l] = allocation + t.block * 4096;
            done[l] = true;
          }
        }
        prefix += count[l];
      }
      const uint32_t in_range = t.c0 >= 4096 ? 0u : std::min<uint32_t>(t.c0 + total, 4096u) - t.c0;
      const uint32_t written_off = in_range - held;
      if (written_off > 0) {
        const uint32_t old = b.free_counter;
        b.free_counter = old + written_off;
        t.completed = old + written_off == 4096;
      }
      t.give_back = t.stale ? held : 0u;
      t.replace = crosses || t.c0 > 4095 || t.stale;
    }
    // rare work, team by team in the order of their first lanes
    for (Team& t : teams) {
      const uint64_t leader_tid = tid[ffs64(t.lanes) - 1];
      if (t.empty) {
        // a class the first look found out of memory takes no block request
        int64_t fresh = t.dry ? -1 : request_new_block_from_tree(t.tree, leader_tid);
        uint32_t& s = slot_ref(t.tree, (uint64_t)t.index);
        if (fresh >= 0 && s == 0u) {
          s = (uint32_t)(fresh + 1);
        } else if (fresh < 0 && t.look_refill) {
          // the look's slot could not be refilled: the next pass walks the
          // pinned slots, where blocks can still be
        } else {
          if (fresh >= 0) free_block((uint64_t)fresh);
          for (uint64_t m = t.lanes; m; m &= m - 1) {
            const int l = ffs64(m) - 1;
            done[l] = true;
            oom[l] = true;
            result[l] = FAIL;
          }
        }
        continue;
      }
      if (t.replace) replace_block(t.tree, t.index, t.block, leader_tid);
      if (t.completed) {  // Q13
        table.blocks[t.block].free_counter = 0;
        free_block(t.block);
      }
      if (t.give_back > 0) {
        Block& b = table.blocks[t.block];
        const uint32_t old = b.free_counter;
        b.free_counter = old + t.give_back;
        if (old + t.give_back == 4096) {
          b.free_counter = 0;
          free_block(t.block);
        }
      }
    }
  }

  // malloc called by the lanes of `lanes` together. Requests above the slice
  // range are served one lane at a time in lane order (they take no team
  // claim); a call site mixing both kinds is not modelled (the order of the
  // two branches is up to the compiler).
  // `spread`: the GPU's malloc under load (first look, overflow slots);
  // false: the reference's pinned walk for every pass.
  int group_malloc(uint64_t lanes, const uint64_t* size, const uint64_t* tid, uint32_t slot, bool spread,
                   uint64_t* out) {
    uint32_t tree[64] = {}, count[64] = {};
    bool large = false, small = false;
    for (uint64_t m = lanes; m; m &= m - 1) {
      const int l = ffs64(m) - 1;
      uint32_t t = tree_id_from_size(size[l]);
      uint32_t c = 1;
      if (t >= num_trees) {
        if (first_bit_bigger(size[l]) - first_bit_bigger(MIN * 4096) >= 0) {
          large = true;
          continue;
        }
        c = (uint32_t)(1ull << (t - (num_trees - 1)));
        t = num_trees - 1;
      }
      small = true;
      tree[l] = t;
      count[l] = c;
    }
    if (large && small) return GAL_ERR_INVALID;
    if (large) {
      for (uint64_t m = lanes; m; m &= m - 1) {
        const int l = ffs64(m) - 1;
        out[l] = malloc(size[l], tid[l], slot);
      }
      return GAL_OK;
    }
    uint64_t result[64];
    bool done[64], oom[64];
    uint32_t calls[64] = {};
    for (int l = 0; l < 64; l++) {
      result[l] = FAIL;
      oom[l] = false;
    }
    // malloc's attempt loop (gallatin.cuh:1263-1268): lanes still without an
    // allocation and not out of memory call the slice claim again together
    for (;;) {
      uint64_t calling = 0;
      for (uint64_t m = lanes; m; m &= m - 1) {
        const int l = ffs64(m) - 1;
        if (result[l] == FAIL && !oom[l] && calls[l] < 10) calling |= 1ull << l;
      }
      if (!calling) break;
      uint32_t tries[64] = {};
      for (int l = 0; l < 64; l++) done[l] = false;
      for (uint64_t m = calling; m; m &= m - 1) calls[ffs64(m) - 1]++;
      uint64_t looking = calling;
      for (bool first = true; looking; first = false) {
        // the look needs more than one lane claiming (of any class)
        team_pass(looking, tree, count, tid, slot, spread && first && (looking & (looking - 1)) != 0, result, done,
                  oom);
        uint64_t still = 0;
        for (uint64_t m = looking; m; m &= m - 1) {
          const int l = ffs64(m) - 1;
          if (!done[l]) tries[l]++;
          if (!done[l] && tries[l] < 500) still |= 1ull << l;
        }
        looking = still;
      }
    }
    for (uint64_t m = lanes; m; m &= m - 1) {
      const int l = ffs64(m) - 1;
      out[l] = result[l] == FAIL ? FAIL : offset_to_bytes(result[l], tree[l]);
    }
    return GAL_OK;
  }

  // free called by the lanes of `lanes` together on slice or whole-block
  // allocations (reference gallatin.cuh:1774-1827): lanes are grouped by
  // block (labeled_partition by block id), each group's rank 0 adds the
  // group size to the block's free counter, and the blocks whose count
  // completes are returned group by group in lane order.
  int group_free(uint64_t lanes, const uint64_t* byte_offset) {
    uint64_t block_of[64] = {};
    for (uint64_t m = lanes; m; m &= m - 1) {
      const int l = ffs64(m) - 1;
      const uint64_t off = byte_offset[l];
      if (off >= table.num_segments * SEG) return GAL_ERR_INVALID;
      const uint64_t segment = off / SEG;
      const uint32_t tree_id = table.chunk_ids[segment];
      if (tree_id >= num_trees) return GAL_ERR_INVALID;
      const uint64_t offset = (off - segment * SEG) / table.tree_alloc_size(tree_id) +
                              segment * table.max_allocations_per_segment();
      block_of[l] = offset / 4096;
    }
    std::vector<uint64_t> completed;
    for (uint64_t todo = lanes; todo;) {
      const uint64_t block = block_of[ffs64(todo) - 1];
      uint32_t n = 0;
      for (uint64_t m = todo; m; m &= m - 1) {
        const int l = ffs64(m) - 1;
        if (block_of[l] == block) {
          n++;
          todo &= ~(1ull << l);
        }
      }
      Block& b = table.blocks[block];
      const uint32_t old = b.free_counter;
      b.free_counter = old + n;
      if (old + n == 4096) completed.push_back(block);
    }
    for (uint64_t block : completed) {
      table.blocks[block].free_counter = 0;
      free_block(block);
    }
    return GAL_OK;
  }

  // gallatin.cuh:726-787 boot_block
  bool boot_block(uint32_t tree, int slot) {
    int64_t block = request_new_block_from_tree(tree, 0);
    if (block < 0) return false;
    uint64_t seg = (uint64_t)block / table.bps;
    if (table.chunk_ids[seg] != tree) return false;
    if (slots[tree][slot] != 0u) return false;
    slots[tree][slot] = (uint32_t)(block + 1);
    return true;
  }

  // gallatin.cuh:324-421 generate_on_device, with the pinned blocks booted in
  // the order of boot_shared_block_container_one_thread (gallatin.cuh:196-221).
  bool boot() {
    num_trees = (uint32_t)(first_bit_bigger(MAX) - first_bit_bigger(MIN) + 1);
    smallest_bits = first_bit_bigger(MIN);
    for (uint32_t i = 0; i < num_trees; i++) {
      sub_trees.emplace_back(table.num_segments, i + seed);
      sub_trees.back().clear_all();
    }
    // shared_block_storage.cuh:251-274: 128 slots for class 0, halving, >= 4
    uint64_t n = 128;
    for (uint32_t i = 0; i < num_trees; i++) {
      slots.emplace_back(n, 0u);
      overflow.emplace_back(overflow_slots(n), 0u);
      n = n / 2;
      if (n < 4) n = 4;
    }
    for (uint32_t tree = 0; tree < num_trees; tree++) {
      for (size_t s = 0; s < slots[tree].size(); s++) {
        if (!boot_block(tree, (int)s)) return false;
      }
    }
    return true;
  }

  // The GPU gives a class with fewer than 64 pinned slots overflow slots up
  // to 64, used only when more than one lane of a wave claims at once; with
  // one lane they stay empty. The snapshot lists each class's pinned slots
  // followed by its overflow slots.
  static uint64_t overflow_slots(uint64_t pinned) { return pinned < 64 ? 64 - pinned : 0; }

  uint64_t total_slots() const {
    uint64_t t = 0;
    for (auto& s : slots) t += s.size() + overflow_slots(s.size());
    return t;
  }

  uint64_t snapshot_bytes() const {
    uint64_t words = 8;
    for (auto& l : segment_tree.layers) words += l.num_blocks;
    for (auto& st : sub_trees)
      for (auto& l : st.layers) words += l.num_blocks;
    uint64_t bytes = words * 8;
    bytes += ((table.num_segments * 2 + 7) / 8) * 8;
    uint64_t nblocks = table.num_segments * table.bps;
    bytes += nblocks * 8 + nblocks * 4 + table.num_segments * 16 + total_slots() * 8;
    return ((bytes + 7) / 8) * 8;
  }

  void snapshot(uint8_t* out) const {
    uint8_t* p = out;
    auto put = [&](const void* src, size_t n) {
      std::memcpy(p, src, n);
      p += n;
    };
    uint64_t header[8] = {table.num_segments, table.bps, num_trees, (uint64_t)segment_tree.num_layers,
                          total_slots(), locks, 0, 0};
    put(header, sizeof(header));
    for (auto& l : segment_tree.layers) put(l.bits.data(), l.bits.size() * 8);
    for (auto& st : sub_trees)
      for (auto& l : st.layers) put(l.bits.data(), l.bits.size() * 8);
    put(table.chunk_ids.data(), table.chunk_ids.size() * 2);
    size_t pad = ((table.num_segments * 2 + 7) / 8) * 8 - table.num_segments * 2;
    std::memset(p, 0, pad);
    p += pad;
    for (auto& b : table.blocks) {
      put(&b.malloc_counter, 4);
      put(&b.free_counter, 4);
    }
    put(table.queues.data(), table.queues.size() * 4);
    for (uint64_t s = 0; s < table.num_segments; s++) {
      put(&table.active[s], 4);
      put(&table.q_alloc[s], 4);
      put(&table.q_free[s], 4);
      put(&table.q_final[s], 4);
    }
    for (size_t c = 0; c < slots.size(); c++) {
      put(slots[c].data(), slots[c].size() * 4);
      if (!overflow[c].empty()) put(overflow[c].data(), overflow[c].size() * 4);
    }
    // the GPU's ahead word per slot: the GPU fetches blocks ahead only for
    // claims under load, which this model does not replay, so every word is 0
    std::memset(p, 0, total_slots() * 4);
    p += total_slots() * 4;
    size_t used = (size_t)(p - out);
    size_t total = snapshot_bytes();
    if (total > used) std::memset(p, 0, total - used);
  }
};

static bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

}  // namespace orc

struct orc_alloc {
  orc::Alloc a;
};
struct orc_veb {
  orc::Veb t;
};

extern "C" {

uint64_t orc_murmur64a(uint64_t key, uint64_t seed) { return orc::murmur64a(key, seed); }
int orc_first_bit_bigger(uint64_t x) { return orc::first_bit_bigger(x); }
int orc_cfcll(uint64_t bits) { return orc::cfcll(bits); }
uint64_t orc_splitmix64(uint64_t x) { return orc::splitmix64(x); }
int orc_reregister(uint32_t slot, uint64_t nb) { return orc::reregister_now(slot, nb); }

int orc_route(uint64_t seg, uint64_t min, uint64_t max, uint64_t size, uint32_t* path, uint32_t* tree,
              uint64_t* count) {
  using namespace orc;
  uint32_t num_trees = (uint32_t)(first_bit_bigger(max) - first_bit_bigger(min) + 1);
  int smallest_bits = first_bit_bigger(min);
  uint32_t tree_id = size < min ? 0u : (uint16_t)(first_bit_bigger(size) - smallest_bits);
  *count = 1;
  if (tree_id < num_trees) {
    *path = 0;
    *tree = tree_id;
    return GAL_OK;
  }
  int block_tree = first_bit_bigger(size) - first_bit_bigger(min * 4096);
  if (block_tree < 0) {
    *path = 0;
    *count = 1ull << (tree_id - (num_trees - 1));
    *tree = num_trees - 1;
  } else if ((uint32_t)block_tree < num_trees) {
    *path = 1;
    *tree = (uint32_t)block_tree;
  } else {
    *path = 2;
    *tree = 0;
    *count = (size - 1) / seg + 1;
  }
  return GAL_OK;
}

orc_veb* orc_veb_create(uint64_t universe, uint64_t seed, int start_full) {
  if (universe == 0) return nullptr;
  orc_veb* t = new (std::nothrow) orc_veb{orc::Veb(universe, seed)};
  if (t && !start_full) t->t.clear_all();
  return t;
}
void orc_veb_destroy(orc_veb* t) { delete t; }
int orc_veb_layers(orc_veb* t) { return t->t.num_layers; }
uint64_t orc_veb_layer_words(orc_veb* t, int layer) { return t->t.layers[layer].num_blocks; }
void orc_veb_get_words(orc_veb* t, int layer, uint64_t* out) {
  std::memcpy(out, t->t.layers[layer].bits.data(), t->t.layers[layer].num_blocks * 8);
}
void orc_veb_set_words(orc_veb* t, int layer, const uint64_t* in) {
  std::memcpy(t->t.layers[layer].bits.data(), in, t->t.layers[layer].num_blocks * 8);
}
int orc_veb_insert(orc_veb* t, uint64_t v) { return t->t.insert_force_update(v); }
int orc_veb_remove(orc_veb* t, uint64_t v) { return t->t.remove(v); }
int orc_veb_query(orc_veb* t, uint64_t v) { return t->t.query(v); }
uint64_t orc_veb_successor(orc_veb* t, uint64_t v) { return t->t.successor_thorough(v); }
uint64_t orc_veb_lock_offset(orc_veb* t, uint64_t start) { return t->t.lock_offset(start); }
uint64_t orc_veb_malloc_first(orc_veb* t) { return t->t.malloc_first(); }
uint64_t orc_veb_find_random(orc_veb* t, uint64_t tid) { return t->t.find_random_valid_index(tid); }
uint64_t orc_veb_gather_multiple(orc_veb* t, uint64_t n) { return t->t.gather_multiple(n); }
void orc_veb_return_multiple(orc_veb* t, uint64_t start, uint64_t n) { t->t.return_multiple(start, n); }

int orc_create(uint64_t segment_size, uint64_t min_size, uint64_t max_size, uint64_t pool_bytes, uint64_t seed,
               orc_alloc** out) {
  using namespace orc;
  if (!out) return GAL_ERR_INVALID;
  *out = nullptr;
  if (!is_pow2(segment_size) || !is_pow2(min_size) || !is_pow2(max_size) || min_size > max_size)
    return GAL_ERR_INVALID;
  if (segment_size < max_size * 4096) return GAL_ERR_INVALID;  // static_assert, gallatin.cuh:300
  uint64_t nseg = pool_bytes / segment_size;                   // get_max_chunks, alloc_utils.cuh:223-226
  if (nseg == 0) return GAL_ERR_INVALID;
  orc_alloc* h = new (std::nothrow) orc_alloc{orc::Alloc(segment_size, min_size, max_size, nseg, seed)};
  if (!h) return GAL_ERR_OOM;
  h->a.pool_bytes = pool_bytes;
  if (!h->a.boot()) {
    delete h;
    return GAL_ERR_OOM;
  }
  *out = h;
  return GAL_OK;
}

void orc_destroy(orc_alloc* a) { delete a; }

int orc_get_info(orc_alloc* h, gal_info_t* info) {
  if (!h || !info) return GAL_ERR_INVALID;
  std::memset(info, 0, sizeof(*info));
  info->segment_size = h->a.SEG;
  info->min_size = h->a.MIN;
  info->max_size = h->a.MAX;
  info->num_segments = h->a.table.num_segments;
  info->blocks_per_segment = h->a.table.bps;
  info->num_trees = h->a.num_trees;
  info->veb_layers = (uint64_t)h->a.segment_tree.num_layers;
  info->total_slots = h->a.total_slots();
  info->snapshot_bytes = h->a.snapshot_bytes();
  return GAL_OK;
}

uint64_t orc_malloc(orc_alloc* h, uint64_t size, uint64_t tid, uint32_t slot) { return h->a.malloc(size, tid, slot); }
int orc_free(orc_alloc* h, uint64_t byte_offset) { return h->a.free(byte_offset); }
uint64_t orc_faults(orc_alloc* h) { return h->a.faults; }

int orc_replay(orc_alloc* h, const gal_op* ops, uint64_t n, uint64_t* results) {
  if (!h || (n && (!ops || !results))) return GAL_ERR_INVALID;
  for (uint64_t i = 0; i < n; i++) {
    const gal_op& op = ops[i];
    if (op.kind == GAL_OP_MALLOC) {
      results[i] = h->a.malloc(op.arg, op.tid, op.slot);
    } else if (op.kind == GAL_OP_FREE) {
      if (op.arg >= i || ops[op.arg].kind != GAL_OP_MALLOC || results[op.arg] == GAL_NULL_OFFSET)
        return GAL_ERR_INVALID;
      int rc = h->a.free(results[op.arg]);
      if (rc != GAL_OK) return rc;
      results[i] = 0;
    } else {
      return GAL_ERR_INVALID;
    }
  }
  return GAL_OK;
}

int orc_group_replay(orc_alloc* h, const gal_group_op* ops, uint64_t n, uint64_t* results) {
  if (!h || (n && (!ops || !results))) return GAL_ERR_INVALID;
  for (uint64_t k = 0; k < n * 64; k++) results[k] = GAL_NULL_OFFSET;
  for (uint64_t i = 0; i < n; i++) {
    const gal_group_op& op = ops[i];
    if (op.lanes == 0) return GAL_ERR_INVALID;
    uint64_t* out = results + i * 64;
    if (op.kind == GAL_OP_MALLOC || op.kind == GAL_OP_MALLOC_SPREAD) {
      int rc = h->a.group_malloc(op.lanes, op.arg, op.tid, op.slot, op.kind == GAL_OP_MALLOC_SPREAD, out);
      if (rc != GAL_OK) return rc;
      continue;
    }
    if (op.kind != GAL_OP_FREE) return GAL_ERR_INVALID;
    uint64_t offsets[64] = {};
    for (uint64_t m = op.lanes; m; m &= m - 1) {
      const int l = __builtin_ctzll(m);
      const uint64_t j = op.arg[l];
      const uint32_t src_kind = ops[j / 64].kind;
      if (j >= i * 64 || !((ops[j / 64].lanes >> (j % 64)) & 1ull) ||
          (src_kind != GAL_OP_MALLOC && src_kind != GAL_OP_MALLOC_SPREAD) || results[j] == GAL_NULL_OFFSET)
        return GAL_ERR_INVALID;
      offsets[l] = results[j];
    }
    int rc = (op.lanes & (op.lanes - 1)) == 0 ? h->a.free(offsets[__builtin_ctzll(op.lanes)])
                                               : h->a.group_free(op.lanes, offsets);
    if (rc != GAL_OK) return rc;
    for (uint64_t m = op.lanes; m; m &= m - 1) out[__builtin_ctzll(m)] = 0;
  }
  return GAL_OK;
}

int orc_group_malloc(orc_alloc* h, uint64_t lanes, const uint64_t* sizes, const uint64_t* tids, uint32_t slot,
                     int spread, uint64_t* out) {
  if (!h || !lanes || !sizes || !tids || !out) return GAL_ERR_INVALID;
  return h->a.group_malloc(lanes, sizes, tids, slot, spread != 0, out);
}

int orc_group_free(orc_alloc* h, uint64_t lanes, const uint64_t* byte_offsets) {
  if (!h || !lanes || !byte_offsets) return GAL_ERR_INVALID;
  if ((lanes & (lanes - 1)) == 0) return h->a.free(byte_offsets[__builtin_ctzll(lanes)]);
  return h->a.group_free(lanes, byte_offsets);
}

int orc_snapshot(orc_alloc* h, void* buf, uint64_t cap, uint64_t* needed) {
  if (!h) return GAL_ERR_INVALID;
  uint64_t need = h->a.snapshot_bytes();
  if (needed) *needed = need;
  if (!buf || cap < need) return GAL_ERR_BUFFER;
  h->a.snapshot((uint8_t*)buf);
  return GAL_OK;
}

}  // extern "C"
