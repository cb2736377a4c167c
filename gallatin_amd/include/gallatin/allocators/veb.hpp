// veb.hpp — 64-ary bit-vector tree over segment indices (device side).
//
// Same structure and update rules as the reference's veb_tree
// (include/gallatin/allocators/veb.cuh:109-1228): layer L-1 holds one bit per
// item, each layer above holds one bit per 64-bit word of the layer below,
// layer 0 is the top summary. Words are only touched with agent-scope atomics.
// The search is word-at-a-time: a successor query is at most L dependent
// 64-bit loads going up and L going down, each resolved with one ffs.
#pragma once

#include "device_utils.hpp"

namespace gallatin {
namespace allocators {

constexpr int kMaxVebLayers = 6;           // universes below 2^36 items
constexpr int kVebHashedAttempts = 15;     // reference VEB_MAX_ATTEMPTS, veb.cuh:59
constexpr int kSuccessorRestarts = 1 << 16;  // safety bound on stale-summary restarts
constexpr int kClaimTries = 256;             // lost races inside one word before moving past it

struct veb_tree {
  uint64_t* bits[kMaxVebLayers];    // [0] top summary ... [num_layers-1] items
  uint64_t universe[kMaxVebLayers]; // items represented by each layer
  uint64_t words[kMaxVebLayers];    // 64-bit words of each layer
  uint64_t seed;                    // seeds the hashed start of find_random_valid_index
  uint64_t total_universe;
  int num_layers;

  // layers = floor((h-1)/6)+1 for h = bit length of the universe
  // (reference veb.cuh:601-605)
  __host__ __device__ static int layers_for(uint64_t universe) {
    int h = 64 - utils::clz64(universe);
    return (h - 1) / 6 + 1;
  }

#if defined(__HIPCC__)
  __device__ uint64_t load_word(int layer, uint64_t w) const { return atom::load(&bits[layer][w]); }

  // first set bit strictly above `low` in word `high` of `layer` (-1: any);
  // `seen` (if given) receives the word as loaded
  __device__ int find_next(int layer, uint64_t high, int low, uint64_t* seen = nullptr) const {
    if (high >= universe[layer]) return -1;
    uint64_t w = load_word(layer, high);
    if (seen) *seen = w;
    if (low >= 0) w &= ~utils::bitmask((uint64_t)low + 1);
    return utils::ffs64(w) - 1;
  }

  __device__ bool query(uint64_t v) const { return (load_word(num_layers - 1, v >> 6) >> (v & 63)) & 1ull; }

  // Clear item v; clear summary bits upward while a word empties. After
  // clearing a summary bit the word below is read again: an insert into it
  // between the emptying and the clear saw an empty word and set the very
  // bit being cleared, and without a restore its item would stay set but
  // unreachable from the summaries (a free segment no search can find). With
  // one lane the word is still empty and nothing is restored.
  __device__ bool remove(uint64_t v) { return (take(v >> 6, (int)(v & 63)) >> (v & 63)) & 1ull; }

  // remove() of bit `low` of item word `high`; returns the word as it was
  // before the clear (the bit was this caller's when it is set there)
  __device__ uint64_t take(uint64_t high, int low) {
    int layer = num_layers - 1;
    uint64_t old = atom::fetch_and(&bits[layer][high], ~(1ull << low));
    const uint64_t before = old;
    if (!(old & (1ull << low))) return before;
    while (utils::popc64(old) == 1 && layer > 0) {
      const uint64_t child = high;
      const int child_layer = layer;
      layer--;
      low = (int)(high & 63);
      high >>= 6;
      old = atom::fetch_and(&bits[layer][high], ~(1ull << low));
      // The word below is read only once the clear is performed, and with
      // an atomic (read_current): read earlier, or served from a copy the
      // XCD's L2 kept, it can miss an insert whose summary update came
      // before the clear (the insert then saw the bit still set and
      // stopped), and the word would keep an item that no search can reach.
      atom::order();
      if (atom::read_current(&bits[child_layer][child]) != 0) {
        atom::fetch_or(&bits[layer][high], 1ull << low);
        break;
      }
    }
    return before;
  }

  // set item v; set summary bits upward while a word was empty
  __device__ bool insert_force_update(uint64_t v) {
    uint64_t high = v >> 6;
    int low = (int)(v & 63);
    int layer = num_layers - 1;
    uint64_t old = atom::fetch_or(&bits[layer][high], 1ull << low);
    if (old & (1ull << low)) return false;
    while (old == 0 && layer > 0) {
      layer--;
      low = (int)(high & 63);
      high >>= 6;
      old = atom::fetch_or(&bits[layer][high], 1ull << low);
    }
    return true;
  }

  // First item strictly above q reachable through the summaries. A summary
  // bit whose word below turns out empty is stepped over by floating up again.
  // `seen` (if given) receives the item word the result was found in, as
  // loaded.
  __device__ uint64_t successor_thorough(uint64_t q, uint64_t* seen = nullptr) const {
    uint64_t high = q >> 6;
    int low = (int)(q & 63);
    int layer = num_layers - 1;
    for (int restart = 0; restart < kSuccessorRestarts; restart++) {
      int f;
      while ((f = find_next(layer, high, low, seen)) == -1) {
        if (layer == 0) return utils::kFail;
        layer--;
        low = (int)(high & 63);
        high >>= 6;
      }
      while (layer != num_layers - 1) {
        layer++;
        high = (high << 6) + (uint64_t)f;
        f = find_next(layer, high, -1, seen);
        if (f == -1) break;
      }
      if (f != -1) return (high << 6) + (uint64_t)f;
      // stale summary bit at (layer-1): continue strictly after it
      if (layer == 0) return utils::kFail;
      layer--;
      low = (int)(high & 63);
      high >>= 6;
    }
    return utils::kFail;
  }

  // true when the tree holds no item as seen from the top summary. Exact for
  // trees changed only by insert_force_update/remove (their summaries track
  // the item layer); the segment tree's runs can leave stale summaries.
  __device__ bool looks_empty() const { return top_looks_empty(top_word()); }
  // the same test split in two, for a caller that overlaps the top word's
  // load with other loads
  __device__ uint64_t top_word() const { return load_word(0, 0); }
  __device__ bool top_looks_empty(uint64_t top) const { return top == 0 && (num_layers > 1 || words[0] == 1); }

  // nothing malloc_first could claim: no item 0 and no successor of 0
  __device__ bool nothing_to_claim() const { return !query(0) && successor_thorough(0) == utils::kFail; }

  // Highest item reachable through the summaries: descend along the highest
  // set bit of each layer; a summary bit whose word below is empty is
  // stepped over by backing up and continuing below it.
  __device__ uint64_t max_item() const {
    int layer = 0;
    uint64_t high = 0;
    int below = 64;  // only bits under `below` of word `high` are candidates
    for (int step = 0; step < kSuccessorRestarts; step++) {
      const uint64_t w = load_word(layer, high) & utils::bitmask((uint64_t)below);
      if (w == 0) {
        if (layer == 0) return utils::kFail;
        below = (int)(high & 63);
        high >>= 6;
        layer--;
        continue;
      }
      const uint64_t b = 63 - utils::clz64(w);
      if (layer == num_layers - 1) return (high << 6) + b;
      high = (high << 6) + b;
      layer++;
      below = 64;
    }
    return utils::kFail;
  }

  // find_random_valid_index of a two-layer tree (at most 4096 items: pools
  // up to 32 GiB at 16 MiB segments): the same attempts in the same order,
  // with the same result on an unchanged tree, but the top word is read once
  // and each item word at most once (a few kept in registers), and the
  // highest item's word and the first hashed start's word are read in one
  // round trip. Written out, the general search read the top word three
  // times and re-read the start's word for the successor step: five to
  // seven dependent loads per block request under load, now two or three.
  __device__ uint64_t find_random_valid_index_2(uint64_t tid) const {
    const uint64_t* items = bits[1];
    const uint64_t top = load_word(0, 0);
    if (top == 0) return utils::kFail;  // looks_empty
    const uint64_t hash1 = utils::murmur64a(tid, seed);
    const uint64_t hash2 = utils::murmur64a(tid, hash1);
    // the item words read last (index, value; ~0: none), newest first; only
    // constant indices, so that they stay in registers
    constexpr int kKept = 4;
    uint64_t kept_w[kKept], kept_v[kKept];
#pragma unroll
    for (int i = 0; i < kKept; i++) kept_w[i] = kept_v[i] = ~0ull;
    auto keep = [&](uint64_t w, uint64_t v) {
#pragma unroll
      for (int i = kKept - 1; i > 0; i--) {
        kept_w[i] = kept_w[i - 1];
        kept_v[i] = kept_v[i - 1];
      }
      kept_w[0] = w;
      kept_v[0] = v;
    };
    auto get = [&](uint64_t w) -> uint64_t {
      uint64_t v = 0;
      bool hit = false;
#pragma unroll
      for (int i = 0; i < kKept; i++)
        if (!hit && kept_w[i] == w) {
          v = kept_v[i];
          hit = true;
        }
      if (!hit) {
        v = atom::load(&items[w]);
        keep(w, v);
      }
      return v;
    };
    {
      const uint64_t wb = (uint64_t)(63 - utils::clz64(top));
      const uint64_t w0 = (hash1 % total_universe) >> 6;
      const uint64_t vb = atom::load(&items[wb]);
      const uint64_t v0 = w0 != wb ? atom::load(&items[w0]) : vb;
      keep(wb, vb);
      if (w0 != wb) keep(w0, v0);
    }
    // max_item: the highest summary bit whose word holds an item
    uint64_t top_item = utils::kFail;
    for (uint64_t sum = top; sum != 0;) {
      const uint64_t w = (uint64_t)(63 - utils::clz64(sum));
      const uint64_t v = get(w);
      if (v != 0) {
        top_item = w * 64 + (uint64_t)(63 - utils::clz64(v));
        break;
      }
      sum &= ~(1ull << w);  // stale summary bit: back up below it
    }
    if (top_item == utils::kFail) return utils::kFail;
    // successor_thorough: first item strictly above q reachable through the
    // summaries (a stale summary bit is stepped over)
    auto successor = [&](uint64_t q) -> uint64_t {
      const uint64_t w = q >> 6;
      const uint64_t above = get(w) & ~utils::bitmask((q & 63) + 1);
      if (above != 0) return w * 64 + (uint64_t)(utils::ffs64(above) - 1);
      for (uint64_t sum = top & ~utils::bitmask(w + 1); sum != 0; sum &= sum - 1) {
        const uint64_t w2 = (uint64_t)(utils::ffs64(sum) - 1);
        const uint64_t v = get(w2);
        if (v != 0) return w2 * 64 + (uint64_t)(utils::ffs64(v) - 1);
      }
      return utils::kFail;
    };
    for (int attempts = 0; attempts < kVebHashedAttempts; attempts++) {
      const uint64_t start = (hash1 + (uint64_t)attempts * hash2) % total_universe;
      if (start > top_item) continue;
      if ((get(start >> 6) >> (start & 63)) & 1ull) return start;  // query
      const uint64_t index = successor(start);
      if (index != utils::kFail) return index;
    }
    return top_item == 0 ? utils::kFail : successor(0);
  }

  // Hashed start positions (reference veb.cuh:946-977); tid is the caller's
  // global thread id. An attempt succeeds exactly when some item lies at or
  // above its start, so starts above the highest item are skipped without
  // touching memory; the attempts that remain are the reference's. With the
  // tree unchanged during the call the result is the reference's; a sparse
  // tree no longer costs a failed successor walk per hashed start.
  __device__ uint64_t find_random_valid_index(uint64_t tid) const {
    if (num_layers == 2) return find_random_valid_index_2(tid);
    if (looks_empty()) return utils::kFail;
    const uint64_t top = max_item();
    if (top == utils::kFail) return utils::kFail;
    uint64_t hash1 = utils::murmur64a(tid, seed);
    uint64_t hash2 = utils::murmur64a(tid, hash1);
    for (int attempts = 0; attempts < kVebHashedAttempts; attempts++) {
      uint64_t start = (hash1 + (uint64_t)attempts * hash2) % total_universe;
      if (start > top) continue;
      if (query(start)) return start;
      uint64_t index = successor_thorough(start);
      if (index != utils::kFail) return index;
    }
    return top == 0 ? utils::kFail : successor_thorough(0);
  }

  // Lowest free item, claimed (reference veb.cuh:918-937 lock_offset(0),
  // up to 15 times, veb.cuh:979-991): word 0 read directly (item 0 and the
  // items above it), then the words above through the summaries. A
  // candidate is taken with one atomic; a lost race goes on with the word
  // as that atomic returned it, not with a new search. Once a race is lost,
  // candidates are set bits picked by a per-wave salt instead of the lowest:
  // claimers that arrive together (a burst of segment attaches) spread over
  // the word rather than all trying its lowest bit, one winner per round
  // trip. A claimer that never loses a race, as a single lane, takes the
  // lowest free item: the reference's result.
  __device__ uint64_t malloc_first() {
    const int items = num_layers - 1;
    bool lost = false;
    uint64_t salt = 0;
    for (int attempts = 0; attempts < kVebHashedAttempts; attempts++) {
      uint64_t high = 0;
      uint64_t seen = load_word(items, 0);
      for (int step = 0; step < kSuccessorRestarts; step++) {
        uint64_t cand = seen;
        for (int k = 0; cand != 0 && k < kClaimTries; k++) {
          int b = utils::ffs64(cand) - 1;
          if (lost) {
            salt = utils::splitmix64(salt ? salt : wave::slot_salt());
            for (uint32_t n = (uint32_t)(salt >> 32) % (uint32_t)utils::popc64(cand); n > 0; n--) cand &= cand - 1;
            b = utils::ffs64(cand) - 1;
          }
          const uint64_t old = take(high, b);
          if ((old >> b) & 1ull) return high * 64 + (uint64_t)b;
          lost = true;
          cand = old;
        }
        const uint64_t v = successor_thorough(high * 64 + 63, &seen);
        if (v == utils::kFail) break;
        high = v >> 6;
      }
    }
    return utils::kFail;
  }

  // set `count` bits from bit 0 of word `w` upward (undo of a partial claim)
  __device__ void rollback(uint64_t w, uint64_t count) {
    uint64_t* b = bits[num_layers - 1];
    while (count > 64) {
      atom::fetch_or(&b[w], ~0ull);
      count -= 64;
      w++;
    }
    atom::fetch_or(&b[w], utils::bitmask(count));
  }

  // Where the gather rule (gather_bottom) would start its successful claim
  // of n items, reading the item words as they are now: the word of the
  // attempt that succeeds, or kFail. The calling lanes (`lanes`, all active,
  // the call uniform) load the words together, one word per lane per round
  // trip, from the top down; the rule itself runs on the loaded words as
  // uniform scalar code, with gather_bottom's control flow.
  __device__ uint64_t gather_plan(uint64_t n, uint64_t lanes) const {
    const uint64_t* b = bits[num_layers - 1];
    const uint32_t m = wave::popc(lanes);
    const uint32_t rank = wave::popc(lanes & wave::lanemask_lt());
    int64_t hi = -1;  // the loaded words are (hi - m, hi]
    uint64_t mine = 0;
    auto get = [&](uint64_t w) -> uint64_t {
      if ((int64_t)w > hi || (int64_t)w <= hi - (int64_t)m) {
        hi = (int64_t)w;
        const int64_t x = hi - (int64_t)rank;
        mine = x >= 0 ? atom::load(&b[x]) : 0ull;
      }
      const int holder = wave::first_lane(wave::ballot(rank == (uint32_t)(hi - (int64_t)w)) & lanes);
      return wave::readlane64(mine, holder);
    };
    uint64_t w = words[num_layers - 1] - 1;
    while (w != 0) {
      const uint64_t start = w;
      uint64_t need = n;
      const int run = utils::trailing_ones(get(w));
      if (run == 0) {
        w--;
        continue;
      }
      if ((uint64_t)run >= need) return start;
      need -= (uint64_t)run;
      bool ok = true;
      while (need > 64) {
        if (w == 0) return utils::kFail;
        w--;
        if (get(w) != ~0ull) {
          ok = false;
          break;
        }
        need -= 64;
      }
      if (!ok) continue;
      if (w == 0) return utils::kFail;
      w--;
      const uint64_t top = utils::bitmask(need) << (64 - need);
      if ((get(w) & top) == top) return start;
    }
    return utils::kFail;
  }

  // Claim n contiguous items in the item layer, searching words from the top
  // down: a run is the trailing ones of a word, extended downward through full
  // words and the high bits of the next lower word (reference veb.cuh:323-452).
  // A run may not extend below word 0 and never starts in word 0. `from`
  // starts the search at a lower word (one that gather_plan found); attempts
  // above it would have failed and left the words as they were. A claim
  // that finds part of a run gone (taken by a concurrent claim, or held for
  // a moment by one that is putting it back) restores what it took and
  // tries again, at most kClaimTries times in all (steps past words without
  // a run are not counted).
  __device__ uint64_t gather_bottom(uint64_t n, uint64_t from = ~0ull) {
    uint64_t* b = bits[num_layers - 1];
    uint64_t w = from < words[num_layers - 1] ? from : words[num_layers - 1] - 1;
    for (int lost = 0; w != 0 && lost < kClaimTries;) {
      uint64_t need = n;
      int run = utils::trailing_ones(atom::load(&b[w]));
      if (run == 0) {
        w--;
        continue;
      }
      if ((uint64_t)run >= need) {
        uint64_t mask = utils::bitmask(need) << (run - need);
        uint64_t got = atom::fetch_and(&b[w], ~mask);
        if ((uint64_t)utils::popc64(got & mask) == need) return w * 64 + (run - need);
        atom::fetch_or(&b[w], got & mask);
        lost++;
        continue;
      }
      uint64_t low_mask = utils::bitmask(run);
      uint64_t got = atom::fetch_and(&b[w], ~low_mask);
      if (utils::popc64(got & low_mask) != run) {
        atom::fetch_or(&b[w], got & low_mask);
        lost++;
        continue;
      }
      need -= run;
      bool ok = true;
      while (need > 64) {
        if (w == 0) {
          rollback(w, n - need);
          return utils::kFail;
        }
        w--;
        got = atom::exch(&b[w], 0ull);
        if (utils::popc64(got) != 64) {
          atom::fetch_or(&b[w], got);
          rollback(w + 1, n - need);
          ok = false;
          break;
        }
        need -= 64;
      }
      if (!ok) {
        lost++;
        continue;
      }
      if (w == 0) {
        rollback(w, n - need);
        return utils::kFail;
      }
      w--;
      uint64_t top_mask = utils::bitmask(need) << (64 - need);
      got = atom::fetch_and(&b[w], ~top_mask);
      if ((uint64_t)utils::popc64(got & top_mask) == need) return w * 64 + 64 - need;
      atom::fetch_or(&b[w], got & top_mask);
      rollback(w + 1, n - need);
      lost++;
    }
    return utils::kFail;
  }

  // claim a run of n items and clear the summary bits of emptied words
  __device__ uint64_t gather_multiple(uint64_t n, uint64_t from = ~0ull) {
    uint64_t found = gather_bottom(n, from);
    if (found == utils::kFail) return found;
    uint64_t start = found, count = n;
    for (int layer = num_layers - 2; layer >= 0; layer--) {
      uint64_t hi = (start + count - 1) / 64 + 1;
      uint64_t lo = start / 64;
      start = lo;
      count = hi - lo;
      for (uint64_t w = start; w < start + count; w++) {
        if (atom::load(&bits[layer + 1][w]) != 0) continue;
        atom::fetch_and(&bits[layer][w >> 6], ~(1ull << (w & 63)));
        // an insert into the word between the check and the clear set this
        // very bit (it saw an empty word): put it back (as remove does,
        // reading the word with an atomic once the clear is performed)
        atom::order();
        if (atom::read_current(&bits[layer + 1][w]) != 0) atom::fetch_or(&bits[layer][w >> 6], 1ull << (w & 63));
      }
    }
    return found;
  }

  // return a run of n items; a summary bit is set only when the word below
  // is completely full again (reference veb.cuh:1128-1154, 1198-1218)
  __device__ void return_multiple(uint64_t start, uint64_t n) {
    uint64_t* b = bits[num_layers - 1];
    uint64_t w = start / 64;
    uint64_t first = start % 64;
    uint64_t left = n;
    while (left + first > 64) {
      uint64_t k = 64 - first;
      atom::fetch_or(&b[w], utils::bitmask(k) << first);
      first = 0;
      w++;
      left -= k;
    }
    atom::fetch_or(&b[w], utils::bitmask(left) << first);
    uint64_t s = start, count = n;
    for (int layer = num_layers - 2; layer >= 0; layer--) {
      uint64_t hi = (s + count - 1) / 64 + 1;
      uint64_t lo = s / 64;
      s = lo;
      count = hi - lo;
      for (uint64_t x = s; x < s + count; x++) {
        if (atom::load(&bits[layer + 1][x]) == ~0ull) atom::fetch_or(&bits[layer][x >> 6], 1ull << (x & 63));
      }
    }
  }
#endif  // __HIPCC__
};

}  // namespace allocators
}  // namespace gallatin
