// chaining_table.hpp — a chained slab hash table whose slabs come from a
// Gallatin allocator: the allocator's client workload (reference
// include/gallatin/data_structs/chaining_table.cuh:57-300, test
// tests/src/chaining_test.cu:80-155).
//
// Buckets hold the head of a chain of slabs. A slab is {next, keys[S],
// vals[S]}: 8 + 16*S bytes (120 B for S = 7, served from the 128-byte
// class). insert(key, val) hashes the key to a bucket (MurmurHash64A, seed),
// walks the chain and claims the first empty key slot by compare-and-swap;
// a full chain gets a new slab from the allocator, published by a CAS on
// the last `next` (the loser of a publish race frees its slab). No
// deduplication: inserting a key twice stores it twice, as in the
// reference. query(key, val) walks the chain.
//
// MI355X: the eight XCD L2s are not coherent with each other, so while
// inserts run, chain pointers and keys are read with agent-scope atomic
// loads, a new slab's empty keys (agent-scope stores) are performed before
// the slab is linked (a wait for completion, not a fence: nothing ordinary
// needs writing back), and a slab's key words are read together (one round
// trip per slab).
#pragma once

#include "../allocators/device_utils.hpp"

namespace gallatin {
namespace data_structs {

template <class Key, class Val, int S>
struct chaining_block {
  chaining_block* next;
  Key keys[S];
  Val vals[S];
};

template <class Key, class Val, int S, class Alloc>
struct chaining_table {
  static_assert(sizeof(Key) == 8 && sizeof(Val) == 8, "64-bit keys and values");
  using block_type = chaining_block<Key, Val, S>;

  block_type** heads;  // one chain per bucket
  uint64_t nbuckets;
  uint64_t seed;
  Key empty;   // the reference's defaultKey
  Alloc* alloc;

#if defined(__HIPCC__)
  __device__ uint64_t bucket_of(Key k) const { return utils::murmur64a((uint64_t)k, seed) % nbuckets; }

  // a fresh slab with every key empty, CAS-published at *where; false when
  // the allocator is out of memory (reference chaining_table.cuh:195-214)
  __device__ bool attach_block(block_type** where) {
    block_type* b = (block_type*)alloc->malloc(sizeof(block_type));
    if (!b) return false;
    atom::store((uint64_t*)&b->next, 0ull);
    for (int i = 0; i < S; i++) atom::store((uint64_t*)&b->keys[i], (uint64_t)empty);
    atom::order();  // the empty keys (agent-scope stores) are performed before the slab is
    if (atom::cas((uint64_t*)where, 0ull, (uint64_t)b) != 0ull) alloc->free(b);
    return true;
  }

  // reference chaining_table.cuh:235-270. false when a new slab was needed
  // and the allocator had none (or after kMaxSteps chain steps, which no
  // table this size reaches).
  static constexpr int kMaxSteps = 1 << 20;
  __device__ bool insert(Key key, Val val) {
    block_type** where = &heads[bucket_of(key)];
    block_type* b = (block_type*)atom::load((uint64_t*)where);
    for (int step = 0; step < kMaxSteps; step++) {
      if (!b) {
        if (!attach_block(where)) return false;
        b = (block_type*)atom::load((uint64_t*)where);
        continue;
      }
      uint64_t k[S];
      for (int i = 0; i < S; i++) k[i] = atom::load((uint64_t*)&b->keys[i]);
      for (int i = 0; i < S; i++) {
        if (k[i] == (uint64_t)empty &&
            atom::cas((uint64_t*)&b->keys[i], (uint64_t)empty, (uint64_t)key) == (uint64_t)empty) {
          atom::store((uint64_t*)&b->vals[i], (uint64_t)val);
          return true;
        }
      }
      where = &b->next;
      b = (block_type*)atom::load((uint64_t*)where);
    }
    return false;
  }

  // reference chaining_table.cuh:272-293 (run after the inserts have finished)
  __device__ bool query(Key key, Val& val) const {
    const block_type* b = heads[bucket_of(key)];
    for (int step = 0; b && step < kMaxSteps; step++) {
      for (int i = 0; i < S; i++)
        if (b->keys[i] == key) {
          val = b->vals[i];
          return true;
        }
      b = b->next;
    }
    return false;
  }

  __device__ uint64_t chain_length(uint64_t bucket) const {
    uint64_t n = 0;
    for (const block_type* b = heads[bucket]; b && n < (uint64_t)kMaxSteps; b = b->next) n++;
    return n;
  }

  // free the chain of one bucket (after the table's last use)
  __device__ uint64_t free_chain(uint64_t bucket) {
    uint64_t n = 0;
    block_type* b = heads[bucket];
    heads[bucket] = nullptr;
    while (b && n < (uint64_t)kMaxSteps) {
      block_type* next = b->next;
      alloc->free(b);
      b = next;
      n++;
    }
    return n;
  }
#endif
};

}  // namespace data_structs
}  // namespace gallatin
