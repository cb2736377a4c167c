// workloads.hpp — the synthetic request streams the benchmark kernels use.
//
// Every random choice is a pure function of (seed, round, index), so a
// request can be regenerated anywhere: on the device, on the host, alone or
// inside a mix.
#pragma once

#include <stdint.h>

#include "allocators/device_utils.hpp"

namespace gallatin {
namespace workloads {

// salt 0: size draw of thread i in round r; salt 1: free coin of live slot i
// after round r; salt 2: payload tag of live slot i
__host__ __device__ inline uint64_t mixed_key(uint64_t seed, uint64_t round, uint64_t i, uint64_t salt) {
  return utils::splitmix64(seed + utils::splitmix64((round << 32) ^ i ^ (salt << 60)));
}

// log-uniform power-of-two sizes 16 B .. 4 KiB
__host__ __device__ inline uint64_t mixed_size(uint64_t seed, uint64_t round, uint64_t i) {
  return 16ull << (mixed_key(seed, round, i, 0) % 9);
}

__host__ __device__ inline bool mixed_coin(uint64_t seed, uint64_t round, uint64_t j) {
  return mixed_key(seed, round, j, 1) & 1ull;
}

// nonzero payload tag (zero marks free memory in the atomic check)
__host__ __device__ inline uint64_t payload_tag(uint64_t seed, uint64_t index) {
  return mixed_key(seed, 0, index, 2) | 1ull;
}

// ---- serving: the header a sequence writes into each of its pages.
// A page's 16-byte header (pages are at least 16 bytes): word 0 links to the
// sequence's previous page, word 1 holds content derived from (request id,
// page index). A sequence folds every content word it reads back into its
// checksum, so the checksum is a pure function of (id, pages) unless another
// sequence wrote into one of its pages. One word of payload written and read
// back per page, like the one tag per allocation of the other workloads.
__host__ __device__ inline uint64_t page_word(uint64_t id, uint64_t page) {
  return utils::splitmix64(utils::splitmix64(id) ^ ((page << 3) | 1));
}
__host__ __device__ inline uint64_t fold(uint64_t h, uint64_t x) { return utils::splitmix64(h ^ x); }

}  // namespace workloads
}  // namespace gallatin
