/*
 * serving_check.c — expected checksums of the serving loop, in C, for large
 * samples (test infrastructure: tests and bench.py's serving verification).
 *
 * The same definition as tests/serving_ref.py (which pins this file in a
 * CPU test): a request of `pages` pages and id `id` folds, into h = id, the
 * content word of page s-1 after page s is allocated (s = 1 .. pages-1),
 * then of pages pages-1 .. 0 as they drain; the result is fold(h, pages),
 * with fold(h, x) = splitmix64(h ^ x) and word(s) = splitmix64(splitmix64(id)
 * ^ (s << 3 | 1)). The page size does not enter: only a page's 16-byte
 * header is written.
 */
#include <stdint.h>

static uint64_t sm64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static uint64_t fold_page(uint64_t h, uint64_t base, uint64_t s) { return sm64(h ^ sm64(base ^ ((s << 3) | 1))); }

void serving_checksums(const uint64_t* ids, const uint32_t* pages, const uint32_t* page_bytes, uint64_t n,
                       uint64_t* out) {
  for (uint64_t i = 0; i < n; i++) {
    (void)page_bytes;
    const uint64_t base = sm64(ids[i]);
    uint64_t h = ids[i];
    for (uint64_t s = 1; s < pages[i]; s++) h = fold_page(h, base, s - 1);
    for (uint64_t s = pages[i]; s-- > 0;) h = fold_page(h, base, s);
    out[i] = sm64(h ^ (uint64_t)pages[i]);
  }
}
