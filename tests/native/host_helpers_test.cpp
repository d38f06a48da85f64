// TEST INFRASTRUCTURE ONLY: checks the host helpers of the library
// (phraya_amd/csrc/phx_host.h) that classify and pack sequences before any
// upload: the word-at-a-time ACGT test against a per-byte table for every
// byte value in every position and for random words, and the 32-base packer
// against the per-base packing for every length 0..40.
#include <cstdio>
#include <cstring>
#include <random>

#include "phx_host.h"

using namespace phx::host;

int main() {
  std::mt19937 rng(1);
  const char* acgt = "ACGT";
  for (int pos = 0; pos < 8; ++pos)
    for (int c = 0; c < 256; ++c) {
      uint8_t b[8] = {'A', 'C', 'G', 'T', 'A', 'C', 'G', 'T'};
      b[pos] = (uint8_t)c;
      uint64_t x;
      std::memcpy(&x, b, 8);
      const bool want = c == 'A' || c == 'C' || c == 'G' || c == 'T';
      if (word_acgt(x) != want) {
        std::printf("word_acgt: byte %d = %d\n", pos, c);
        return 1;
      }
    }
  for (int it = 0; it < 2000000; ++it) {
    uint8_t b[8];
    bool want = true;
    for (int k = 0; k < 8; ++k) {
      const int c = rng() % 4 == 0 ? (int)(rng() % 256) : acgt[rng() % 4];
      b[k] = (uint8_t)c;
      want &= c == 'A' || c == 'C' || c == 'G' || c == 'T';
    }
    uint64_t x;
    std::memcpy(&x, b, 8);
    if (word_acgt(x) != want || is_acgt(b, 8) != want) {
      std::printf("word_acgt: random word %d\n", it);
      return 1;
    }
  }
  for (int it = 0; it < 200000; ++it) {
    uint8_t s[40];
    const size_t n = rng() % 41;
    for (size_t k = 0; k < n; ++k) s[k] = (uint8_t)acgt[rng() % 4];
    uint64_t w[3], v[3] = {0, 0, 0};
    pack_into(w, s, n, false);
    for (size_t k = 0; k < n; ++k) v[k >> 5] |= (uint64_t)acgt_table()[s[k]] << ((k & 31) * 2);
    for (size_t k = 0; k < words_for(n, false); ++k)
      if (w[k] != v[k]) {
        std::printf("pack_into: length %zu\n", n);
        return 1;
      }
  }
  std::puts("host helpers ok");
  return 0;
}
