// rng_host.cpp -- host side of the MRG32k3a streams: seeding and the
// base-256 jump-ahead table (see rng.h).
#include <string.h>
#include <mutex>
#include "rng.h"

// 64-bit seed -> 6 state words: SplitMix64 (S. Vigna) expansion, reduced
// modulo m1 (first component) and m2 (second); an all-zero component is
// replaced by (1, 0, 0).
void mrg_seed_state(uint64_t seed, MrgState &g) {
  uint64_t x = seed;
  for (int i = 0; i < 6; i++) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    g.s[i] = (uint32_t) (z % (i < 3 ? (uint64_t) MRG_M1 : (uint64_t) MRG_M2));
  }
  if (!g.s[0] && !g.s[1] && !g.s[2]) g.s[0] = 1;
  if (!g.s[3] && !g.s[4] && !g.s[5]) g.s[3] = 1;
}

static void matmul(const uint32_t *A, const uint32_t *B, uint64_t m,
    uint32_t *C) {
  uint32_t T[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      uint64_t s = 0;
      for (int k = 0; k < 3; k++)
        s = (s + ((uint64_t) A[3 * i + k] * B[3 * k + j]) % m) % m;
      T[3 * i + j] = (uint32_t) s;
    }
  memcpy(C, T, sizeof(T));
}

// tab[c][d][b] = A_c^(b * 256^d) mod m_c
const uint32_t *mrg_host_jump_table() {
  static uint32_t tab[MRG_JUMP_WORDS];
  static std::once_flag once;
  std::call_once(once, [] {
    for (int c = 0; c < 2; c++) {
      const uint64_t m = c ? MRG_M2 : MRG_M1;
      uint32_t P[9] = {0, 1, 0, 0, 0, 1, 0, 0, 0};     // A^(256^d)
      if (c == 0) { P[6] = MRG_M1 - MRG_A13N; P[7] = MRG_A12; P[8] = 0; }
      else { P[6] = MRG_M2 - MRG_A23N; P[7] = 0; P[8] = MRG_A21; }
      for (int d = 0; d < 8; d++) {
        uint32_t *row = tab + ((size_t) (c * 8 + d) * 256) * 9;
        const uint32_t I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        memcpy(row, I, sizeof(I));
        for (int b = 1; b < 256; b++) matmul(row + (b - 1) * 9, P, m, row + b * 9);
        uint32_t Q[9];
        matmul(row + 255 * 9, P, m, Q);                 // P^256
        memcpy(P, Q, sizeof(Q));
      }
    }
  });
  return tab;
}
