// rng.h -- MRG32k3a streams with O(1) jump-ahead, for device and host code.
//
// Generator core: P. L'Ecuyer, Operations Research 47 (1999) 159 (moduli,
// multipliers, output (p1 - p2) mod m1 scaled by 1/(m1 + 1)).  Jump-ahead:
// the state advances by the 3x3 transition matrices A1, A2 (mod m1, m2);
// skipping n draws multiplies by A^n = prod_d A^(b_d * 256^d) over the
// base-256 digits b_d of n, read from a table of 2 x 8 x 256 matrices
// (147 KB, built once per process on the host and kept in device memory).
// One draw of the reference's prand interface (get_double / get_double_pos)
// is one generator step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MRG_M1 4294967087u      // 2^32 - 209
#define MRG_M2 4294944443u      // 2^32 - 22853
#define MRG_A12 1403580u
#define MRG_A13N 810728u
#define MRG_A21 527612u
#define MRG_A23N 1370589u
#define MRG_NORM 2.328306549295727688e-10

struct MrgState {
  uint32_t s[6];                // x1[n-3], x1[n-2], x1[n-1], x2[n-3..n-1]
};

// x mod (2^32 - c) for x < 2^63, c <= 22853: two folds and one subtraction.
__host__ __device__ __forceinline__ uint32_t mrg_mod(uint64_t x, uint32_t c) {
  uint64_t t = (x >> 32) * c + (x & 0xffffffffull);
  t = (t >> 32) * c + (t & 0xffffffffull);
  const uint64_t m = 0x100000000ull - c;
  if (t >= m) t -= m;
  return (uint32_t) t;
}

// One step; returns the uniform draw in (0, 1).
__host__ __device__ __forceinline__ double mrg_next(MrgState &g) {
  uint64_t t1 = (uint64_t) MRG_A12 * g.s[1] + (uint64_t) MRG_A13N * (MRG_M1 - g.s[0]);
  uint32_t p1 = mrg_mod(t1, 209u);
  g.s[0] = g.s[1]; g.s[1] = g.s[2]; g.s[2] = p1;
  uint64_t t2 = (uint64_t) MRG_A21 * g.s[5] + (uint64_t) MRG_A23N * (MRG_M2 - g.s[3]);
  uint32_t p2 = mrg_mod(t2, 22853u);
  g.s[3] = g.s[4]; g.s[4] = g.s[5]; g.s[5] = p2;
  int64_t d = (int64_t) p1 - (int64_t) p2;
  if (d <= 0) d += MRG_M1;
  return (double) d * MRG_NORM;
}

// v <- A v (mod 2^32 - c), A a row-major 3x3 of residues.
__host__ __device__ __forceinline__ void mrg_matvec(const uint32_t *A,
    uint32_t *v, uint32_t c) {
  uint32_t r[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    uint64_t a = mrg_mod((uint64_t) A[3 * i] * v[0], c);
    uint64_t b = mrg_mod((uint64_t) A[3 * i + 1] * v[1], c);
    uint64_t d = mrg_mod((uint64_t) A[3 * i + 2] * v[2], c);
    r[i] = mrg_mod(a + b + d, c);
  }
  v[0] = r[0]; v[1] = r[1]; v[2] = r[2];
}

// Jump table layout: tab[((comp * 8 + digit) * 256 + b) * 9 + e].
#define MRG_JUMP_WORDS (2 * 8 * 256 * 9)

__host__ __device__ __forceinline__ void mrg_jump(MrgState &g, uint64_t n,
    const uint32_t *__restrict__ tab) {
  for (int d = 0; d < 8 && n; d++, n >>= 8) {
    const uint32_t b = (uint32_t) (n & 255u);
    if (b) {
      mrg_matvec(tab + ((0 * 8 + d) * 256 + b) * 9, g.s, 209u);
      mrg_matvec(tab + ((1 * 8 + d) * 256 + b) * 9, g.s + 3, 22853u);
    }
  }
}

// Host: seed -> initial state (SplitMix64 expansion of the 64-bit seed; the
// reference's prand seeding is not available, so this map is an assumption),
// and the jump table.
void mrg_seed_state(uint64_t seed, MrgState &g);
const uint32_t *mrg_host_jump_table();

// MT19937 draws from a buffer of tempered output words (two words per draw,
// genrand_res53): get_double in [0, 1), get_double_pos in (0, 1).
__host__ __device__ __forceinline__ double mt_draw(const uint32_t *w, uint64_t d) {
  const uint32_t a = w[2 * d] >> 5, b = w[2 * d + 1] >> 6;
  return ((double) a * 67108864.0 + (double) b) * (1.0 / 9007199254740992.0);
}
__host__ __device__ __forceinline__ double mt_draw_pos(const uint32_t *w, uint64_t d) {
  const uint32_t a = w[2 * d] >> 5, b = w[2 * d + 1] >> 6;
  return ((double) a * 67108864.0 + (double) b + 0.5) * (1.0 / 9007199254740992.0);
}
