/*
 * orc_rng.c -- random streams of the CPU oracle (test infrastructure only).
 *
 * The reference draws every random number through the `prand` library
 * (submodule libEZmock/prand, .gitmodules:1-3), which is not present.  Its
 * call sites fix the interface:
 *   prand_init(type, seed, nstream, 0, &err)        EZmock.c:85
 *   rng->reset(state, seed, step, &err)             perturb.c:113-118,
 *                                                   pop_tracer.c:236,781,919,
 *                                                   1071
 *   rng->get_double(state), get_double_pos(state)   perturb.c:155-163,
 *                                                   pop_tracer.c:73-74
 * `reset` reseeds and skips `step` draws, so a stream position is a pure
 * function of (seed, step).  The generator cores are the published ones:
 *   MRG32k3a: P. L'Ecuyer, Operations Research 47 (1999) 159 -- recurrence,
 *             moduli and multipliers; output (p1-p2) mod m1 scaled by
 *             1/(m1+1) as in L'Ecuyer's RngStreams package.
 *   MT19937:  M. Matsumoto & T. Nishimura, ACM TOMACS 8 (1998) 3 --
 *             init_genrand / init_by_array / genrand_int32 / genrand_res53.
 * ASSUMED (parity unpinned against prand): the 64-bit seed -> state map
 * (SplitMix64 expansion for MRG32k3a; init_genrand or init_by_array for
 * MT19937), get_double / get_double_pos conversions, and the unit of `step`
 * (one call of get_double = one MRG32k3a step = two MT19937 outputs).
 */
#include "orc.h"
#include <string.h>

#define MRG_M1 4294967087ULL
#define MRG_M2 4294944443ULL
#define MRG_A12 1403580ULL
#define MRG_A13N 810728ULL
#define MRG_A21 527612ULL
#define MRG_A23N 1370589ULL
#define MRG_NORM 2.328306549295727688e-10

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* One-step transition matrices (state vector (x[n-3], x[n-2], x[n-1])). */
static void mrg_base(int comp, uint64_t A[9]) {
  memset(A, 0, 9 * sizeof(uint64_t));
  A[1] = 1; A[5] = 1;
  if (comp == 0) { A[6] = MRG_M1 - MRG_A13N; A[7] = MRG_A12; A[8] = 0; }
  else { A[6] = MRG_M2 - MRG_A23N; A[7] = 0; A[8] = MRG_A21; }
}

static void mat_mul(const uint64_t *A, const uint64_t *B, uint64_t m,
    uint64_t *C) {
  uint64_t T[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      uint64_t s = 0;
      for (int k = 0; k < 3; k++) s = (s + (A[3 * i + k] * B[3 * k + j]) % m) % m;
      T[3 * i + j] = s;
    }
  memcpy(C, T, sizeof(T));
}

static void mat_vec(const uint64_t *A, uint64_t *v, uint64_t m) {
  uint64_t t[3];
  for (int i = 0; i < 3; i++) {
    uint64_t s = 0;
    for (int k = 0; k < 3; k++) s = (s + (A[3 * i + k] * v[k]) % m) % m;
    t[i] = s;
  }
  memcpy(v, t, sizeof(t));
}

/* A^(2^log2n) for component comp, by repeated squaring. */
void orc_mrg_matpow(int comp, int log2n, uint64_t out[9]) {
  uint64_t m = comp ? MRG_M2 : MRG_M1;
  mrg_base(comp, out);
  for (int i = 0; i < log2n; i++) mat_mul(out, out, m, out);
}

static uint64_t jump_tab[2][64][9];
static int jump_ready = 0;

static void jump_init(void) {
  if (jump_ready) return;
#pragma omp critical(orc_jump_init)
  {
    if (!jump_ready) {
      for (int c = 0; c < 2; c++) {
        uint64_t m = c ? MRG_M2 : MRG_M1;
        mrg_base(c, jump_tab[c][0]);
        for (int b = 1; b < 64; b++)
          mat_mul(jump_tab[c][b - 1], jump_tab[c][b - 1], m, jump_tab[c][b]);
      }
      jump_ready = 1;
    }
  }
}

void orc_mrg_jump(orc_rng_t *r, uint64_t step) {
  jump_init();
  for (int b = 0; b < 64 && step; b++, step >>= 1) {
    if (step & 1) {
      mat_vec(jump_tab[0][b], r->s, MRG_M1);
      mat_vec(jump_tab[1][b], r->s + 3, MRG_M2);
    }
  }
}

void orc_mrg_set_state(orc_rng_t *r, const uint64_t s[6]) {
  r->type = ORC_RNG_MRG32K3A;
  memcpy(r->s, s, 6 * sizeof(uint64_t));
}

void orc_mrg_get_state(const orc_rng_t *r, uint64_t s[6]) {
  memcpy(s, r->s, 6 * sizeof(uint64_t));
}

static inline double mrg_next(orc_rng_t *r) {
  int64_t p1 = (int64_t) (MRG_A12 * r->s[1]) - (int64_t) (MRG_A13N * r->s[0]);
  p1 %= (int64_t) MRG_M1;
  if (p1 < 0) p1 += MRG_M1;
  r->s[0] = r->s[1]; r->s[1] = r->s[2]; r->s[2] = (uint64_t) p1;
  int64_t p2 = (int64_t) (MRG_A21 * r->s[5]) - (int64_t) (MRG_A23N * r->s[3]);
  p2 %= (int64_t) MRG_M2;
  if (p2 < 0) p2 += MRG_M2;
  r->s[3] = r->s[4]; r->s[4] = r->s[5]; r->s[5] = (uint64_t) p2;
  return (p1 > p2) ? (double) (p1 - p2) * MRG_NORM
                   : (double) (p1 - p2 + (int64_t) MRG_M1) * MRG_NORM;
}

/* ------------------------------------------------------------ MT19937 --- */
#define MT_N 624
#define MT_M 397

static void mt_init_genrand(orc_rng_t *r, uint32_t s) {
  r->mt[0] = s;
  for (int i = 1; i < MT_N; i++)
    r->mt[i] = 1812433253U * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t) i;
  r->mti = MT_N;
}

static void mt_init_by_array(orc_rng_t *r, const uint32_t *key, int len) {
  mt_init_genrand(r, 19650218U);
  int i = 1, j = 0;
  int k = (MT_N > len) ? MT_N : len;
  for (; k; k--) {
    r->mt[i] = (r->mt[i] ^ ((r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) * 1664525U))
        + key[j] + (uint32_t) j;
    i++; j++;
    if (i >= MT_N) { r->mt[0] = r->mt[MT_N - 1]; i = 1; }
    if (j >= len) j = 0;
  }
  for (k = MT_N - 1; k; k--) {
    r->mt[i] = (r->mt[i] ^ ((r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) * 1566083941U))
        - (uint32_t) i;
    i++;
    if (i >= MT_N) { r->mt[0] = r->mt[MT_N - 1]; i = 1; }
  }
  r->mt[0] = 0x80000000U;
  r->mti = MT_N;
}

uint32_t orc_mt_next(orc_rng_t *r) {
  static const uint32_t mag01[2] = {0U, 0x9908b0dfU};
  uint32_t y;
  if (r->mti >= MT_N) {
    int kk;
    for (kk = 0; kk < MT_N - MT_M; kk++) {
      y = (r->mt[kk] & 0x80000000U) | (r->mt[kk + 1] & 0x7fffffffU);
      r->mt[kk] = r->mt[kk + MT_M] ^ (y >> 1) ^ mag01[y & 1U];
    }
    for (; kk < MT_N - 1; kk++) {
      y = (r->mt[kk] & 0x80000000U) | (r->mt[kk + 1] & 0x7fffffffU);
      r->mt[kk] = r->mt[kk + (MT_M - MT_N)] ^ (y >> 1) ^ mag01[y & 1U];
    }
    y = (r->mt[MT_N - 1] & 0x80000000U) | (r->mt[0] & 0x7fffffffU);
    r->mt[MT_N - 1] = r->mt[MT_M - 1] ^ (y >> 1) ^ mag01[y & 1U];
    r->mti = 0;
  }
  y = r->mt[r->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680U;
  y ^= (y << 15) & 0xefc60000U;
  y ^= (y >> 18);
  return y;
}

/* ---------------------------------------------------------- interface --- */
int orc_rng_reset(orc_rng_t *r, int type, uint64_t seed, uint64_t step) {
  r->type = type;
  if (type == ORC_RNG_MRG32K3A) {
    uint64_t x = seed;
    for (int i = 0; i < 3; i++) r->s[i] = splitmix64(&x) % MRG_M1;
    for (int i = 3; i < 6; i++) r->s[i] = splitmix64(&x) % MRG_M2;
    if (!r->s[0] && !r->s[1] && !r->s[2]) r->s[0] = 1;
    if (!r->s[3] && !r->s[4] && !r->s[5]) r->s[3] = 1;
    orc_mrg_jump(r, step);
    return 0;
  }
  else if (type == ORC_RNG_MT19937) {
    if (seed <= 0xffffffffULL) mt_init_genrand(r, (uint32_t) seed);
    else {
      uint32_t key[2] = {(uint32_t) (seed & 0xffffffffULL),
                         (uint32_t) (seed >> 32)};
      mt_init_by_array(r, key, 2);
    }
    for (uint64_t i = 0; i < 2 * step; i++) (void) orc_mt_next(r);
    return 0;
  }
  return 1;
}

double orc_rng_double(orc_rng_t *r) {
  if (r->type == ORC_RNG_MRG32K3A) return mrg_next(r);
  uint32_t a = orc_mt_next(r) >> 5, b = orc_mt_next(r) >> 6;
  return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

double orc_rng_double_pos(orc_rng_t *r) {
  if (r->type == ORC_RNG_MRG32K3A) return mrg_next(r);
  uint32_t a = orc_mt_next(r) >> 5, b = orc_mt_next(r) >> 6;
  return (a * 67108864.0 + b + 0.5) * (1.0 / 9007199254740992.0);
}
