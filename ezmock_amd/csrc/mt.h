// mt.h -- MT19937 streams on the device (jump-ahead + segment generation).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MT_JUMP_WORDS 624       // coefficient words of a jump polynomial

// Window (x_0 .. x_623) of the raw recurrence after init_genrand(seed) when
// seed < 2^32, else init_by_array({seed low, seed high}) (the published
// MT19937 initialisations, mt19937ar.c).  Host and device.
__host__ __device__ inline void mt_seed_window(uint64_t seed, uint32_t *w) {
  const int n = 624;
  const uint32_t s0 = (seed <= 0xffffffffull) ? (uint32_t) seed : 19650218u;
  w[0] = s0;
  for (int i = 1; i < n; i++) w[i] = 1812433253u * (w[i - 1] ^ (w[i - 1] >> 30)) + (uint32_t) i;
  if (seed <= 0xffffffffull) return;
  const uint32_t key[2] = {(uint32_t) (seed & 0xffffffffull), (uint32_t) (seed >> 32)};
  int i = 1, j = 0;
  for (int k = n; k; k--) {
    w[i] = (w[i] ^ ((w[i - 1] ^ (w[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t) j;
    i++; j++;
    if (i >= n) { w[0] = w[n - 1]; i = 1; }
    if (j >= 2) j = 0;
  }
  for (int k = n - 1; k; k--) {
    w[i] = (w[i] ^ ((w[i - 1] ^ (w[i - 1] >> 30)) * 1566083941u)) - (uint32_t) i;
    i++;
    if (i >= n) { w[0] = w[n - 1]; i = 1; }
  }
  w[0] = 0x80000000u;
}

// Coefficients of x^e mod phi (bit i of word i/32 = coefficient of x^i),
// cached; nullptr if the characteristic polynomial failed its check.
const uint32_t *mt_jump_poly(uint64_t e);

// Words first .. first + n - 1 of the tempered output stream of `seed`, in
// two steps, so that the segment windows of a stream (the jump trees:
// latency-bound, seed-dependent only) can be built ahead of time, on another
// stream, and reused: nseg windows at words first + s J (win and the tree's
// scratch scr: mt_window_words(nseg) u32 each, owned by the caller, so that
// builds on several streams share no stream-ordered pool memory), then the
// words [first, first + n) into out (device, n u32) from windows built for
// at least mt_segments(n) segments.  Both return a HIP error code.
size_t mt_segments(uint64_t n);
size_t mt_window_words(size_t nseg);
hipError_t mt_stream_windows(uint64_t seed, uint64_t first, size_t nseg, uint32_t *win,
    uint32_t *scr, hipStream_t st);
hipError_t mt_stream_fill(const uint32_t *win, uint64_t n, uint32_t *out, hipStream_t st);
