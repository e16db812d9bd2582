// mt.h -- MT19937 streams on the device (jump-ahead + segment generation).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MT_JUMP_WORDS 624       // coefficient words of a jump polynomial

// Window (x_0 .. x_623) of the raw recurrence after init_genrand(seed) when
// seed < 2^32, else init_by_array({seed low, seed high}).
void mt_seed_window(uint64_t seed, uint32_t *w);

// Coefficients of x^e mod phi (bit i of word i/32 = coefficient of x^i),
// cached; nullptr if the characteristic polynomial failed its check.
const uint32_t *mt_jump_poly(uint64_t e);

// Words first .. first + n - 1 of the tempered output stream of `seed`
// into out (device, n u32), on stream st.  Returns a HIP error code.
hipError_t mt_stream_words(uint64_t seed, uint64_t first, uint64_t n,
    uint32_t *out, hipStream_t st);
