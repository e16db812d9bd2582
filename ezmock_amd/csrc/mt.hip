// mt.hip -- MT19937 output streams on gfx950.
//
// A stream segment of words [first, first + n) is produced in three steps:
//   1. the seed window is moved to word `first` with one jump (x^first mod
//      phi, from mt_host.cpp);
//   2. segment start windows every J words are built level by level: at a
//      level of stride 16^l J, window c of the coarser level gives windows
//      16 c + k by one jump x^(k 16^l J) each, all in parallel, so S segments
//      need log16(S) jumps in sequence;
//   3. one workgroup per segment runs the twist recurrence in LDS and writes
//      its J tempered words.
// A jump applies g(A) = sum_i g_i A^i to a window with a blocked Horner
// scheme: acc <- A^B acc + sum_{r<B} g_{qB+r} A^r W over blocks q of B = 224
// coefficients; A^r W is the input window shifted by r words, so each block
// is a word-wise XOR over the set coefficients, and A^B acc is one parallel
// generation of B words (B <= 227 keeps every new word a function of the
// old window only).
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.h"
#include "mt.h"

#define MT_N 624
#define MT_M 397
#define MT_TPB 256
#define MT_JTPB 640                    // jump workgroups: 4 x 156 XOR slots
#define MT_B 224
#define MT_SEG_WORDS (1ull << 18)
#define MT_FAN 16

__device__ __forceinline__ uint32_t mt_f(uint32_t x0, uint32_t x1, uint32_t xm) {
  const uint32_t y = (x0 & 0x80000000u) | (x1 & 0x7fffffffu);
  return xm ^ (y >> 1) ^ ((x1 & 1u) ? 0x9908b0dfu : 0u);
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t y) {
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

// S <- g(A) S for the window S[0, 624) (S has room for 624 + B words),
// with the coefficients gl in LDS; acc and nw are LDS scratch.
__device__ void mt_jump_lds(const uint32_t *gl, uint32_t *S, uint32_t *acc,
    uint32_t *nw) {
  const int tid = threadIdx.x;
  int qtop = -1;                         // highest block with a set coefficient
  for (int i = MT_JUMP_WORDS - 1; i >= 0; i--)
    if (gl[i]) { qtop = (i * 32) / MT_B; break; }
  // S[624 .. 624 + B - 2] = next words of the input window
  for (int j = tid; j < MT_B - 1; j += MT_JTPB)
    S[MT_N + j] = mt_f(S[j], S[j + 1], S[j + MT_M]);
  for (int i = tid; i < MT_N; i += MT_JTPB) acc[i] = 0;
  __syncthreads();
  for (int q = qtop; q >= 0; q--) {
    // acc <- A^B acc
    if (q != qtop) {
      for (int j = tid; j < MT_B; j += MT_JTPB) nw[j] = mt_f(acc[j], acc[j + 1], acc[j + MT_M]);
      __syncthreads();
      uint32_t keep[(MT_N + MT_JTPB - 1) / MT_JTPB];
#pragma unroll
      for (int u = 0; u < (MT_N + MT_JTPB - 1) / MT_JTPB; u++) {
        const int i = tid + u * MT_JTPB;
        if (i < MT_N) keep[u] = (i + MT_B < MT_N) ? acc[i + MT_B] : nw[i + MT_B - MT_N];
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < (MT_N + MT_JTPB - 1) / MT_JTPB; u++) {
        const int i = tid + u * MT_JTPB;
        if (i < MT_N) acc[i] = keep[u];
      }
    }
    // acc ^= sum_r g_{qB + r} A^r W (B = 224 = 7 coefficient words).
    // Thread t < 624 owns words 4 (t % 156) .. + 3 and coefficient words
    // 2 (t / 156), + 1: per coefficient word it reads the 35 window words its
    // four need (nine aligned 16-byte LDS reads) and applies all 32
    // coefficients with masks, independent operations instead of a serial
    // loop over the set bits; the four partial sums of a word meet in LDS
    // XOR atomics.
    if (tid < MT_N) {
      const int i0 = 4 * (tid % (MT_N / 4));
      const int part = tid / (MT_N / 4);
      uint32_t t0 = 0, t1 = 0, t2 = 0, t3 = 0;
      for (int wq = 2 * part; wq < 2 * part + 2 && wq < MT_B / 32; wq++) {
        const int gw = q * (MT_B / 32) + wq;
        const uint32_t bits = (gw < MT_JUMP_WORDS) ? gl[gw] : 0u;
        if (!bits) continue;            // the same for every thread
        const uint4 *src = (const uint4 *) (S + i0 + wq * 32);   // up to S[847]
        uint32_t w[36];
#pragma unroll
        for (int v = 0; v < 9; v++) {
          const uint4 x = src[v];
          w[4 * v] = x.x; w[4 * v + 1] = x.y; w[4 * v + 2] = x.z; w[4 * v + 3] = x.w;
        }
#pragma unroll
        for (int r = 0; r < 32; r++) {
          const uint32_t m = 0u - ((bits >> r) & 1u);
          t0 ^= w[r] & m;
          t1 ^= w[r + 1] & m;
          t2 ^= w[r + 2] & m;
          t3 ^= w[r + 3] & m;
        }
      }
      if (t0) atomicXor(&acc[i0], t0);
      if (t1) atomicXor(&acc[i0 + 1], t1);
      if (t2) atomicXor(&acc[i0 + 2], t2);
      if (t3) atomicXor(&acc[i0 + 3], t3);
    }
    __syncthreads();
  }
  for (int i = tid; i < MT_N; i += MT_JTPB) S[i] = (qtop >= 0) ? acc[i] : 0u;
  __syncthreads();
}

// One workgroup per output window o = c fan + k (o < nout):
// out[o] = g_k(A) in[c], g_k = polys[k] (k = 0: a copy of in[c]).
__global__ void __launch_bounds__(MT_JTPB)
k_mt_fan(const uint32_t *__restrict__ polys, const uint32_t *__restrict__ in,
    int fan, size_t nout, uint32_t *__restrict__ out) {
  __shared__ uint32_t gl[MT_JUMP_WORDS];
  __shared__ __align__(16) uint32_t S[MT_N + MT_B];
  __shared__ uint32_t acc[MT_N], nw[MT_B];
  const int tid = threadIdx.x;
  const size_t o = blockIdx.x;
  if (o >= nout) return;                 // uniform over the workgroup
  const size_t c = o / fan;
  const int k = (int) (o % fan);
  for (int i = tid; i < MT_N; i += MT_JTPB) S[i] = in[c * MT_N + i];
  if (k > 0)
    for (int i = tid; i < MT_JUMP_WORDS; i += MT_JTPB) gl[i] = polys[(size_t) k * MT_JUMP_WORDS + i];
  __syncthreads();
  if (k > 0) mt_jump_lds(gl, S, acc, nw);
  for (int i = tid; i < MT_N; i += MT_JTPB) out[o * MT_N + i] = S[i];
}

// The seed window, on the device (one thread), so that a stream needs no
// host synchronisation.
__global__ void __launch_bounds__(64) k_mt_seed(uint64_t seed, uint32_t *__restrict__ w) {
  __shared__ uint32_t sw[MT_N];
  if (threadIdx.x == 0) mt_seed_window(seed, sw);
  __syncthreads();
  for (int i = threadIdx.x; i < MT_N; i += 64) w[i] = sw[i];
}

// One workgroup per segment s: words [s J, min((s + 1) J, n)) of the stream
// whose window at word s J is starts[s].
__global__ void __launch_bounds__(MT_TPB)
k_mt_fill(const uint32_t *__restrict__ starts, uint64_t J, uint64_t n,
    uint32_t *__restrict__ out) {
  __shared__ uint32_t a[MT_N], b[MT_N];
  const int tid = threadIdx.x;
  const uint64_t s = blockIdx.x;
  for (int i = tid; i < MT_N; i += MT_TPB) a[i] = starts[s * MT_N + i];
  __syncthreads();
  const uint64_t w0 = s * J;
  const uint64_t w1 = (w0 + J < n) ? w0 + J : n;
  uint32_t *cur = a, *nxt = b;
  for (uint64_t w = w0; w < w1; w += MT_N) {
    // new[j] = f(x_j, x_{j+1}, x_{j+397}) in three dependent phases
    for (int j = tid; j < MT_N - MT_M; j += MT_TPB)
      nxt[j] = mt_f(cur[j], cur[j + 1], cur[j + MT_M]);
    __syncthreads();
    for (int j = MT_N - MT_M + tid; j < 2 * (MT_N - MT_M); j += MT_TPB)
      nxt[j] = mt_f(cur[j], cur[j + 1], nxt[j + MT_M - MT_N]);
    __syncthreads();
    for (int j = 2 * (MT_N - MT_M) + tid; j < MT_N; j += MT_TPB)
      nxt[j] = mt_f(cur[j], (j + 1 < MT_N) ? cur[j + 1] : nxt[0], nxt[j + MT_M - MT_N]);
    __syncthreads();
    for (int j = tid; j < MT_N; j += MT_TPB)
      if (w + j < w1) out[w + j] = mt_temper(nxt[j]);
    uint32_t *t = cur; cur = nxt; nxt = t;
    __syncthreads();
  }
}

namespace {
std::mutex g_mu;
// (device, stride, fan) -> fan polynomials x^(k stride) mod phi, k < fan
std::map<std::tuple<int, uint64_t, int>, uint32_t *> g_dev_polys;

const uint32_t *device_fan_polys(uint64_t stride, int fan) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const auto key = std::make_tuple(dev, stride, fan);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_dev_polys.find(key);
    if (it != g_dev_polys.end()) return it->second;
  }
  std::vector<uint32_t> h((size_t) fan * MT_JUMP_WORDS, 0);
  for (int k = 1; k < fan; k++) {
    const uint32_t *g = mt_jump_poly((uint64_t) k * stride);
    if (!g) return nullptr;
    std::copy(g, g + MT_JUMP_WORDS, h.begin() + (size_t) k * MT_JUMP_WORDS);
  }
  uint32_t *d = nullptr;
  if (hipMalloc(&d, h.size() * sizeof(uint32_t)) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
    (void) hipFree(d);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  auto res = g_dev_polys.emplace(key, d);
  if (!res.second) (void) hipFree(d);    // another thread stored it first
  return res.first->second;
}
}  // namespace

size_t mt_segments(uint64_t n) { return (size_t) ((n + MT_SEG_WORDS - 1) / MT_SEG_WORDS); }

size_t mt_window_words(size_t nseg) { return (nseg + MT_FAN) * MT_N; }

// Windows of the stream at words first + s J, s < nseg, into win, with scr
// as the other half of the ping-pong (both mt_window_words(nseg) words): the
// seed window, one jump to `first`, then log16(nseg) levels of fan-out jumps.
hipError_t mt_stream_windows(uint64_t seed, uint64_t first, size_t nseg, uint32_t *win,
    uint32_t *scr, hipStream_t st) {
  if (nseg == 0) return hipSuccess;
  const uint64_t J = MT_SEG_WORDS;
  int L = 0;
  uint64_t cap = 1;
  while (cap < nseg) { cap *= MT_FAN; L++; }
  hipError_t e;
  // levels ping-pong between win and scr; the root goes where the last
  // level ends in win
  uint32_t *wa = (L % 2 == 0) ? win : scr, *wb = (L % 2 == 0) ? scr : win;
  hipLaunchKernelGGL(k_mt_seed, dim3(1), dim3(64), 0, st, seed, wa);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (first) {                           // window at word `first`
    const uint32_t *g = device_fan_polys(first, 2);
    if (!g) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mt_fan, dim3(2), dim3(MT_JTPB), 0, st, g, wa, 2, (size_t) 2, wb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(wa, wb + MT_N, MT_N * sizeof(uint32_t), hipMemcpyDeviceToDevice, st)) != hipSuccess)
      return e;
  }
  // level l: windows at stride J FAN^(L-1-l), each one jump x^(k stride)
  // (k < FAN) from a window of the coarser level
  for (int l = 0; l < L; l++) {
    uint64_t stride = J;
    for (int k = 0; k < L - 1 - l; k++) stride *= MT_FAN;
    const uint64_t need = (nseg + stride / J - 1) / (stride / J);
    const uint32_t *g = device_fan_polys(stride, MT_FAN);
    if (!g) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mt_fan, dim3((unsigned) need), dim3(MT_JTPB), 0, st, g, wa,
        MT_FAN, (size_t) need, wb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t *t = wa; wa = wb; wb = t;
  }
  return hipSuccess;
}

// Words [first, first + n) into out from the windows win of
// mt_stream_windows(seed, first, nseg >= mt_segments(n)).
hipError_t mt_stream_fill(const uint32_t *win, uint64_t n, uint32_t *out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mt_fill, dim3((unsigned) mt_segments(n)), dim3(MT_TPB), 0, st, win,
      (uint64_t) MT_SEG_WORDS, n, out);
  return hipGetLastError();
}
