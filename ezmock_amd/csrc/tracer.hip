// tracer.hip -- effective-bias tracer population on gfx950.
//
// Stages of EZmock_populate_tracer (pop_tracer.c:1140-1219):
//   k_count_valid + scan + k_bias   bias_model, pop_tracer.c:149-348: valid
//                                   cell of rank r uses the Box-Muller pair
//                                   r/2 drawn at 2N + 2(r/2).
//   k_compact + radix sort + k_walk n_tracer_in_cell, :359-514: the PDF
//                                   bins are rank ranges of the sorted tracer
//                                   densities, each extended over the ties at
//                                   its lowest value (:436-446).
//   k_assign_nt + scan              tracers per cell and their offsets.
//   k_positions                     generate_tracers, :881-1007: tracer t at
//                                   draws 5N + 3t.
//   k_velocities                    disp_cic normalisation (:676-682) fused
//                                   into assign_vel (:697-753) and
//                                   add_vel_scatter (:1043-1113).
// Compiled with -ffp-contract=off (IEEE operation order of the reference).
#include "common.h"
#include "rng.h"
#include "kernels.h"

#define CPB EZ_CELLS_PER_BLOCK
#define TPB 256
#define CPT (CPB / TPB)

// ------------------------------------------------------------ block scan
template <class T>
__device__ __forceinline__ T block_excl_scan(T v, T &total, T *ws) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < TPB / 64; i++) {
    T s = ws[i];
    if (i < w) off += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

// ------------------------------------------------------------ u32 scans
__global__ void __launch_bounds__(TPB)
k_scan_chunks(const uint32_t *__restrict__ in, uint64_t *__restrict__ out,
    size_t n, uint64_t *__restrict__ sums) {
  __shared__ uint64_t ws[TPB / 64];
  const size_t base = (size_t) blockIdx.x * CPB + (size_t) threadIdx.x * CPT;
  uint32_t v[CPT];
  uint64_t s = 0;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const uint32_t t = in[base + q < n ? base + q : n - 1];   // no branch per load
    v[q] = (base + q < n) ? t : 0u;
    s += v[q];
  }
  uint64_t tot;
  uint64_t off = block_excl_scan<uint64_t>(s, tot, ws);
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    if (base + q < n) out[base + q] = off;
    off += v[q];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// In-place exclusive scan of up to CPB u64 values by one block.
__global__ void __launch_bounds__(TPB)
k_scan_top(uint64_t *__restrict__ s, size_t n, uint64_t *__restrict__ total) {
  __shared__ uint64_t ws[TPB / 64];
  const size_t base = (size_t) threadIdx.x * CPT;
  uint64_t v[CPT], t = 0;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const uint64_t x = s[base + q < n ? base + q : n - 1];
    v[q] = (base + q < n) ? x : 0;
    t += v[q];
  }
  uint64_t tot;
  uint64_t off = block_excl_scan<uint64_t>(t, tot, ws);
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    if (base + q < n) s[base + q] = off;
    off += v[q];
  }
  if (threadIdx.x == 0 && total) *total = tot;
}

__global__ void k_scan_add(uint64_t *__restrict__ out, size_t n,
    const uint64_t *__restrict__ sums) {
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / CPB];
}

// Exclusive scan of 1 <= n <= SCAN1_MAX u32 in one launch, by one workgroup
// of SCAN1_NT threads: the mock's block-count scans (32768 values at 512^3,
// 4096 at 256^3) take one launch instead of the three of the chunked scan
// (13.9 us against 8.1 + 4.7 + 5.1 us at 32768 values, where the one CU's
// vector work sets the time; one 256^3 instance 694 -> 721 mocks/s).  Values are
// loaded coalesced into LDS; thread t sums its segment [t S, t S + S) with
// S = ceil(n / SCAN1_NT), the segment sums are scanned across the block,
// each segment is turned into its prefixes relative to the segment start
// (u32, since a segment sum below 2^32 is checked), and the outputs are
// stored coalesced.  A segment sum of 2^32 or more switches the block to
// 64-bit prefixes stored per segment.  (A version keeping 16 uint4 per
// thread in registers spilled and took 84 us.)
#define SCAN1_NT 1024
#define SCAN1_MAX 32768
#define SCAN1_PAD(e) ((e) + ((e) >> 5))

__global__ void __launch_bounds__(SCAN1_NT)
k_scan_one(const uint32_t *__restrict__ in, uint64_t *__restrict__ out, uint32_t n,
    FastDiv fS, uint64_t *__restrict__ total) {
  __shared__ uint32_t buf[SCAN1_PAD(SCAN1_MAX)];
  __shared__ uint64_t toff[SCAN1_NT];
  __shared__ uint64_t ws[SCAN1_NT / 64];
  __shared__ uint32_t wide;
  const uint32_t t = threadIdx.x;
  if (t == 0) wide = 0;
  constexpr int SMAX = SCAN1_MAX / SCAN1_NT;
  const uint32_t S = fS.d;                              // ceil(n / SCAN1_NT) <= SMAX
  // coalesced loads, all in flight together (indices clamped), then the
  // LDS stores of the values inside [0, n)
  uint32_t v[SMAX];
#pragma unroll
  for (int r = 0; r < SMAX; r++) {
    const uint32_t e = (uint32_t) r * SCAN1_NT + t;
    v[r] = in[e < n ? e : n - 1];
  }
#pragma unroll
  for (int r = 0; r < SMAX; r++) {
    const uint32_t e = (uint32_t) r * SCAN1_NT + t;
    if (e < n) buf[SCAN1_PAD(e)] = v[r];
  }
  __syncthreads();
  const uint32_t b = t * S;
  uint64_t s = 0;
#pragma unroll 8
  for (uint32_t q = 0; q < S; q++)
    if (b + q < n) s += buf[SCAN1_PAD(b + q)];
  if (s >> 32) atomicOr(&wide, 1u);
  const int lane = (int) (t & 63), w = (int) (t >> 6);
  uint64_t x = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  uint64_t off = x - s, tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN1_NT / 64; i++) {
    const uint64_t a = ws[i];
    if (i < w) off += a;
    tot += a;
  }
  if (t == 0) *total = tot;
  if (wide) {                                 // uniform after the barrier
    for (uint32_t q = 0; q < S && b + q < n; q++) {
      out[b + q] = off;
      off += buf[SCAN1_PAD(b + q)];
    }
    return;
  }
  uint32_t acc = 0;
#pragma unroll 8
  for (uint32_t q = 0; q < S; q++) {
    if (b + q < n) {
      const uint32_t v = buf[SCAN1_PAD(b + q)];
      buf[SCAN1_PAD(b + q)] = acc;
      acc += v;
    }
  }
  toff[t] = off;
  __syncthreads();
#pragma unroll 8
  for (uint32_t r = 0; r < S; r++) {
    const uint32_t e = r * SCAN1_NT + t;
    if (e < n) out[e] = toff[fastdiv_q(e, fS)] + buf[SCAN1_PAD(e)];
  }
}

size_t scan_tmp_words(size_t n) { return (n + CPB - 1) / CPB + 1; }

// Exclusive scan of n <= CPB*CPB u32 into u64; out[n] is not written, the
// grand total lands in tmp[nchunks].
hipError_t launch_scan_u32(const uint32_t *in, uint64_t *out, size_t n,
    uint64_t *tmp, hipStream_t st) {
  const size_t nch = (n + CPB - 1) / CPB;
  if (nch > CPB) return hipErrorInvalidValue;
  if (nch == 0) return hipSuccess;
  if (n <= SCAN1_MAX) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(SCAN1_NT), 0, st, in, out, (uint32_t) n,
        fastdiv_make((uint32_t) ((n + SCAN1_NT - 1) / SCAN1_NT)), tmp + nch);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_scan_chunks, dim3((unsigned) nch), dim3(TPB), 0, st, in, out, n, tmp);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(TPB), 0, st, tmp, nch, tmp + nch);
  hipLaunchKernelGGL(k_scan_add, dim3(ez_blocks(n, 256)), dim3(256), 0, st, out, n, (const uint64_t *) tmp);
  return hipGetLastError();
}

// ------------------------------------------------------------ bias model
__global__ void __launch_bounds__(TPB)
k_count_valid(const real *__restrict__ rho, real thres, size_t N,
    uint32_t *__restrict__ blockcnt, const PopCtl *__restrict__ ctl) {
  __shared__ uint32_t ws[TPB / 64];
  if (ctl) thres = ctl->bp.thres;
  const size_t base = (size_t) blockIdx.x * CPB;
  // unconditional loads at clamped indices: all in flight together (a
  // guarded load waits for its data at the branch join)
  real v[CPT];
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = base + (size_t) q * TPB + threadIdx.x;
    v[q] = rho[i < N ? i : N - 1];
  }
  uint32_t c = 0;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = base + (size_t) q * TPB + threadIdx.x;
    if (i < N && v[q] >= thres) c++;
  }
  uint32_t tot;
  (void) block_excl_scan<uint32_t>(c, tot, ws);
  if (threadIdx.x == 0) blockcnt[blockIdx.x] = tot;
}

hipError_t launch_count_valid(const real *rho, real thres, size_t N,
    uint32_t *blockcnt, hipStream_t st, const PopCtl *ctl) {
  hipLaunchKernelGGL(k_count_valid, dim3(ez_blocks(N, CPB)), dim3(TPB), 0, st,
      rho, thres, N, blockcnt, ctl);
  return hipGetLastError();
}

// Histogram bin of a tracer density, pop_tracer.c:265-268: imax - ilogb(v)
// clamped to [0, 63], the subtraction wrapping as 32-bit int arithmetic.
__device__ __forceinline__ int hist_bin(real v, int imax) {
  int d = (int) ((unsigned) imax - (unsigned) ilogb(v));
  if (d < 0) d = 0;
  else if (d >= EZ_NUM_LOGB_BIN) d = EZ_NUM_LOGB_BIN - 1;
  return d;
}

// Sequential draws from one stream position on: an MRG32k3a state, or
// MT19937 output words generated beforehand (w != nullptr, two words per
// draw, d = the draw's index in w).
struct Draws {
  MrgState g;
  const uint32_t *w;
  uint64_t d;
  __device__ __forceinline__ double next() { return w ? mt_draw(w, d++) : mrg_next(g); }
  __device__ __forceinline__ double next_pos() { return w ? mt_draw_pos(w, d++) : mrg_next(g); }
};

// pop_tracer.c:70-79
__device__ __forceinline__ void gauss_pair(Draws &s, double sigma,
    double &g1, double &g2) {
  const double r = -2 * log(s.next_pos());
  const double th = s.next() * 2 * EZ_PI;
  // one argument reduction for both (the values of separate sin and cos)
  double sn, cs;
  sincos(th, &sn, &cs);
  g1 = sigma * sqrt(r) * cs;
  g2 = sigma * sqrt(r) * sn;
}

// LDS tile with one pad word per 16 so that thread-sequential access
// (thread t reads cells 16t .. 16t+15) is bank-conflict free.
#define PADIDX(c) ((c) + ((c) >> 4))

__global__ void __launch_bounds__(TPB)
k_bias(const real *__restrict__ rho, real *__restrict__ rhot,
    const uint64_t *__restrict__ blockoff, unsigned long long *__restrict__ hist,
    uint32_t *__restrict__ blockhist, BiasParams bp_in, const uint32_t *__restrict__ jtab,
    const uint32_t *__restrict__ mtw, const PopCtl *__restrict__ ctl) {
  const BiasParams bp = ctl ? ctl->bp : bp_in;
  __shared__ real tile[PADIDX(CPB)];
  __shared__ uint16_t vcell[CPB];       // block-local rank -> cell
  __shared__ uint32_t lh[EZ_NUM_LOGB_BIN];
  __shared__ uint64_t ws[TPB / 64];
  __shared__ MrgState base;
  const size_t b0 = (size_t) blockIdx.x * CPB;
  const int tid = threadIdx.x;
  if (tid < EZ_NUM_LOGB_BIN) lh[tid] = 0;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int c = q * TPB + tid;
    const size_t i = b0 + c;
    const real t = rho[i < bp.N ? i : bp.N - 1];
    tile[PADIDX(c)] = (i < bp.N) ? t : (real) 0;
  }
  __syncthreads();
  // thread tid owns cells c = 16 tid .. 16 tid + 15 of the block: rank the
  // valid ones, clear the others
  uint64_t cnt = 0;
  uint32_t vmask = 0;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int c = tid * CPT + q;
    if (b0 + c < bp.N && tile[PADIDX(c)] >= bp.thres) { cnt++; vmask |= 1u << q; }
  }
  uint64_t tot;
  const uint64_t excl = block_excl_scan<uint64_t>(cnt, tot, ws);
  {
    int l = (int) excl;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const int c = tid * CPT + q;
      if ((vmask >> q) & 1) vcell[l++] = (uint16_t) c;
      else tile[PADIDX(c)] = 0;
    }
  }
  const uint64_t rb = blockoff[blockIdx.x];
  const uint64_t pb = rb >> 1;
  // pairs pb .. pe cover global ranks rb .. rb + tot - 1
  const uint64_t npair = tot ? ((rb + tot - 1) >> 1) - pb + 1 : 0;
  const uint64_t ppt = (npair + TPB - 1) / TPB;
  if (tid == 0 && npair && !mtw) {
    MrgState g = bp.seed_state;            // advanced to draw 2N
    mrg_jump(g, 2 * pb, jtab);
    base = g;
  }
  __syncthreads();
  // thread tid draws pairs pb + tid ppt .. pb + (tid + 1) ppt - 1 in order:
  // one Box-Muller pair per iteration on every lane
  const uint64_t m0 = (uint64_t) tid * ppt;
  if (m0 < npair) {
    Draws ds;
    ds.w = mtw;
    ds.d = 2 * (pb + m0);                  // MT19937 words start at draw 2N
    if (!mtw) {
      ds.g = base;
      mrg_jump(ds.g, 2 * m0, jtab);
    }
    const uint64_t m1 = (m0 + ppt < npair) ? m0 + ppt : npair;
#pragma unroll 1
    for (uint64_t m = m0; m < m1; m++) {
      double gg[2];
      gauss_pair(ds, bp.lambda, gg[0], gg[1]);
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint64_t r = 2 * (pb + m) + h;          // global rank
        if (r < rb || r >= rb + tot) continue;
        const double gp = (gg[h] >= 0) ? 1 + gg[h] : exp(gg[h]);
        const int c = vcell[r - rb];
        const real v = tile[PADIDX(c)];
        const real rt = (v > bp.rho_sat) ? bp.rho_sat : v;
        const double tmp = (1 - exp((double) rt * bp.iexp)) * gp;
        const real o = (real) tmp;
        tile[PADIDX(c)] = o;
        atomicAdd(&lh[hist_bin(o, bp.imax)], 1u);
      }
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int q = 0; q < CPT; q++) {
    const int c = q * TPB + tid;
    const size_t i = b0 + c;
    if (i < bp.N) rhot[i] = tile[PADIDX(c)];
  }
  if (tid < EZ_NUM_LOGB_BIN) {
    if (lh[tid]) atomicAdd(&hist[tid], (unsigned long long) lh[tid]);
    if (blockhist) blockhist[(size_t) blockIdx.x * EZ_NUM_LOGB_BIN + tid] = lh[tid];
  }
}

hipError_t launch_bias(const real *rho, real *rhot, const uint64_t *blockoff,
    unsigned long long *hist, uint32_t *blockhist, const BiasParams &bp,
    const uint32_t *jtab, const uint32_t *mtw, hipStream_t st, const PopCtl *ctl) {
  BiasParams b2 = bp;
  if (!ctl) mrg_jump(b2.seed_state, (uint64_t) 2 * bp.N, mrg_host_jump_table());
  hipLaunchKernelGGL(k_bias, dim3(ez_blocks(bp.N, CPB)), dim3(TPB), 0, st,
      rho, rhot, blockoff, hist, blockhist, b2, jtab, mtw, ctl);
  return hipGetLastError();
}

// pop_tracer.c:361-369 on the device: the first log2 bin where the
// cumulative count of tracer densities reaches ncell, its threshold
// 2^(imax_bias - imax) (REAL_MIN for the last bin) and the candidates to
// sort; one thread (64 bins).
__global__ void __launch_bounds__(64)
k_select(const unsigned long long *__restrict__ hist, const uint64_t *__restrict__ nvalid_p,
    PopCtl *__restrict__ ctl) {
  if (threadIdx.x) return;
  const uint64_t nvalid = *nvalid_p, ncell = ctl->ncell;
  uint32_t fl = 0;
  if (nvalid < ncell) fl |= POP_RHO_LESS;
  uint64_t c = 0;
  int im;
  for (im = 0; im < EZ_NUM_LOGB_BIN; im++) {
    c += hist[im];
    if (c >= ncell) break;
  }
  if (im == EZ_NUM_LOGB_BIN) {
    fl |= POP_NO_BIN;
    im = EZ_NUM_LOGB_BIN - 1;
  }
  ctl->imax = im;
  ctl->rho_thres = (im == EZ_NUM_LOGB_BIN - 1) ? (real) EZ_REAL_MIN
      : (real) ldexp(1.0, ctl->bp.imax - im);
  ctl->nvalid = nvalid;
  ctl->cnt = c;
  ctl->nsort = c > ctl->kcap ? ctl->kcap : c;
  ctl->nfine = ctl->nsort;
  ctl->kfine = ~(rkey) 0;
  if (c > ctl->kcap) fl |= POP_KEYS_FULL;
  ctl->flags = fl;
}

hipError_t launch_select(const unsigned long long *hist, const uint64_t *nvalid, PopCtl *ctl,
    hipStream_t st) {
  hipLaunchKernelGGL(k_select, dim3(1), dim3(64), 0, st, hist, nvalid, ctl);
  return hipGetLastError();
}

// Log2 histogram of a tracer-density array (pop_tracer.c:265-268 with
// ilogb_max 5; zero densities are not counted): 64 u64 counters, which the
// caller clears.  Lanes of a wave that share the first active lane's bin add
// with one LDS atomic (densities concentrate in a few bins).
__global__ void __launch_bounds__(TPB)
k_hist_rhot(const real *__restrict__ rhot, size_t N, unsigned long long *__restrict__ hist) {
  __shared__ uint32_t h[EZ_NUM_LOGB_BIN];
  if (threadIdx.x < EZ_NUM_LOGB_BIN) h[threadIdx.x] = 0;
  __syncthreads();
  const size_t b0 = (size_t) blockIdx.x * CPB;
  real v[CPT];
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = b0 + (size_t) q * TPB + threadIdx.x;
    const real t = rhot[i < N ? i : N - 1];
    v[q] = (i < N) ? t : (real) 0;
  }
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const bool on = v[q] != 0;
    const int d = on ? hist_bin(v[q], 5) : 0;
    uint64_t act = __ballot(on);
    while (act) {
      const int leader = __ffsll((unsigned long long) act) - 1;
      const int ld = __shfl(d, leader, 64);
      const uint64_t peers = __ballot(on && d == ld) & act;
      if (lane == leader) atomicAdd(&h[ld], (uint32_t) __popcll(peers));
      act &= ~peers;
    }
  }
  __syncthreads();
  if (threadIdx.x < EZ_NUM_LOGB_BIN && h[threadIdx.x])
    atomicAdd(&hist[threadIdx.x], (unsigned long long) h[threadIdx.x]);
}

hipError_t launch_hist_rhot(const real *rhot, size_t N, unsigned long long *hist,
    hipStream_t st) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_hist_rhot, dim3(ez_blocks(N, CPB)), dim3(TPB), 0, st, rhot, N, hist);
  return hipGetLastError();
}

// -------------------------------------------------------------- compact
// Keys: bitwise NOT of the bits of a positive real, so that ascending key
// order is descending density.
__device__ __forceinline__ rkey real_key(real v) {
#ifdef SINGLE_PREC
  return ~__float_as_uint(v);
#else
  return ~(uint64_t) __double_as_longlong(v);
#endif
}

// One device atomic per block of 4096 cells reserves the block's range.
__global__ void __launch_bounds__(TPB)
k_compact(const real *__restrict__ rhot, real thres, size_t N,
    rkey *__restrict__ keys, size_t cap, unsigned long long *__restrict__ counter) {
  __shared__ uint32_t ws[TPB / 64];
  __shared__ unsigned long long base;
  const size_t b0 = (size_t) blockIdx.x * CPB;
  real v[CPT];
  uint32_t c = 0;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = b0 + (size_t) q * TPB + threadIdx.x;
    const real t = rhot[i < N ? i : N - 1];
    v[q] = (i < N) ? t : (real) 0;
    if (i < N && v[q] >= thres) c++;
  }
  uint32_t tot;
  const uint32_t excl = block_excl_scan<uint32_t>(c, tot, ws);
  if (threadIdx.x == 0 && tot) base = atomicAdd(counter, (unsigned long long) tot);
  __syncthreads();
  if (!c) return;
  size_t pos = base + excl;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = b0 + (size_t) q * TPB + threadIdx.x;
    if (i < N && v[q] >= thres) {
      if (pos < cap) keys[pos] = real_key(v[q]);
      pos++;
    }
  }
}

hipError_t launch_compact(const real *rhot, real thres, size_t N, rkey *keys,
    size_t cap, unsigned long long *counter, hipStream_t st) {
  hipLaunchKernelGGL(k_compact, dim3(ez_blocks(N, CPB)), dim3(TPB), 0, st,
      rhot, thres, N, keys, cap, counter);
  return hipGetLastError();
}

// Candidate compaction without a shared counter: the bias kernel's
// per-block histograms give each block's number of candidates (the cells in
// log2 bins <= imax, i.e. rhot >= 2^(imax_bias - imax)); their exclusive
// scan gives every block its output range.
__global__ void k_block_cand(const uint32_t *__restrict__ blockhist, size_t nblk,
    int imax, uint32_t *__restrict__ blockcnt, const PopCtl *__restrict__ ctl) {
  const size_t b = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  if (ctl) imax = ctl->imax;
  uint32_t c = 0;
  for (int i = 0; i <= imax; i++) c += blockhist[b * EZ_NUM_LOGB_BIN + i];
  blockcnt[b] = c;
}

hipError_t launch_block_cand(const uint32_t *blockhist, size_t nblk, int imax,
    uint32_t *blockcnt, hipStream_t st, const PopCtl *ctl) {
  hipLaunchKernelGGL(k_block_cand, dim3(ez_blocks(nblk, 256)), dim3(256), 0, st,
      blockhist, nblk, imax, blockcnt, ctl);
  return hipGetLastError();
}

// flag[0] |= 1 if a block's candidates disagree with its histogram count.
__global__ void __launch_bounds__(TPB)
k_compact_at(const real *__restrict__ rhot, real thres, size_t N,
    const uint64_t *__restrict__ blockoff, const uint32_t *__restrict__ blockcnt,
    rkey *__restrict__ keys, size_t cap, unsigned long long *__restrict__ flag,
    const PopCtl *__restrict__ ctl) {
  __shared__ uint32_t ws[TPB / 64];
  if (ctl) thres = ctl->rho_thres;
  const size_t b0 = (size_t) blockIdx.x * CPB;
  real v[CPT];
  uint32_t c = 0;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = b0 + (size_t) q * TPB + threadIdx.x;
    const real t = rhot[i < N ? i : N - 1];
    v[q] = (i < N) ? t : (real) 0;
    if (i < N && v[q] >= thres) c++;
  }
  uint32_t tot;
  const uint32_t excl = block_excl_scan<uint32_t>(c, tot, ws);
  if (threadIdx.x == 0 && tot != blockcnt[blockIdx.x]) atomicOr(flag, 1ull);
  if (!c) return;
  size_t pos = blockoff[blockIdx.x] + excl;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = b0 + (size_t) q * TPB + threadIdx.x;
    if (i < N && v[q] >= thres) {
      if (pos < cap) keys[pos] = real_key(v[q]);
      pos++;
    }
  }
}

hipError_t launch_compact_at(const real *rhot, real thres, size_t N,
    const uint64_t *blockoff, const uint32_t *blockcnt, rkey *keys, size_t cap,
    unsigned long long *flag, hipStream_t st, const PopCtl *ctl) {
  hipLaunchKernelGGL(k_compact_at, dim3(ez_blocks(N, CPB)), dim3(TPB), 0, st,
      rhot, thres, N, blockoff, blockcnt, keys, cap, flag, ctl);
  return hipGetLastError();
}

// ------------------------------------------------- candidate refinement
// The threshold bin imax covers [t, 2t), t = rho_thres a power of two
// (except bin 0, open above, and the last bin, REAL_MIN: not refined).  Its
// candidates are counted by their top four mantissa bits; the refined
// threshold t (1 + s/16) is the largest whose count of candidates at or
// above it still reaches ncell, so the ncell-th largest density and its
// ties stay above it and the walk over the kept keys gives the same PDF.
#ifdef SINGLE_PREC
#define KEY_MANT 23
#else
#define KEY_MANT 52
#endif

__device__ __forceinline__ bool refine_on(const PopCtl *ctl) {
  return ctl->imax > 0 && ctl->imax < EZ_NUM_LOGB_BIN - 1 && !(ctl->flags & POP_KEYS_FULL);
}

// Per block b of CPB keys: the candidates of the threshold bin by sub-bin
// (bc[s nb + b], s < 16) and those above the bin (bc[16 nb + b]), with plain
// stores: the block totals are combined by one workgroup (k_refine_plan).
#define REFINE_W 17

__global__ void __launch_bounds__(TPB)
k_key_subhist(const rkey *__restrict__ keys, const PopCtl *__restrict__ ctl,
    uint32_t *__restrict__ bc) {
  const size_t nb = gridDim.x;
  __shared__ uint32_t h[REFINE_W];
  if (!refine_on(ctl)) return;                      // uniform
  const uint64_t n = ctl->nsort;
  const size_t b0 = (size_t) blockIdx.x * CPB;
  if (b0 >= n) return;
  if (threadIdx.x < REFINE_W) h[threadIdx.x] = 0;
  __syncthreads();
  const rkey tb = ~real_key(ctl->rho_thres);        // bits of t
  const rkey eb = tb >> KEY_MANT;                   // its exponent field
  rkey kv[CPT];
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = b0 + (size_t) q * TPB + threadIdx.x;
    kv[q] = keys[i < n ? i : n - 1];
  }
  // counts in registers, summed over the wave, one LDS add per bin and wave
  // (the keys crowd into few sub-bins: per-key atomics serialise)
  uint32_t c[REFINE_W];
#pragma unroll
  for (int b = 0; b < REFINE_W; b++) c[b] = 0;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = b0 + (size_t) q * TPB + threadIdx.x;
    const rkey vb = ~kv[q];
    const bool in = i < n && (vb >> KEY_MANT) == eb;
    const uint32_t sb = (uint32_t) (vb >> (KEY_MANT - 4)) & 15u;
#pragma unroll
    for (int b = 0; b < 16; b++) c[b] += (in && sb == (uint32_t) b) ? 1u : 0u;
    c[16] += (i < n && !in) ? 1u : 0u;
  }
#pragma unroll
  for (int b = 0; b < REFINE_W; b++) {
    uint32_t v = c[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&h[b], v);
  }
  __syncthreads();
  if (threadIdx.x < REFINE_W) bc[(size_t) threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// pop_tracer.c:361-369 refined: the sub-bin s of the threshold bin where
// the count of candidates at or above t (1 + s/16) first reaches ncell.
__device__ __forceinline__ int refine_threshold(const uint64_t *__restrict__ sub,
    PopCtl *__restrict__ ctl) {
  uint64_t inbin = 0;
  for (int s = 0; s < 16; s++) inbin += sub[s];
  uint64_t acc = ctl->nsort - inbin;                // candidates above the bin
  int s = 15;
  for (; s > 0; s--) {
    if (acc + sub[s] >= ctl->ncell) break;
    acc += sub[s];
  }
  acc += sub[s];
  const rkey tb = ~real_key(ctl->rho_thres);
  ctl->kfine = ~(tb | ((rkey) s << (KEY_MANT - 4)));
  ctl->nfine = acc;
  return s;
}

// One workgroup: the block counts of k_key_subhist summed, the refined
// threshold, and each key block's output offset for k_key_keep (exclusive
// scan of the blocks' kept counts, in rounds of REFINE_NT blocks; all keys
// of a block when the bin is not refined).  At 512^3 (2137 key blocks) the
// three kernels take 23 + 24 + 25 us, against 43 + 5 + 34 us with device
// atomics (16 sub-bin totals on one line, one output counter).
#define REFINE_NT 1024

__device__ __forceinline__ uint64_t refine_kept(const uint32_t *__restrict__ bc, size_t nb,
    size_t b, bool on, int s, uint64_t n) {
  if (!on) {
    const uint64_t r = n - (uint64_t) b * CPB;
    return r < CPB ? r : CPB;
  }
  uint32_t k = bc[16 * nb + b];
  for (int j = s; j < 16; j++) k += bc[(size_t) j * nb + b];
  return k;
}

__global__ void __launch_bounds__(REFINE_NT)
k_refine_plan(const uint32_t *__restrict__ bc, size_t nb, PopCtl *__restrict__ ctl,
    uint64_t *__restrict__ boff) {
  __shared__ uint64_t red[REFINE_NT / 64][16];
  __shared__ uint64_t sub[16];
  __shared__ uint64_t ws[REFINE_NT / 64];
  __shared__ int ssel;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint64_t n = ctl->nsort;
  const size_t nba = (size_t) ((n + CPB - 1) / CPB);      // <= nb
  const bool on = refine_on(ctl);                   // uniform
  if (on) {
    uint64_t a[16];
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] = 0;
    for (size_t b = tid; b < nba; b += REFINE_NT)
#pragma unroll
      for (int j = 0; j < 16; j++) a[j] += bc[(size_t) j * nb + b];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      uint64_t v = a[j];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[w][j] = v;
    }
    __syncthreads();
    if (tid < 16) {
      uint64_t v = 0;
      for (int i = 0; i < REFINE_NT / 64; i++) v += red[i][tid];
      sub[tid] = v;
    }
    __syncthreads();
    if (tid == 0) ssel = refine_threshold(sub, ctl);
  }
  __syncthreads();
  const int s = on ? ssel : 0;
  uint64_t carry = 0;
  for (size_t b0 = 0; b0 < nba; b0 += REFINE_NT) {           // uniform rounds
    const size_t b = b0 + tid;
    const uint64_t cnt = b < nba ? refine_kept(bc, nb, b, on, s, n) : 0;
    uint64_t x = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    uint64_t off = carry + x - cnt, rt = 0;
    for (int i = 0; i < REFINE_NT / 64; i++) {
      const uint64_t v = ws[i];
      if (i < w) off += v;
      rt += v;
    }
    if (b < nba) boff[b] = off;
    carry += rt;
    __syncthreads();                                          // ws reused
  }
  if (tid == 0 && ctl->nfine > ctl->kfcap) ctl->flags |= POP_KEYS_FULL;
}

// Keys at or above the refined threshold (key <= kfine) of each block, at
// the block's offset (the sort that follows orders them).
__global__ void __launch_bounds__(TPB)
k_key_keep(const rkey *__restrict__ keys, const PopCtl *__restrict__ ctl,
    const uint64_t *__restrict__ boff, rkey *__restrict__ out) {
  __shared__ uint32_t ws[TPB / 64];
  const uint64_t n = ctl->nsort;
  const size_t b0 = (size_t) blockIdx.x * CPB;
  if (b0 >= n) return;
  const rkey kf = ctl->kfine;
  rkey kv[CPT];
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = b0 + (size_t) threadIdx.x * CPT + q;
    kv[q] = keys[i < n ? i : n - 1];
  }
  uint32_t c = 0;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = b0 + (size_t) threadIdx.x * CPT + q;
    if (i < n && kv[q] <= kf) c++;
  }
  uint32_t tot;
  uint32_t pos = block_excl_scan<uint32_t>(c, tot, ws);
  const uint64_t bb = boff[blockIdx.x];
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = b0 + (size_t) threadIdx.x * CPT + q;
    if (i < n && kv[q] <= kf) out[bb + pos++] = kv[q];
  }
}

hipError_t launch_refine(const rkey *keys, size_t kcap, PopCtl *ctl, uint32_t *bc,
    uint64_t *boff, rkey *out, hipStream_t st) {
  const unsigned nb = (unsigned) ez_blocks(kcap, CPB);
  hipLaunchKernelGGL(k_key_subhist, dim3(nb), dim3(TPB), 0, st, keys, (const PopCtl *) ctl, bc);
  hipLaunchKernelGGL(k_refine_plan, dim3(1), dim3(REFINE_NT), 0, st, (const uint32_t *) bc,
      (size_t) nb, ctl, boff);
  hipLaunchKernelGGL(k_key_keep, dim3(nb), dim3(TPB), 0, st, keys, (const PopCtl *) ctl,
      (const uint64_t *) boff, out);
  return hipGetLastError();
}

// ----------------------------------------------------------- radix sort
// LSD, 8-bit digits, 4096 keys per block; stable ranks within a block from
// wave ballots (lanes with equal digit) and per-wave digit counts.
#define SORT_ITEMS CPT

__global__ void __launch_bounds__(TPB)
k_sort_hist(const rkey *__restrict__ in, size_t n, int shift,
    uint32_t *__restrict__ gh, size_t nblk, const uint64_t *__restrict__ np) {
  __shared__ uint32_t h[256];
  if (np) n = *np;
  const size_t base = (size_t) blockIdx.x * CPB;
  if (base >= n) {                        // past the keys: an empty column
    gh[(size_t) threadIdx.x * nblk + blockIdx.x] = 0;
    return;
  }
  h[threadIdx.x] = 0;
  __syncthreads();
  rkey kv[SORT_ITEMS];
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; it++) {
    const size_t e = base + (size_t) it * TPB + threadIdx.x;
    kv[it] = in[e < n ? e : n - 1];               // all loads in flight
  }
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; it++) {
    const size_t e = base + (size_t) it * TPB + threadIdx.x;
    if (e < n) atomicAdd(&h[(uint32_t) (kv[it] >> shift) & 255u], 1u);
  }
  __syncthreads();
  gh[(size_t) threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(TPB)
k_sort_scatter(const rkey *__restrict__ in, rkey *__restrict__ out, size_t n,
    int shift, const uint64_t *__restrict__ gofs, size_t nblk,
    const uint64_t *__restrict__ np) {
  __shared__ uint64_t bofs[256];
  __shared__ uint32_t wcnt[TPB / 64][256];
  __shared__ uint32_t wofs[TPB / 64][256];
  __shared__ uint32_t run[256];
  if (np) n = *np;
  if ((size_t) blockIdx.x * CPB >= n) return;     // uniform over the block
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  bofs[tid] = gofs[(size_t) tid * nblk + blockIdx.x];
  run[tid] = 0;
#pragma unroll
  for (int q = 0; q < TPB / 64; q++) wcnt[q][tid] = 0;
  __syncthreads();
  const size_t base = (size_t) blockIdx.x * CPB;
  const uint64_t lt = (1ull << lane) - 1;
  rkey kv[SORT_ITEMS];                    // all key loads in flight first
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; it++) {
    const size_t e = base + (size_t) it * TPB + tid;
    kv[it] = in[e < n ? e : n - 1];
  }
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; it++) {
    const size_t e = base + (size_t) it * TPB + tid;
    const bool valid = e < n;
    const rkey key = valid ? kv[it] : (rkey) 0;
    const uint32_t d = (uint32_t) (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; bit++) {
      const bool on = (d >> bit) & 1u;
      const uint64_t bb = __ballot(valid && on);
      peers &= on ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t) __popcll(peers & lt);
    if (valid && rank == 0) wcnt[w][d] = (uint32_t) __popcll(peers);
    __syncthreads();
    {
      uint32_t r = run[tid];
#pragma unroll
      for (int q = 0; q < TPB / 64; q++) {
        const uint32_t c = wcnt[q][tid];
        wofs[q][tid] = r;
        r += c;
        wcnt[q][tid] = 0;
      }
      run[tid] = r;
    }
    __syncthreads();
    if (valid) out[bofs[d] + wofs[w][d] + rank] = key;
  }
}

size_t sort_hist_words(size_t n) { return 256 * ((n + CPB - 1) / CPB); }

// Sorts keys[0..n) ascending; tmp has n entries; ghist holds
// sort_hist_words(n) u32; gofs the same number of u64; scan_tmp
// scan_tmp_words(sort_hist_words(n)) u64.
hipError_t launch_sort(rkey *keys, rkey *tmp, size_t n, uint32_t *ghist,
    uint64_t *gofs, uint64_t *scan_tmp, hipStream_t st, const uint64_t *np) {
  if (n == 0) return hipSuccess;
  const size_t nblk = (n + CPB - 1) / CPB;
  rkey *src = keys, *dst = tmp;
  const int npass = (int) sizeof(rkey);
  for (int p = 0; p < npass; p++) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_sort_hist, dim3((unsigned) nblk), dim3(TPB), 0, st,
        src, n, shift, ghist, nblk, np);
    hipError_t e = launch_scan_u32(ghist, gofs, 256 * nblk, scan_tmp, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sort_scatter, dim3((unsigned) nblk), dim3(TPB), 0, st,
        src, dst, n, shift, gofs, nblk, np);
    rkey *t = src; src = dst; dst = t;
  }
  // npass is even (4 or 8): the result is back in `keys`.
  return hipGetLastError();
}

// ------------------------------------------------------------- PDF walk
// One wave.  sorted: ascending keys (descending density), M of them.
// For PDF bin p (nmax - p tracers per cell): its last rank is s + num[p] - 1
// and the bin takes every further key equal to the key there.
// Outputs: vb[p] = largest key of bin p (0 when bins 0..p are empty),
// cnt[p] = cells in bin p, res = {ncell, ntot, error}.
__global__ void __launch_bounds__(64)
k_walk(const rkey *__restrict__ sorted, size_t M, const uint64_t *__restrict__ num,
    int nmax, rkey *__restrict__ vb, uint64_t *__restrict__ cnt,
    uint64_t *__restrict__ res, PopCtl *__restrict__ ctl,
    const uint64_t *__restrict__ ncand) {
  const int lane = threadIdx.x;
  if (ctl) {
    M = ctl->nfine;
    nmax = ctl->nmax;
  }
  size_t s = 0;
  uint64_t ntot = 0;
  int err = 0;
  rkey last = 0;
  for (int p = 0; p < nmax; p++) {
    const uint64_t n = num[p];
    if (n == 0) {
      if (lane == 0) { vb[p] = last; cnt[p] = 0; }
      continue;
    }
    const size_t e = s + n;
    if (e > M) { err = 1; break; }
    const rkey v = sorted[e - 1];
    size_t pos = e;
    for (;;) {
      const size_t q = pos + lane;
      const bool stop = (q >= M) || (sorted[q] != v);
      const uint64_t m = __ballot(stop);
      if (m) { pos += __ffsll((unsigned long long) m) - 1; break; }
      pos += 64;
    }
    if (lane == 0) { vb[p] = v; cnt[p] = pos - s; }
    ntot += (uint64_t) (pos - s) * (uint64_t) (nmax - p);
    last = v;
    s = pos;
  }
  if (lane == 0) {
    res[0] = s; res[1] = ntot; res[2] = err;
    if (ctl) {
      uint32_t fl = ctl->flags;
      if (err) fl |= POP_WALK_ERR;
      if (ntot > ctl->ccap) fl |= POP_CAT_FULL;
      if (ncand && *ncand != ctl->cnt) fl |= POP_CAND_ERR;
      ctl->ncell_out = s;
      ctl->ntot = ntot;
      ctl->flags = fl;
    }
  }
}

hipError_t launch_walk(const rkey *sorted, size_t M, const uint64_t *num,
    int nmax, rkey *vb, uint64_t *cnt, uint64_t *res, hipStream_t st,
    PopCtl *ctl, const uint64_t *ncand) {
  hipLaunchKernelGGL(k_walk, dim3(1), dim3(64), 0, st, sorted, M, num, nmax,
      vb, cnt, res, ctl, ncand);
  return hipGetLastError();
}

// ------------------------------------------------------ tracers per cell
__global__ void __launch_bounds__(TPB)
k_assign_nt(const real *__restrict__ rhot, size_t N, const rkey *__restrict__ vb,
    int nmax, int16_t *__restrict__ nt, uint32_t *__restrict__ blocktr,
    const PopCtl *__restrict__ ctl) {
  __shared__ rkey sb[EZ_MAX_NUM_TRACER_PER_CELL];
  __shared__ uint32_t ws[TPB / 64];
  if (ctl) nmax = ctl->nmax;
  for (int p = threadIdx.x; p < nmax; p += TPB) sb[p] = vb[p];
  __syncthreads();
  const rkey vlast = sb[nmax - 1];
  const size_t base = (size_t) blockIdx.x * CPB;
  uint32_t sum = 0;
  real vv[CPT];
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const size_t i = base + (size_t) q * TPB + threadIdx.x;
    vv[q] = rhot[i < N ? i : N - 1];              // all loads in flight
  }
#pragma unroll 4
  for (int q = 0; q < CPT; q++) {
    const size_t i = base + (size_t) q * TPB + threadIdx.x;
    if (i >= N) break;
    const real v = vv[q];
    int16_t c = 0;
    if (v > 0) {
      const rkey u = real_key(v);
      if (u <= vlast) {
        int lo = 0, hi = nmax - 1;     // first p with u <= vb[p]
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (u <= sb[mid]) hi = mid; else lo = mid + 1;
        }
        c = (int16_t) (nmax - lo);
      }
    }
    nt[i] = c;
    sum += (uint32_t) c;
  }
  uint32_t tot;
  (void) block_excl_scan<uint32_t>(sum, tot, ws);
  if (threadIdx.x == 0) blocktr[blockIdx.x] = tot;
}

hipError_t launch_assign_nt(const real *rhot, size_t N, const rkey *vb,
    int nmax, int16_t *nt, uint32_t *blocktr, hipStream_t st, const PopCtl *ctl) {
  hipLaunchKernelGGL(k_assign_nt, dim3(ez_blocks(N, CPB)), dim3(TPB), 0, st,
      rhot, N, vb, nmax, nt, blocktr, ctl);
  return hipGetLastError();
}

// ------------------------------------------------------------ positions
// Per block of CPB cells: the tracer counts are ranked (block scan), then
// each thread generates a contiguous run of the block's tracers, so every
// lane draws one tracer per iteration whatever the occupancy of its cells.
// seed: the stream already advanced to draw 5N.
__global__ void __launch_bounds__(TPB)
k_positions(const int16_t *__restrict__ nt, const uint64_t *__restrict__ blockofs,
    FastDiv fNg, double Lbox, real gs, size_t N, MrgState seed,
    const uint32_t *__restrict__ jtab, const uint32_t *__restrict__ mtw,
    real *__restrict__ x, real *__restrict__ y, real *__restrict__ z, size_t ntot,
    const PopCtl *__restrict__ ctl) {
  __shared__ int16_t tile[PADIDX(CPB)];
  __shared__ uint32_t toff[TPB];
  __shared__ uint64_t ws[TPB / 64];
  const size_t b0 = (size_t) blockIdx.x * CPB;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int c = q * TPB + tid;
    const size_t i = b0 + c;
    const int16_t t = nt[i < N ? i : N - 1];
    tile[PADIDX(c)] = (i < N) ? t : (int16_t) 0;
  }
  __syncthreads();
  uint64_t cnt = 0;
#pragma unroll
  for (int q = 0; q < CPT; q++) cnt += (uint64_t) tile[PADIDX(tid * CPT + q)];
  uint64_t tot;
  const uint64_t excl = block_excl_scan<uint64_t>(cnt, tot, ws);
  if (tot == 0) return;                 // uniform over the block
  toff[tid] = (uint32_t) excl;          // tot <= CPB * 1024
  const uint64_t tb = blockofs[blockIdx.x];
  __syncthreads();
  const uint32_t K = (uint32_t) ((tot + TPB - 1) / TPB);
  uint32_t l = (uint32_t) tid * K;
  if (l >= tot) return;
  const uint32_t lend = (l + K < tot) ? l + K : (uint32_t) tot;
  // cell of tracer l: the last thread range starting at or before l, then
  // its cells
  int lo = 0, hi = TPB - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (toff[mid] <= l) lo = mid; else hi = mid - 1;
  }
  int c = lo * CPT;
  uint32_t acc = toff[lo];
  int m = tile[PADIDX(c)];
  while (acc + (uint32_t) m <= l) { acc += (uint32_t) m; c++; m = tile[PADIDX(c)]; }
  int n = (int) (l - acc);
  Draws ds;
  ds.w = mtw;
  ds.d = 3 * (tb + l);                  // MT19937 words start at draw 5N
  // one jump from the stream start per thread (a block-base jump by one
  // thread, a barrier and a jump per thread from it chain five table digits
  // instead of three: 177 -> 172 us at 512^3)
  if (!mtw) {
    ds.g = ctl ? ctl->pos_seed : seed;
    mrg_jump(ds.g, 3 * (tb + l), jtab);
  }
  uint32_t k, j;
  uint32_t ij = fastdiv_qr((uint32_t) (b0 + c), fNg, k);
  uint32_t i = fastdiv_qr(ij, fNg, j);
#pragma unroll 1
  for (;;) {
    double r = ds.next() - 0.5;
    real xx = (real) (((int) i + r) * (double) gs);
    r = ds.next() - 0.5;
    real yy = (real) (((int) j + r) * (double) gs);
    r = ds.next() - 0.5;
    real zz = (real) (((int) k + r) * (double) gs);
    if (xx < 0) xx = (real) ((double) xx + Lbox);
    if (yy < 0) yy = (real) ((double) yy + Lbox);
    if (zz < 0) zz = (real) ((double) zz + Lbox);
    const size_t t = tb + l;
    if (t < ntot) { x[t] = xx; y[t] = yy; z[t] = zz; }
    if (++l >= lend) break;
    if (++n == m) {                     // next occupied cell of the block
      n = 0;
      do { c++; m = tile[PADIDX(c)]; } while (m <= 0);
      ij = fastdiv_qr((uint32_t) (b0 + c), fNg, k);
      i = fastdiv_qr(ij, fNg, j);
    }
  }
}

hipError_t launch_positions(const int16_t *nt, const uint64_t *blockofs,
    int Ng, double Lbox, real gs, MrgState seed, const uint32_t *jtab,
    const uint32_t *mtw, real *x, real *y, real *z, size_t ntot, hipStream_t st,
    const PopCtl *ctl) {
  const size_t N = (size_t) Ng * Ng * Ng;
  if (N >= 0xffffffffull) return hipErrorInvalidValue;
  MrgState s5 = seed;
  if (!ctl) mrg_jump(s5, (uint64_t) 5 * N, mrg_host_jump_table());
  hipLaunchKernelGGL(k_positions, dim3(ez_blocks(N, CPB)), dim3(TPB), 0, st,
      nt, blockofs, fastdiv_make((uint32_t) Ng), Lbox, gs, N, s5, jtab, mtw, x, y, z, ntot,
      ctl);
  return hipGetLastError();
}

// ----------------------------------------------------------- velocities
// One thread = PPT tracer pairs.  For each tracer: CIC gather of the
// momentum fields normalised by rho (disp_cic, pop_tracer.c:680-681, where
// rho > REAL_TOL), weighted by rho (assign_vel), times vfac; then the
// Gaussian scatter of pair q from draws 8N + 6q.
#define PPT 4

// CIC corners of a tracer and their {rho, u} vectors: loads issued here,
// used by corners_vel (so that the loads of several tracers are in flight
// together).
struct Corners {
  real4 f[8];
  real wx0, wx1, wy0, wy1, wz0, wz1;
};

__device__ __forceinline__ void corners_load(const real4 *__restrict__ vel,
    int Ng, real igs, real px, real py, real pz, Corners &k) {
  const real xx = px * igs, yy = py * igs, zz = pz * igs;
  int x0 = (int) xx, y0 = (int) yy, z0 = (int) zz;
  k.wx1 = xx - (real) x0; k.wx0 = 1 - k.wx1;
  k.wy1 = yy - (real) y0; k.wy0 = 1 - k.wy1;
  k.wz1 = zz - (real) z0; k.wz0 = 1 - k.wz1;
  if (x0 == Ng) x0 = 0;
  if (y0 == Ng) y0 = 0;
  if (z0 == Ng) z0 = 0;
  if ((unsigned) x0 >= (unsigned) Ng) x0 = 0;    // only for non-finite input
  if ((unsigned) y0 >= (unsigned) Ng) y0 = 0;
  if ((unsigned) z0 >= (unsigned) Ng) z0 = 0;
  const int x1 = (x0 == Ng - 1) ? 0 : x0 + 1;
  const int y1 = (y0 == Ng - 1) ? 0 : y0 + 1;
  const int z1 = (z0 == Ng - 1) ? 0 : z0 + 1;
  const size_t c[8] = {
    vel_brick_index(x0, y0, z0, Ng), vel_brick_index(x0, y0, z1, Ng),
    vel_brick_index(x0, y1, z0, Ng), vel_brick_index(x0, y1, z1, Ng),
    vel_brick_index(x1, y0, z0, Ng), vel_brick_index(x1, y0, z1, Ng),
    vel_brick_index(x1, y1, z0, Ng), vel_brick_index(x1, y1, z1, Ng)};
  // per corner {rho, u0, u1, u2}: u already normalised by rho (disp_cic)
#pragma unroll
  for (int q = 0; q < 8; q++) k.f[q] = vel[c[q]];
}

__device__ __forceinline__ void corners_vel(const Corners &k, double vfac, real *v) {
  const real4 *f = k.f;
  const real wx0 = k.wx0, wx1 = k.wx1, wy0 = k.wy0, wy1 = k.wy1, wz0 = k.wz0, wz1 = k.wz1;
  real r[8];
  // r_q = rho * wx * wy * wz, evaluated left to right as assign_vel does
  r[0] = f[0].x * wx0 * wy0 * wz0; r[1] = f[1].x * wx0 * wy0 * wz1;
  r[2] = f[2].x * wx0 * wy1 * wz0; r[3] = f[3].x * wx0 * wy1 * wz1;
  r[4] = f[4].x * wx1 * wy0 * wz0; r[5] = f[5].x * wx1 * wy0 * wz1;
  r[6] = f[6].x * wx1 * wy1 * wz0; r[7] = f[7].x * wx1 * wy1 * wz1;
  const real rs = r[0] + r[1] + r[2] + r[3] + r[4] + r[5] + r[6] + r[7];
  real sx = f[0].y * r[0] + f[1].y * r[1] + f[2].y * r[2] + f[3].y * r[3]
          + f[4].y * r[4] + f[5].y * r[5] + f[6].y * r[6] + f[7].y * r[7];
  real sy = f[0].z * r[0] + f[1].z * r[1] + f[2].z * r[2] + f[3].z * r[3]
          + f[4].z * r[4] + f[5].z * r[5] + f[6].z * r[6] + f[7].z * r[7];
  real sz = f[0].w * r[0] + f[1].w * r[1] + f[2].w * r[2] + f[3].w * r[3]
          + f[4].w * r[4] + f[5].w * r[5] + f[6].w * r[6] + f[7].w * r[7];
  const double sc = vfac / (double) rs;
  v[0] = (real) ((double) sx * sc);
  v[1] = (real) ((double) sy * sc);
  v[2] = (real) ((double) sz * sc);
}

__global__ void __launch_bounds__(TPB)
k_velocities(const real *__restrict__ x, const real *__restrict__ y,
    const real *__restrict__ z, const real4 *__restrict__ vel, size_t ntot, size_t natt,
    int Ng, real igs, double vfac, double sigma, uint64_t base_draw, MrgState seed,
    const uint32_t *__restrict__ jtab, const uint32_t *__restrict__ mtw,
    real *__restrict__ vx,
    real *__restrict__ vy, real *__restrict__ vz, const PopCtl *__restrict__ ctl) {
  if (ctl) {                            // ntot: the catalogue capacity
    if (ctl->ntot < ntot) ntot = ctl->ntot;
    seed = ctl->vel_seed;
    sigma = ctl->sigma_v;
  }
  const size_t npair = (ntot + 1) / 2;
  // the XCD-aware order over the blocks the catalogue needs (the grid may
  // be sized by a capacity), so that every XCD gets its share
  const unsigned nact = (unsigned) ((npair + (size_t) TPB * PPT - 1) / ((size_t) TPB * PPT));
  if (blockIdx.x >= nact) return;       // uniform: blocks past the catalogue
  const size_t q0blk = (size_t) xcd_item(blockIdx.x, nact) * TPB * PPT;
  const size_t q0 = q0blk + (size_t) threadIdx.x * PPT;
  if (q0 >= npair) return;
  Draws g;
  g.w = mtw;
  g.d = 6 * q0;                         // MT19937 words start at draw 8N
  if (!mtw) {                           // one jump per thread from the stream
    g.g = seed;
    mrg_jump(g.g, base_draw + 6 * q0, jtab);
  }
#pragma unroll 1
  for (int p = 0; p < PPT; p++) {
    const size_t q = q0 + p;
    if (q >= npair) break;
    const size_t a = 2 * q, b = a + 1;
    const bool hb = b < ntot;             // else: the last, unpaired tracer
    const size_t bb = hb ? b : a;
    // both tracers' corner loads first, then the draws and sums
    Corners ka, kb;
    corners_load(vel, Ng, igs, x[a], y[a], z[a], ka);
    corners_load(vel, Ng, igs, x[bb], y[bb], z[bb], kb);
    real va[3], vb3[3];
    if (a < natt) { va[0] = vx[a]; va[1] = vy[a]; va[2] = vz[a]; }
    else corners_vel(ka, vfac, va);
    double g1, g2;
    gauss_pair(g, sigma, g1, g2);
    va[0] = (real) ((double) va[0] + g1);
    va[1] = (real) ((double) va[1] + g2);
    gauss_pair(g, sigma, g1, g2);
    va[2] = (real) ((double) va[2] + g1);
    vx[a] = va[0]; vy[a] = va[1]; vz[a] = va[2];
    if (hb) {
      if (b < natt) { vb3[0] = vx[b]; vb3[1] = vy[b]; vb3[2] = vz[b]; }
      else corners_vel(kb, vfac, vb3);
      vb3[0] = (real) ((double) vb3[0] + g2);
      gauss_pair(g, sigma, g1, g2);
      vb3[1] = (real) ((double) vb3[1] + g1);
      vb3[2] = (real) ((double) vb3[2] + g2);
      vx[b] = vb3[0]; vy[b] = vb3[1]; vz[b] = vb3[2];
    }
  }
}

hipError_t launch_velocities(const real *x, const real *y, const real *z,
    const real4 *vel, size_t ntot, size_t natt, int Ng, real igs, double vfac, double sigma_v,
    MrgState seed, const uint32_t *jtab, const uint32_t *mtw, real *vx, real *vy,
    real *vz, hipStream_t st, const PopCtl *ctl) {
  const size_t N = (size_t) Ng * Ng * Ng;
  const size_t npair = (ntot + 1) / 2;
  if (npair == 0) return hipSuccess;
  MrgState s8 = seed;                       // draw 8N
  if (!ctl) mrg_jump(s8, (uint64_t) 8 * N, mrg_host_jump_table());
  hipLaunchKernelGGL(k_velocities, dim3(ez_blocks(npair, TPB * PPT)), dim3(TPB),
      0, st, x, y, z, vel, ntot, natt, Ng, igs, vfac, sigma_v,
      (uint64_t) 0, s8, jtab, mtw, vx, vy, vz, ctl);
  return hipGetLastError();
}

// ------------------------------------------- tracers attached to particles
// pop_tracer.c:517-582 (NGP particle counts of the tracer cells) and
// :804-879 (the selection).  Particles are numbered in Lagrangian order
// q = (i Ng + j) Ng + k; each block takes CPB consecutive particles, CPT per
// thread in a contiguous run, so block-scan ranks follow Lagrangian order.

// Wrapped grid coordinate of particle q and its nearest grid point cell
// (pop_tracer.c:536-557), in the reference's operation order.
__device__ __forceinline__ size_t att_ngp(const real *__restrict__ p0,
    const real *__restrict__ p1, const real *__restrict__ p2, size_t q,
    const AttGeom &g, real pc[3]) {
  uint32_t k, j;
  const uint32_t ij = fastdiv_qr((uint32_t) q, g.fNg, k);
  const uint32_t i = fastdiv_qr(ij, g.fNg, j);
  const uint32_t lat[3] = {i, j, k};
  const real ps[3] = {p0[q], p1[q], p2[q]};
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    real x = ps[d] * g.igs + (real) (int) lat[d];
    if (x < 0) x += (real) g.Ng;
    else if (x >= (real) g.Ng) x -= (real) g.Ng;
    pc[d] = x;
    c[d] = (int) ((double) x + 0.5);
    if (c[d] >= g.Ng) c[d] = 0;
    // displacements beyond a box length or not finite (undefined in the
    // reference): keep the index inside the grid
    if ((unsigned) c[d] >= (unsigned) g.Ng) c[d] = ((c[d] % g.Ng) + g.Ng) % g.Ng;
  }
  return ((size_t) c[0] * g.Ng + c[1]) * g.Ng + c[2];
}

// Particles per tracer cell (u32 counters, cleared by the caller); ovf is
// set when a cell would exceed the reference's uint16 counter.
__global__ void __launch_bounds__(TPB)
k_att_count(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, AttGeom g, const int16_t *__restrict__ nt,
    uint32_t *__restrict__ np, uint32_t *__restrict__ ovf) {
  const size_t q0 = (size_t) blockIdx.x * CPB + (size_t) threadIdx.x * CPT;
  for (int u = 0; u < CPT; u++) {
    const size_t q = q0 + u;
    if (q >= g.N) break;
    real pc[3];
    const size_t c = att_ngp(p0, p1, p2, q, g, pc);
    if (nt[c] == 0) continue;
    if (atomicAdd(&np[c], 1u) >= 65535u) *ovf = 1u;
  }
}

// Contested cells (more particles than tracers: the selection draws) per
// block of CPB cells; WRITE: give each its rank cid, record (nt, np) in
// cinfo[cid] and mark np[c] = 0x80000000 | cid.
template <int WRITE>
__global__ void __launch_bounds__(TPB)
k_att_cells(const int16_t *__restrict__ nt, uint32_t *__restrict__ np, size_t N,
    uint32_t *__restrict__ blockcnt, const uint64_t *__restrict__ blockoff,
    uint32_t *__restrict__ cinfo) {
  __shared__ uint32_t ws[TPB / 64];
  const size_t c0 = (size_t) blockIdx.x * CPB + (size_t) threadIdx.x * CPT;
  uint32_t n = 0;
  for (int u = 0; u < CPT; u++) {
    const size_t c = c0 + u;
    if (c < N && nt[c] > 0 && np[c] > (uint32_t) nt[c]) n++;
  }
  uint32_t tot;
  const uint32_t excl = block_excl_scan<uint32_t>(n, tot, ws);
  if (!WRITE) {
    if (threadIdx.x == 0) blockcnt[blockIdx.x] = tot;
    return;
  }
  uint32_t r = (uint32_t) blockoff[blockIdx.x] + excl;
  for (int u = 0; u < CPT; u++) {
    const size_t c = c0 + u;
    if (c < N && nt[c] > 0 && np[c] > (uint32_t) nt[c]) {
      cinfo[r] = (uint32_t) (uint16_t) nt[c] | (np[c] << 16);
      np[c] = 0x80000000u | r;
      r++;
    }
  }
}

// Class of each particle: 0 its cell hosts no tracer, 1 attached without a
// draw (the cell has no more particles than tracers), 2 contested; per
// block the number of contested particles.
__global__ void __launch_bounds__(TPB)
k_att_classify(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, AttGeom g, const int16_t *__restrict__ nt,
    const uint32_t *__restrict__ np, uint8_t *__restrict__ cls,
    uint32_t *__restrict__ blockcnt) {
  __shared__ uint32_t ws[TPB / 64];
  const size_t q0 = (size_t) blockIdx.x * CPB + (size_t) threadIdx.x * CPT;
  uint32_t n = 0;
  for (int u = 0; u < CPT; u++) {
    const size_t q = q0 + u;
    if (q >= g.N) break;
    real pc[3];
    const size_t c = att_ngp(p0, p1, p2, q, g, pc);
    const uint8_t k = nt[c] == 0 ? 0 : ((np[c] & 0x80000000u) ? 2 : 1);
    cls[q] = k;
    n += k == 2;
  }
  uint32_t tot;
  (void) block_excl_scan<uint32_t>(n, tot, ws);
  if (threadIdx.x == 0) blockcnt[blockIdx.x] = tot;
}

// The contested particles in Lagrangian order: plist[rank] = their cell's
// cid.
__global__ void __launch_bounds__(TPB)
k_att_list(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, AttGeom g, const uint32_t *__restrict__ np,
    const uint8_t *__restrict__ cls, const uint64_t *__restrict__ blockoff,
    uint32_t *__restrict__ plist) {
  __shared__ uint32_t ws[TPB / 64];
  const size_t q0 = (size_t) blockIdx.x * CPB + (size_t) threadIdx.x * CPT;
  uint32_t n = 0;
  for (int u = 0; u < CPT; u++)
    if (q0 + u < g.N && cls[q0 + u] == 2) n++;
  uint32_t tot;
  uint64_t r = blockoff[blockIdx.x] + block_excl_scan<uint32_t>(n, tot, ws);
  for (int u = 0; u < CPT; u++) {
    const size_t q = q0 + u;
    if (q >= g.N || cls[q] != 2) continue;
    real pc[3];
    const size_t c = att_ngp(p0, p1, p2, q, g, pc);
    plist[r++] = np[c] & 0x7fffffffu;
  }
}

// Selected particles (class 1, or class 2 with its selection flag) per
// block; WRITE: the attached tracers at their rank: position = wrapped grid
// coordinate x gs, velocity = displacement x vfac (pop_tracer.c:856-863).
template <int WRITE>
__global__ void __launch_bounds__(TPB)
k_att_select(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, AttGeom g, double vfac, const uint8_t *__restrict__ cls,
    const uint64_t *__restrict__ contoff, const uint8_t *__restrict__ sel,
    uint32_t *__restrict__ blockcnt, const uint64_t *__restrict__ blockoff,
    real *__restrict__ x, real *__restrict__ y, real *__restrict__ z,
    real *__restrict__ vx, real *__restrict__ vy, real *__restrict__ vz) {
  __shared__ uint32_t ws[TPB / 64];
  const size_t q0 = (size_t) blockIdx.x * CPB + (size_t) threadIdx.x * CPT;
  uint32_t nc = 0;
  for (int u = 0; u < CPT; u++)
    if (q0 + u < g.N && cls[q0 + u] == 2) nc++;
  uint32_t tot;
  uint64_t rc = contoff[blockIdx.x] + block_excl_scan<uint32_t>(nc, tot, ws);
  uint32_t take = 0;                    // bit u: particle q0 + u is selected
  for (int u = 0; u < CPT; u++) {
    const size_t q = q0 + u;
    if (q >= g.N) break;
    const uint8_t k = cls[q];
    if (k == 1 || (k == 2 && sel[rc])) take |= 1u << u;
    if (k == 2) rc++;
  }
  uint64_t r = block_excl_scan<uint32_t>((uint32_t) __popc(take), tot, ws);
  if (!WRITE) {
    if (threadIdx.x == 0) blockcnt[blockIdx.x] = tot;
    return;
  }
  r += blockoff[blockIdx.x];
  for (int u = 0; u < CPT; u++) {
    if (!((take >> u) & 1u)) continue;
    const size_t q = q0 + u;
    real pc[3];
    (void) att_ngp(p0, p1, p2, q, g, pc);
    x[r] = pc[0] * g.gs;
    y[r] = pc[1] * g.gs;
    z[r] = pc[2] * g.gs;
    vx[r] = (real) ((double) p0[q] * vfac);
    vy[r] = (real) ((double) p1[q] * vfac);
    vz[r] = (real) ((double) p2[q] * vfac);
    r++;
  }
}

// Tracers per cell left for uniform positions: max(nt - np, 0), i.e. nt - np
// in uncontested tracer cells and 0 in contested ones; per-block sums.
__global__ void __launch_bounds__(TPB)
k_att_rest(const int16_t *__restrict__ nt, const uint32_t *__restrict__ np, size_t N,
    int16_t *__restrict__ rest, uint32_t *__restrict__ blockcnt) {
  __shared__ uint32_t ws[TPB / 64];
  const size_t c0 = (size_t) blockIdx.x * CPB + (size_t) threadIdx.x * CPT;
  uint32_t n = 0;
  for (int u = 0; u < CPT; u++) {
    const size_t c = c0 + u;
    if (c >= N) break;
    const int t = nt[c];
    const uint32_t p = np[c];
    const int r = (t > 0 && !(p & 0x80000000u)) ? t - (int) p : 0;
    rest[c] = (int16_t) r;
    n += (uint32_t) r;
  }
  uint32_t tot;
  (void) block_excl_scan<uint32_t>(n, tot, ws);
  if (threadIdx.x == 0) blockcnt[blockIdx.x] = tot;
}

hipError_t launch_att_count(const real *p0, const real *p1, const real *p2,
    const AttGeom &g, const int16_t *nt, uint32_t *np, uint32_t *ovf, hipStream_t st) {
  hipLaunchKernelGGL(k_att_count, dim3(ez_blocks(g.N, CPB)), dim3(TPB), 0, st, p0, p1, p2,
      g, nt, np, ovf);
  return hipGetLastError();
}

hipError_t launch_att_cells(const int16_t *nt, uint32_t *np, size_t N, uint32_t *blockcnt,
    const uint64_t *blockoff, uint32_t *cinfo, hipStream_t st) {
  if (!blockoff)
    hipLaunchKernelGGL(k_att_cells<0>, dim3(ez_blocks(N, CPB)), dim3(TPB), 0, st, nt, np, N,
        blockcnt, blockoff, cinfo);
  else
    hipLaunchKernelGGL(k_att_cells<1>, dim3(ez_blocks(N, CPB)), dim3(TPB), 0, st, nt, np, N,
        blockcnt, blockoff, cinfo);
  return hipGetLastError();
}

hipError_t launch_att_classify(const real *p0, const real *p1, const real *p2,
    const AttGeom &g, const int16_t *nt, const uint32_t *np, uint8_t *cls,
    uint32_t *blockcnt, hipStream_t st) {
  hipLaunchKernelGGL(k_att_classify, dim3(ez_blocks(g.N, CPB)), dim3(TPB), 0, st, p0, p1, p2,
      g, nt, np, cls, blockcnt);
  return hipGetLastError();
}

hipError_t launch_att_list(const real *p0, const real *p1, const real *p2,
    const AttGeom &g, const uint32_t *np, const uint8_t *cls, const uint64_t *blockoff,
    uint32_t *plist, hipStream_t st) {
  hipLaunchKernelGGL(k_att_list, dim3(ez_blocks(g.N, CPB)), dim3(TPB), 0, st, p0, p1, p2,
      g, np, cls, blockoff, plist);
  return hipGetLastError();
}

hipError_t launch_att_select(const real *p0, const real *p1, const real *p2,
    const AttGeom &g, double vfac, const uint8_t *cls, const uint64_t *contoff,
    const uint8_t *sel, uint32_t *blockcnt, const uint64_t *blockoff, real *const cat[6],
    hipStream_t st) {
  if (!blockoff)
    hipLaunchKernelGGL(k_att_select<0>, dim3(ez_blocks(g.N, CPB)), dim3(TPB), 0, st, p0, p1,
        p2, g, vfac, cls, contoff, sel, blockcnt, blockoff, cat[0], cat[1], cat[2], cat[3],
        cat[4], cat[5]);
  else
    hipLaunchKernelGGL(k_att_select<1>, dim3(ez_blocks(g.N, CPB)), dim3(TPB), 0, st, p0, p1,
        p2, g, vfac, cls, contoff, sel, blockcnt, blockoff, cat[0], cat[1], cat[2], cat[3],
        cat[4], cat[5]);
  return hipGetLastError();
}

hipError_t launch_att_rest(const int16_t *nt, const uint32_t *np, size_t N, int16_t *rest,
    uint32_t *blockcnt, hipStream_t st) {
  hipLaunchKernelGGL(k_att_rest, dim3(ez_blocks(N, CPB)), dim3(TPB), 0, st, nt, np, N, rest,
      blockcnt);
  return hipGetLastError();
}

// -------------------------------------------------------- test helpers
__global__ void k_uniform(MrgState seed, uint64_t start, size_t n,
    const uint32_t *__restrict__ jtab, double *__restrict__ out) {
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  const size_t i0 = t * 64;
  if (i0 >= n) return;
  MrgState g = seed;
  mrg_jump(g, start + i0, jtab);
  for (size_t i = i0; i < i0 + 64 && i < n; i++) out[i] = mrg_next(g);
}

__global__ void k_uniform_mt(const uint32_t *__restrict__ w, size_t n, int pos,
    double *__restrict__ out) {
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = pos ? mt_draw_pos(w, i) : mt_draw(w, i);
}

hipError_t launch_uniform_mt(const uint32_t *w, size_t n, int pos, double *out,
    hipStream_t st) {
  hipLaunchKernelGGL(k_uniform_mt, dim3(ez_blocks(n, 256)), dim3(256), 0, st, w, n,
      pos, out);
  return hipGetLastError();
}

hipError_t launch_uniform(MrgState seed, uint64_t start, size_t n,
    const uint32_t *jtab, double *out, hipStream_t st) {
  const size_t nthr = (n + 63) / 64;
  hipLaunchKernelGGL(k_uniform, dim3(ez_blocks(nthr, 256)), dim3(256), 0, st,
      seed, start, n, jtab, out);
  return hipGetLastError();
}

// save_res.c:147-158 on the device.
__global__ void k_rsd(const real *__restrict__ z, const real *__restrict__ vz,
    size_t n, double fac, double Lbox, double *__restrict__ out) {
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double zr = (double) z[i] + fac * (double) vz[i];
  if (zr >= Lbox) zr -= Lbox;
  if (zr < 0) zr += Lbox;
  out[i] = zr;
}

hipError_t launch_rsd(const real *z, const real *vz, size_t n, double fac,
    double Lbox, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_rsd, dim3(ez_blocks(n, 256)), dim3(256), 0, st, z, vz,
      n, fac, Lbox, out);
  return hipGetLastError();
}
