// field.hip -- initial-condition spectrum and CIC density on gfx950.
//
//   k_pk_table    P(k) at every integer |n|^2 of the grid (pk_interp of
//                 linear_pk.c:194-202 evaluated once per distinct |k|).
//   k_gen_rhok    noise + Zel'dovich spectrum, perturb.c:78-359.
// (CIC painting is in cic.hip.)  Compiled with -ffp-contract=off.
#include "common.h"
#include "rng.h"
#include "kernels.h"
#include "gen.h"

// ------------------------------------------------------------ P(k) table
// linear_pk.c:36-51 (bin search) and math/cspline.c:91-105 (evaluation).
__device__ __forceinline__ double pk_eval(const double *__restrict__ kt,
    const double *__restrict__ Pt, const double *__restrict__ P2t, int n,
    double v) {
  int l = 0, u = n - 1, idx = -1;
  while (l <= u) {
    const int i = (l + u) >> 1;
    if (i >= n - 1) { idx = (kt[n - 1] == v) ? n - 1 : -1; break; }
    if (kt[i + 1] <= v) l = i + 1;
    else if (kt[i] > v) u = i - 1;
    else { idx = i; break; }
  }
  if (idx == n - 1) return Pt[idx];
  if (idx < 0) {
    if (v >= kt[n - 1]) return Pt[n - 1];
    return Pt[0];
  }
  const int j = idx + 1;
  const double h = kt[j] - kt[idx];
  const double dl = v - kt[idx];
  const double dr = kt[j] - v;
  const double h2 = h * h;
  const double lin = dl * Pt[j] + dr * Pt[idx];
  const double cr = (dl * dl - h2) * dl * P2t[j];
  const double cl = (dr * dr - h2) * dr * P2t[idx];
  return (lin + 0x1.5555555555555p-3 * (cr + cl)) / h;
}

__global__ void k_pk_table(double *__restrict__ tab, int nsq,
    const double *__restrict__ kt, const double *__restrict__ Pt,
    const double *__restrict__ P2t, int n, int logpk, double kfac,
    double logkfac) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsq) return;
  if (q == 0) { tab[0] = 0; return; }
  const double ksq = (double) q;
  const double kmod = logpk ? log(ksq) * 0.5 + logkfac : kfac * sqrt(ksq);
  tab[q] = pk_eval(kt, Pt, P2t, n, kmod);
}

// ------------------------------------------------------------ spectrum
// One thread = CH consecutive cells of the half-complex grid (reference
// order, k fastest, Ngk = Ng/2 + 1 per row; Ng even).  Cell idx >= 1
// consumes draws (idx-1)*2 (amplitude, then phase) or (idx-1) with fixed
// amplitudes: MRG32k3a by jump-ahead per thread, MT19937 (MT) from the
// stream words generated beforehand (mtw, two words per draw).  Output: rhok[i][j][kk < Ng/2] and rhok_ny[i][j] (the
// kz = Ng/2 plane); optional copy in the reference layout.
// The block's 256 CH cells map to one contiguous range of rhok (minus the
// Nyquist cells): values are staged in LDS and stored by consecutive lanes
// to consecutive addresses.  Mirror writes of the kz = 0 and Nyquist planes
// (2 cells in Ngk) go straight to memory.
#define GEN_TPB 256
#define GEN_SKIP 0xffffu

#define GEN_CHUNK (sizeof(real) == 4 ? 16 : 8)     // cells per thread

// MRG32k3a state at the first draw of each k_gen_rhok workgroup, one thread
// per workgroup, so that no workgroup waits on a serial jump of its own.
__global__ void __launch_bounds__(256)
k_gen_base(MrgState seed, uint64_t mult, size_t nblk, size_t cells_per_blk,
    const uint32_t *__restrict__ jtab, MrgState *__restrict__ base,
    const MrgState *__restrict__ seedp) {
  const size_t b = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const size_t blk0 = b * cells_per_blk;
  MrgState g = seedp ? *seedp : seed;
  mrg_jump(g, (blk0 == 0 ? 0 : blk0 - 1) * mult, jtab);
  base[b] = g;
}

template <int CH>
__global__ void __launch_bounds__(GEN_TPB)
k_gen_rhok(cplx *__restrict__ rhok, cplx *__restrict__ rhok_ny,
    real *__restrict__ ref_copy, GenParams gp, const double *__restrict__ ptab,
    const uint32_t *__restrict__ jtab, const MrgState *__restrict__ bases) {
  constexpr int NB = GEN_TPB * CH;
  __shared__ cplx sv[NB + NB / CH];             // one pad slot per thread
  __shared__ uint16_t so[NB];
  const int Ng = gp.Ng, Ngh = Ng >> 1, Ngk = Ngh + 1;
  const size_t blk0 = (size_t) blockIdx.x * NB;
  // rhok index of a cell k < Ngh: row * Ngh + k = idx - row
  const size_t obase = blk0 - blk0 / (size_t) Ngk;
  const size_t idx0 = blk0 + (size_t) threadIdx.x * CH;
  size_t row = idx0 / (size_t) Ngk;
  int k = (int) (idx0 - row * (size_t) Ngk);
  int i = (int) (row / (size_t) Ng);
  int j = (int) (row - (size_t) i * Ng);
  const uint64_t mult = gp.fixamp ? 1 : 2;
  const uint64_t d = (idx0 == 0 ? 0 : (idx0 - 1)) * mult;   // first draw
  // the workgroup's MRG32k3a state at the draw of its first cell (from
  // k_gen_base), then a short jump per thread (< 2 base-256 digits)
  const uint64_t dblk = (blk0 == 0 ? 0 : (blk0 - 1)) * mult;
  MrgState g = bases[blockIdx.x];
  if (idx0 < gp.Nh) mrg_jump(g, d - dblk, jtab);
  for (int c = 0; c < CH; c++) {
    const size_t idx = idx0 + c;
    const int e = threadIdx.x * CH + c;
    uint16_t off = GEN_SKIP;
    if (idx < gp.Nh) {
      GenCell gc{0, 0, true, false, 0, 0};
      if (idx != 0) {
        const double amp = gp.fixamp ? 1.0 : mrg_next(g);
        const double ph = mrg_next(g);
        gc = gen_cell(i, j, k, amp, ph, gp, ptab);
      }
      if (gc.write) {
        const cplx v = {gc.re, gc.im};
        if (k < Ngh) {
          sv[e + e / CH] = v;
          off = (uint16_t) (idx - row - obase);
        }
        else rhok_ny[row] = v;
        if (ref_copy) { ref_copy[2 * idx] = gc.re; ref_copy[2 * idx + 1] = gc.im; }
        gen_mirror(gc, i, j, k, gp, rhok, rhok_ny, ref_copy);
      }
    }
    so[e] = off;
    if (++k == Ngk) {
      k = 0; row++;
      if (++j == Ng) { j = 0; i++; }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NB; e += GEN_TPB) {
    const uint16_t o = so[e];
    if (o != GEN_SKIP) rhok[obase + o] = sv[e + e / CH];
  }
}

// MT19937: the draws come from a buffer of stream words, so a thread's
// cells need not be consecutive.  Lane-consecutive cells make the word
// reads and the spectrum stores coalesced, with no LDS staging (per-thread
// runs of 16 cells read each word load as 64 separate lines).
template <int CH>
__global__ void __launch_bounds__(GEN_TPB)
k_gen_rhok_mt(cplx *__restrict__ rhok, cplx *__restrict__ rhok_ny,
    real *__restrict__ ref_copy, GenParams gp, const double *__restrict__ ptab,
    const uint32_t *__restrict__ mtw) {
  const int Ng = gp.Ng, Ngh = Ng >> 1, Ngk = Ngh + 1;
  const uint64_t mult = gp.fixamp ? 1 : 2;
  const size_t blk0 = (size_t) blockIdx.x * GEN_TPB * CH;
  for (int c = 0; c < CH; c++) {
    const size_t idx = blk0 + (size_t) c * GEN_TPB + threadIdx.x;
    if (idx >= gp.Nh) break;
    const size_t row = idx / (size_t) Ngk;
    const int k = (int) (idx - row * (size_t) Ngk);
    const int i = (int) (row / (size_t) Ng);
    const int j = (int) (row - (size_t) i * Ng);
    GenCell gc{0, 0, true, false, 0, 0};
    if (idx != 0) {
      uint64_t d = (idx - 1) * mult;
      const double amp = gp.fixamp ? 1.0 : mt_draw_pos(mtw, d++);
      const double ph = mt_draw(mtw, d);
      gc = gen_cell(i, j, k, amp, ph, gp, ptab);
    }
    if (!gc.write) continue;
    const cplx v = {gc.re, gc.im};
    if (k < Ngh) rhok[idx - row] = v;
    else rhok_ny[row] = v;
    if (ref_copy) { ref_copy[2 * idx] = gc.re; ref_copy[2 * idx + 1] = gc.im; }
    gen_mirror(gc, i, j, k, gp, rhok, rhok_ny, ref_copy);
  }
}

size_t gen_base_count(size_t Nh) {
  constexpr int CH = GEN_CHUNK;
  return ez_blocks((Nh + CH - 1) / CH, GEN_TPB);
}

hipError_t launch_gen_rhok(cplx *rhok, cplx *rhok_ny, real *ref_copy,
    const GenParams &gp, const double *ptab, const uint32_t *jtab,
    const uint32_t *mtw, MrgState *bases, hipStream_t st, const MrgState *seedp) {
  if (gp.Ng & 1) return hipErrorInvalidValue;
  constexpr int CH = GEN_CHUNK;
  const size_t nthr = (gp.Nh + CH - 1) / CH;
  const size_t nblk = ez_blocks(nthr, GEN_TPB);
  if (mtw)
    hipLaunchKernelGGL((k_gen_rhok_mt<CH>), dim3(nblk), dim3(GEN_TPB), 0, st,
        rhok, rhok_ny, ref_copy, gp, ptab, mtw);
  else {
    if (!bases) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gen_base, dim3(ez_blocks(nblk, 256)), dim3(256), 0, st,
        gp.seed_state, (uint64_t) (gp.fixamp ? 1 : 2), nblk, (size_t) GEN_TPB * CH, jtab,
        bases, seedp);
    hipLaunchKernelGGL((k_gen_rhok<CH>), dim3(nblk), dim3(GEN_TPB), 0,
        st, rhok, rhok_ny, ref_copy, gp, ptab, jtab, (const MrgState *) bases);
  }
  return hipGetLastError();
}

// The kz = 0 and Nyquist planes of the spectrum only (rhok[i][j][0] and
// rhok_ny), one thread per (i, j), each cell from its own draws by a
// jump-ahead: the rest of the spectrum is generated by the first FFT pass
// itself (fft_first_pass3_gen), which reads these two planes through the
// packed kz = 0 slots.  seedp (optional): the seed in device memory.
__global__ void __launch_bounds__(256)
k_gen_planes(cplx *__restrict__ rhok, cplx *__restrict__ rhok_ny, GenParams gp,
    const double *__restrict__ ptab, const uint32_t *__restrict__ jtab,
    const MrgState *__restrict__ seedp) {
  const int Ng = gp.Ng, Ngh = Ng >> 1, Ngk = Ngh + 1;
  const size_t row = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= (size_t) Ng * Ng) return;
  const int i = (int) (row / (size_t) Ng), j = (int) (row % (size_t) Ng);
  const uint64_t mult = gp.fixamp ? 1 : 2;
  const MrgState seed = seedp ? *seedp : gp.seed_state;
#pragma unroll 1
  for (int h = 0; h < 2; h++) {
    const int k = h ? Ngh : 0;
    const size_t idx = row * (size_t) Ngk + k;
    GenCell gc{0, 0, true, false, 0, 0};
    if (idx != 0) {
      MrgState g = seed;
      mrg_jump(g, (idx - 1) * mult, jtab);
      const double amp = gp.fixamp ? 1.0 : mrg_next(g);
      const double ph = mrg_next(g);
      gc = gen_cell(i, j, k, amp, ph, gp, ptab);
    }
    if (!gc.write) continue;
    const cplx v = {gc.re, gc.im};
    if (h) rhok_ny[row] = v;
    else rhok[row * (size_t) Ngh] = v;
    gen_mirror(gc, i, j, k, gp, rhok, rhok_ny, nullptr);
  }
}

hipError_t launch_gen_planes(cplx *rhok, cplx *rhok_ny, const GenParams &gp,
    const double *ptab, const uint32_t *jtab, hipStream_t st, const MrgState *seedp) {
  if (gp.Ng & 1) return hipErrorInvalidValue;
  const size_t n = (size_t) gp.Ng * gp.Ng;
  hipLaunchKernelGGL(k_gen_planes, dim3(ez_blocks(n, 256)), dim3(256), 0, st,
      rhok, rhok_ny, gp, ptab, jtab, seedp);
  return hipGetLastError();
}

hipError_t launch_pk_table(double *tab, int nsq, const double *kt,
    const double *Pt, const double *P2t, int n, int logpk, double kfac,
    double logkfac, hipStream_t st) {
  hipLaunchKernelGGL(k_pk_table, dim3(ez_blocks(nsq, 256)), dim3(256), 0, st,
      tab, nsq, kt, Pt, P2t, n, logpk, kfac, logkfac);
  return hipGetLastError();
}

// ----------------------------------------------------------- utilities
__global__ void k_fill(uint32_t *__restrict__ p, uint32_t v, size_t n) {
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  const size_t s = (size_t) gridDim.x * blockDim.x;
  for (size_t q = i; q < n; q += s) p[q] = v;
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t st) {
  // bytes is a multiple of 4 for every buffer of the engine
  const size_t n = bytes / 4;
  unsigned nb = ez_blocks(n, 256);
  if (nb > 65536) nb = 65536;
  if (nb == 0) nb = 1;
  hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, st, (uint32_t *) p, 0u, n);
  return hipGetLastError();
}

// White-noise path, perturb.c:456-554: rhok = i sqrt(P) L^-1.5 / (k_f |n|^2)
// * delta_k with delta_k from the r2c of the input field.  Input: the r2c
// result in the packed layout (src[i][j][kk], kk = 0 slot = X0 + i XNy
// after unpacking by k_unpack_r2c); output in the generator's layout.
__global__ void k_gen_rhok_wn(cplx *__restrict__ rhok, cplx *__restrict__ rhok_ny,
    const cplx *__restrict__ dk, const cplx *__restrict__ dk_ny, GenParams gp,
    const double *__restrict__ ptab, real *__restrict__ ref_copy) {
  const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= gp.Nh) return;
  const int Ng = gp.Ng, Ngh = Ng >> 1, Ngk = Ngh + 1;
  const size_t plane = (size_t) Ng * Ngk;
  const int i = (int) (idx / plane);
  const size_t rem = idx - (size_t) i * plane;
  const int j = (int) (rem / Ngk);
  const int k = (int) (rem - (size_t) j * Ngk);
  const size_t ij = (size_t) i * Ng + j;
  cplx d = (k < Ngh) ? dk[ij * Ngh + k] : dk_ny[ij];
  real re = 0, im = 0;
  if (idx != 0) {
    const int ki = (i <= Ngh) ? i : i - Ng;
    const int kj = (j <= Ngh) ? j : j - Ng;
    const int ksqi = ki * ki + kj * kj + k * k;
    double P = ptab[ksqi];
    P = gp.logpk ? exp(P * 0.5) : sqrt(P);
    P *= gp.Pfac / ((double) ksqi * gp.kfac);
    re = (real) (-P * d.y);
    im = (real) (P * d.x);
  }
  const cplx v = {re, im};
  if (k < Ngh) rhok[ij * Ngh + k] = v;
  else rhok_ny[ij] = v;
  if (ref_copy) { ref_copy[2 * idx] = re; ref_copy[2 * idx + 1] = im; }
}

hipError_t launch_gen_rhok_wn(cplx *rhok, cplx *rhok_ny, const cplx *dk,
    const cplx *dk_ny, const GenParams &gp, const double *ptab,
    real *ref_copy, hipStream_t st) {
  hipLaunchKernelGGL(k_gen_rhok_wn, dim3(ez_blocks(gp.Nh, 256)), dim3(256), 0,
      st, rhok, rhok_ny, dk, dk_ny, gp, ptab, ref_copy);
  return hipGetLastError();
}

// After the forward (i, j) transforms of a packed r2c field, slot kk = 0
// holds A + i B with A = FFT2(x0), B = FFT2(xNy) of two real planes.
// A(k) = (P(k) + conj(P(-k)))/2, B(k) = (P(k) - conj(P(-k)))/(2i).
__global__ void k_unpack_r2c(cplx *__restrict__ buf, cplx *__restrict__ ny,
    int Ng) {
  const size_t ij = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  const size_t plane = (size_t) Ng * Ng;
  if (ij >= plane) return;
  const int Ngh = Ng >> 1;
  const int i = (int) (ij / Ng), j = (int) (ij % Ng);
  const int ni = (i == 0) ? 0 : Ng - i, nj = (j == 0) ? 0 : Ng - j;
  const cplx p = buf[ij * Ngh];
  const cplx q = buf[((size_t) ni * Ng + nj) * Ngh];
  const cplx a = {(real) 0.5 * (p.x + q.x), (real) 0.5 * (p.y - q.y)};
  const cplx b = {(real) 0.5 * (p.y + q.y), (real) 0.5 * (q.x - p.x)};
  ny[ij] = b;
  // in-place write of A must not race with the read of the mirror: the
  // caller passes a separate output plane for A through `ny + plane`.
  ny[plane + ij] = a;
}

__global__ void k_store_slot0(cplx *__restrict__ buf, const cplx *__restrict__ a,
    int Ng) {
  const size_t ij = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (ij >= (size_t) Ng * Ng) return;
  buf[ij * (Ng >> 1)] = a[ij];
}

hipError_t launch_unpack_r2c(cplx *buf, cplx *ny2, int Ng, hipStream_t st) {
  const size_t plane = (size_t) Ng * Ng;
  hipLaunchKernelGGL(k_unpack_r2c, dim3(ez_blocks(plane, 256)), dim3(256), 0,
      st, buf, ny2, Ng);
  hipLaunchKernelGGL(k_store_slot0, dim3(ez_blocks(plane, 256)), dim3(256), 0,
      st, buf, ny2 + plane, Ng);
  return hipGetLastError();
}

__global__ void k_scale(real *__restrict__ p, real s, size_t n) {
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] *= s;
}

hipError_t launch_scale(real *p, real s, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_scale, dim3(ez_blocks(n, 256)), dim3(256), 0, st, p, s, n);
  return hipGetLastError();
}

// ------------------------------------------------ plane-wave FFT checks
// Largest |out - expected| and |expected| over a real field that a c2r
// transform produced from a spectrum of a few modes (i0, j0, kk0) with
// 0 < kk0 < Ng/2: expected(x, y, z) = sum_m 2 w_m Re(A_m e^{i theta_m}),
// theta_m = 2 pi ((i0 x + j0 y + kk0 z) mod Ng) / Ng, evaluated in double
// with the phase reduced exactly in integers.  res[0], res[1]: bit patterns
// of the two maxima (non-negative doubles order as unsigned integers);
// cleared by the caller.  One workgroup per (x, y) row.
__global__ void __launch_bounds__(256)
k_plane_wave_check(const real *__restrict__ out, int Ng, FftModes md,
    unsigned long long *__restrict__ res) {
  const size_t row = blockIdx.x;
  const int x = (int) (row / (size_t) Ng), y = (int) (row % (size_t) Ng);
  double emax = 0, rmax = 0;
  for (int z = threadIdx.x; z < Ng; z += blockDim.x) {
    double e = 0;
    for (int m = 0; m < md.n; m++) {
      const long long ph = ((long long) md.ijk[m][0] * x + (long long) md.ijk[m][1] * y +
                            (long long) md.ijk[m][2] * z) % Ng;
      double sn, cs;
      sincos(2 * EZ_PI * (double) ph / Ng, &sn, &cs);
      e += 2 * md.w[m] * (md.re[m] * cs - md.im[m] * sn);
    }
    const double d = fabs((double) out[row * Ng + z] - e);
    emax = d > emax ? d : emax;
    rmax = fabs(e) > rmax ? fabs(e) : rmax;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double a = __shfl_xor(emax, o, 64), b = __shfl_xor(rmax, o, 64);
    emax = a > emax ? a : emax;
    rmax = b > rmax ? b : rmax;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&res[0], (unsigned long long) __double_as_longlong(emax));
    atomicMax(&res[1], (unsigned long long) __double_as_longlong(rmax));
  }
}

hipError_t launch_plane_wave_check(const real *out, int Ng, const FftModes &md,
    unsigned long long *res, hipStream_t st) {
  hipLaunchKernelGGL(k_plane_wave_check, dim3((unsigned) ((size_t) Ng * Ng)), dim3(256), 0, st,
      out, Ng, md, res);
  return hipGetLastError();
}
