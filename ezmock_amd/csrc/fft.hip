// fft.hip -- 3-D real FFTs for gfx950, replacing the reference's FFTW plans
// (fftw_def.h:29-64; c2r executed at perturb.c:361,397,398,556,592,593,
// r2c at dens_field.c:197-201).
//
// Data layout (device, complex `cplx`):  C[i][j][kk], kk in [0, Ng/2), i.e.
// exactly Ng^3 reals -- the same bytes as the real-space output, so every
// transform runs in place in the output buffer.  The kz = 0 and kz = Ng/2
// planes of the half spectrum are both Hermitian in (i, j) after the
// c2r's implicit projection, so they are packed into the kk = 0 slot as
// P = H(kz=0) + i H(kz=Ny); after the (i, j) transforms the real part of the
// slot is the kz=0 column and the imaginary part the Nyquist column.
//
// Passes (each reads and writes the field once, HBM-bound):
//   strided pass, axis i or j: one workgroup = TK consecutive kk columns
//     (128-byte rows) x all L points, staged in LDS; Stockham radix-8/4/2/3
//     stages in LDS with register staging; optionally fused with the
//     spectrum multiply of the first pass (MODE 1).
//   row pass (c2r): TR contiguous rows; Hermitian pre-processing
//     Z[m] = X[m] + conj(X[L/2-m]) + i W^m (X[m] - conj(X[L/2-m])), an L/2
//     complex FFT, and the row of L reals is the complex result.
#include <map>
#include <mutex>
#include <utility>

#include "common.h"
#include "fft.h"
#include "rng.h"
#include "gen.h"

// ------------------------------------------------------------ radix plan
__host__ __device__ constexpr int pow2_exp(int L) {
  int e = 0;
  while (L % 2 == 0) { L /= 2; e++; }
  return e;
}
__host__ __device__ constexpr int pow3_exp(int L) {
  int e = 0;
  while (L % 3 == 0) { L /= 3; e++; }
  return e;
}
// Radix sequence: radix-8 stages, the power-of-two remainder as 4 or 4*4,
// then radix-3 stages.
__host__ __device__ constexpr int n_stages(int L) {
  // rem 1: one radix-8 stage becomes 4*4 (or a lone radix-2); rem 2: a 4.
  const int e = pow2_exp(L), n8 = e / 3, rem = e % 3;
  return n8 + (rem != 0 ? 1 : 0) + pow3_exp(L);
}
__host__ __device__ constexpr int radix_at(int L, int s) {
  int e = pow2_exp(L), n8 = e / 3, rem = e % 3;
  int seq[40] = {0};
  int n = 0;
  if (rem == 1 && n8 > 0) {
    for (int i = 0; i < n8 - 1; i++) seq[n++] = 8;
    seq[n++] = 4; seq[n++] = 4;
  }
  else {
    for (int i = 0; i < n8; i++) seq[n++] = 8;
    if (rem == 1) seq[n++] = 2;
    if (rem == 2) seq[n++] = 4;
  }
  for (int i = 0; i < pow3_exp(L); i++) seq[n++] = 3;
  return seq[s];
}
__host__ __device__ constexpr int prefix_prod(int L, int s) {
  int p = 1;
  for (int i = 0; i < s; i++) p *= radix_at(L, i);
  return p;
}

// ------------------------------------------------------- complex helpers
template <class C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <class C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <class C> __device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by SIGN * i
template <class C, int SIGN> __device__ __forceinline__ C cmuli(C a) {
  return (SIGN > 0) ? C{-a.y, a.x} : C{a.y, -a.x};
}
template <class C> __device__ __forceinline__ C cconj(C a) { return C{a.x, -a.y}; }

// In-register DFT of size R, X[q] = sum_t x[t] exp(SIGN 2 pi i q t / R).
template <class C, int R, int SIGN> struct Dft;

template <class C, int SIGN> struct Dft<C, 2, SIGN> {
  __device__ __forceinline__ static void run(C *x) {
    C a = x[0], b = x[1];
    x[0] = cadd(a, b); x[1] = csub(a, b);
  }
};
template <class C, int SIGN> struct Dft<C, 4, SIGN> {
  __device__ __forceinline__ static void run(C *x) {
    C a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    C c = cadd(x[1], x[3]), d = cmuli<C, SIGN>(csub(x[1], x[3]));
    x[0] = cadd(a, c); x[2] = csub(a, c);
    x[1] = cadd(b, d); x[3] = csub(b, d);
  }
};
template <class C, int SIGN> struct Dft<C, 8, SIGN> {
  __device__ __forceinline__ static void run(C *x) {
    using R = decltype(x[0].x);
    const R h = (R) 0.70710678118654752440;
    C e[4] = {x[0], x[2], x[4], x[6]}, o[4] = {x[1], x[3], x[5], x[7]};
    Dft<C, 4, SIGN>::run(e);
    Dft<C, 4, SIGN>::run(o);
    // twiddles w8^(SIGN q), q = 1..3
    C o1 = cmul(o[1], C{h, (R) SIGN * h});
    C o2 = cmuli<C, SIGN>(o[2]);
    C o3 = cmul(o[3], C{-h, (R) SIGN * h});
    x[0] = cadd(e[0], o[0]); x[4] = csub(e[0], o[0]);
    x[1] = cadd(e[1], o1);   x[5] = csub(e[1], o1);
    x[2] = cadd(e[2], o2);   x[6] = csub(e[2], o2);
    x[3] = cadd(e[3], o3);   x[7] = csub(e[3], o3);
  }
};
template <class C, int SIGN> struct Dft<C, 3, SIGN> {
  __device__ __forceinline__ static void run(C *x) {
    using R = decltype(x[0].x);
    const R s3 = (R) 0.86602540378443864676;
    C a = x[0], b = x[1], c = x[2];
    C t = cadd(b, c);
    C m = C{a.x - (R) 0.5 * t.x, a.y - (R) 0.5 * t.y};
    C d = csub(b, c);
    C r = cmuli<C, SIGN>(C{s3 * d.x, s3 * d.y});
    x[0] = cadd(a, t);
    x[1] = cadd(m, r);
    x[2] = csub(m, r);
  }
};

// ----------------------------------------------------------- LDS stages
// NCOL transforms of length L; element (n, col) at lds[n*SN + col*SC].
// COLFAST: consecutive threads take consecutive columns (SC == 1 layouts).
template <class C, int L, int NCOL, int NT, int SIGN, int SN, int SC,
          bool COLFAST, int S>
__device__ __forceinline__ void lds_stage(C *lds, const C *__restrict__ tw) {
  constexpr int R = radix_at(L, S);
  constexpr int Ls = prefix_prod(L, S);
  constexpr int r = L / Ls, rp = r / R;
  constexpr int NBC = L / R;
  constexpr int NB = NCOL * NBC;
  constexpr int IT = (NB + NT - 1) / NT;
  const int tid = threadIdx.x;
  C v[IT][R];
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int b = tid + it * NT;
    if (IT * NT == NB || b < NB) {
      const int col = COLFAST ? b % NCOL : b / NBC;
      const int bb = COLFAST ? b / NCOL : b % NBC;
      const int j = bb / rp, k = bb % rp;
#pragma unroll
      for (int t = 0; t < R; t++) v[it][t] = lds[(j * r + k + rp * t) * SN + col * SC];
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int b = tid + it * NT;
    if (IT * NT == NB || b < NB) {
      const int col = COLFAST ? b % NCOL : b / NBC;
      const int bb = COLFAST ? b / NCOL : b % NBC;
      const int j = bb / rp, k = bb % rp;
      if (Ls > 1) {
#pragma unroll
        for (int t = 1; t < R; t++) {
          C w = tw[j * t * rp];
          if (SIGN < 0) w.y = -w.y;
          v[it][t] = cmul(v[it][t], w);
        }
      }
      Dft<C, R, SIGN>::run(v[it]);
#pragma unroll
      for (int q = 0; q < R; q++) lds[((j + Ls * q) * rp + k) * SN + col * SC] = v[it][q];
    }
  }
  __syncthreads();
}

template <class C, int L, int NCOL, int NT, int SIGN, int SN, int SC,
          bool COLFAST, int S = 0>
__device__ __forceinline__ void lds_fft(C *lds, const C *__restrict__ tw) {
  if constexpr (S < n_stages(L)) {
    lds_stage<C, L, NCOL, NT, SIGN, SN, SC, COLFAST, S>(lds, tw);
    lds_fft<C, L, NCOL, NT, SIGN, SN, SC, COLFAST, S + 1>(lds, tw);
  }
}

// ------------------------------------------------------- strided passes
// Spectrum multiply of the first c2r pass (axis i): psi_a(k) = rhok(k) * n_a
// with the c2r projection on the kz = 0 and kz = Ny planes (lines where the
// pair is anti-Hermitian vanish: i = Ng/2 for a = x, j = Ng/2 for a = y, the
// whole Nyquist plane for a = z) and the packing P = H0 + i HNy at kk = 0.
template <class C>
__device__ __forceinline__ C spec_mul(C h, int comp, int ki, int kj, int kz) {
  using R = decltype(C{}.x);
  const R n = (comp == 0) ? (R) ki : (comp == 1) ? (R) kj :
              (comp == 2) ? (R) kz : (R) 1;
  return C{h.x * n, h.y * n};
}

// Packed kk = 0 slot from the kz = 0 values at k and -k (p0, q0) and the
// Nyquist values (p1, q1), for component comp.
template <class C>
__device__ __forceinline__ C spec_pack(C p0, C q0, C p1, C q1, int comp, int ki,
    int kj, int kni, int knj, int Ngh) {
  using R = decltype(C{}.x);
  const C a0 = spec_mul(p0, comp, ki, kj, 0);
  const C b0 = spec_mul(q0, comp, kni, knj, 0);
  const C a1 = spec_mul(p1, comp, ki, kj, Ngh);
  const C b1 = spec_mul(q1, comp, kni, knj, Ngh);
  const C h0 = C{(a0.x + b0.x) * (R) 0.5, (a0.y - b0.y) * (R) 0.5};
  const C hn = C{(a1.x + b1.x) * (R) 0.5, (a1.y - b1.y) * (R) 0.5};
  return C{h0.x - hn.y, h0.y + hn.x};
}

// kz = 0 and kz = Ny planes: Hermitian part (H(k) + conj H(-k)) / 2 of each
// plane (exact for the generator's mirrored modes; anti-Hermitian lines of
// psi_x, psi_y and the psi_z Nyquist plane vanish), packed as H0 + i HNy.
// comp 3 = unit multiplier (plain c2r of rhok).
template <class C>
__device__ __forceinline__ C spec_load(const C *__restrict__ rhok,
    const C *__restrict__ rhok_ny, int i, int j, int kk, int Ng, int Ngh,
    int comp) {
  using R = decltype(C{}.x);
  const int ki = (i <= Ngh) ? i : i - Ng;
  const int kj = (j <= Ngh) ? j : j - Ng;
  if (kk > 0) return spec_mul(rhok[((size_t) i * Ng + j) * Ngh + kk], comp, ki, kj, kk);
  const int ni = (i == 0) ? 0 : Ng - i, nj = (j == 0) ? 0 : Ng - j;
  const int kni = (ni <= Ngh) ? ni : ni - Ng, knj = (nj <= Ngh) ? nj : nj - Ng;
  const size_t p = (size_t) i * Ng + j, q = (size_t) ni * Ng + nj;
  return spec_pack(rhok[p * Ngh], rhok[q * Ngh], rhok_ny[p], rhok_ny[q], comp,
      ki, kj, kni, knj, Ngh);
}

// MODE 0: in place on buf; MODE 1: first pass from the generated spectrum
// (either axis).  Grid: x = kk tile, y = outer line.  The transform index n
// walks `stride`.
template <class C, int L, int TK, int NT, int SIGN, int MODE>
__global__ void __launch_bounds__(NT)
k_fft_strided(C *__restrict__ buf, const C *__restrict__ rhok,
    const C *__restrict__ rhok_ny, const C *__restrict__ tw, int Ng, int Ngh,
    size_t stride, size_t outer_stride, int comp, int axis) {
  __shared__ C lds[L * TK];
  // 1-D grid of (Ngh / TK) x Ng tiles, in XCD-aware order: the workgroups
  // of one XCD take consecutive tiles, so column tiles that share cache
  // lines (TK * sizeof(C) < 128 B) run together behind one L2
  const unsigned nx = (unsigned) (Ngh / TK);
  const unsigned t = xcd_item(blockIdx.x, gridDim.x);
  const int kk0 = (int) (t % nx) * TK;
  const int o = (int) (t / nx);
  const size_t base = (size_t) o * outer_stride + kk0;
  constexpr int E = L * TK;
  if (MODE == 0) {
#pragma unroll 16
    for (int e = threadIdx.x; e < E; e += NT) {
      const int n = e / TK, c = e % TK;
      lds[n * TK + c] = buf[base + (size_t) n * stride + c];
    }
  }
  else {
    // rhok has buf's layout: plain, batched loads times n_a; the kz = 0 /
    // Nyquist slot (kk = 0, first column tile only) is redone afterwards
    const int ko = (o <= Ngh) ? o : o - Ng;
#pragma unroll 16
    for (int e = threadIdx.x; e < E; e += NT) {
      const int n = e / TK, c = e % TK;
      const int kn = (n <= Ngh) ? n : n - Ng;
      const C h = rhok[base + (size_t) n * stride + c];
      lds[n * TK + c] = axis ? spec_mul(h, comp, ko, kn, kk0 + c)
                             : spec_mul(h, comp, kn, ko, kk0 + c);
    }
    if (kk0 == 0) {                     // uniform over the block
      __syncthreads();
      for (int n = threadIdx.x; n < L; n += NT)
        lds[n * TK] = axis ? spec_load(rhok, rhok_ny, o, n, 0, Ng, Ngh, comp)
                           : spec_load(rhok, rhok_ny, n, o, 0, Ng, Ngh, comp);
    }
  }
  __syncthreads();
  lds_fft<C, L, TK, NT, SIGN, TK, 1, true>(lds, tw);
#pragma unroll 4
  for (int e = threadIdx.x; e < E; e += NT) {
    const int n = e / TK, c = e % TK;
    buf[base + (size_t) n * stride + c] = lds[n * TK + c];
  }
}

// kz = 0 / Nyquist slot (kk = 0) of the three components, spec_load for
// every (i, j), stored as slot0[comp][j][i] so that the first pass reads
// its column contiguously: the four mirrored loads of a slot are strided
// across the whole spectrum and are done once here instead of in the
// strided pass.
#define SP_T 16
__global__ void __launch_bounds__(SP_T * SP_T)
k_slot0_pack(const cplx *__restrict__ rhok, const cplx *__restrict__ rhok_ny,
    int Ng, cplx *__restrict__ slot0) {
  __shared__ cplx tile[3][SP_T][SP_T + 1];
  const int Ngh = Ng >> 1;
  const size_t plane = (size_t) Ng * Ng;
  const int i0 = blockIdx.x * SP_T, j0 = blockIdx.y * SP_T;
  {
    // loads with j fastest: 2 KB apart, within one page
    const int i = i0 + (int) threadIdx.x / SP_T, j = j0 + (int) threadIdx.x % SP_T;
    if (i < Ng && j < Ng) {
      const int ni = (i == 0) ? 0 : Ng - i, nj = (j == 0) ? 0 : Ng - j;
      const int ki = (i <= Ngh) ? i : i - Ng, kj = (j <= Ngh) ? j : j - Ng;
      const int kni = (ni <= Ngh) ? ni : ni - Ng, knj = (nj <= Ngh) ? nj : nj - Ng;
      const size_t p = (size_t) i * Ng + j, q = (size_t) ni * Ng + nj;
      const cplx p0 = rhok[p * Ngh], q0 = rhok[q * Ngh];
      const cplx p1 = rhok_ny[p], q1 = rhok_ny[q];
#pragma unroll
      for (int comp = 0; comp < 3; comp++)
        tile[comp][i - i0][j - j0] = spec_pack(p0, q0, p1, q1, comp, ki, kj, kni, knj, Ngh);
    }
  }
  __syncthreads();
  // stores with i fastest: slot0[comp][j][i]
  const int i = i0 + (int) threadIdx.x % SP_T, j = j0 + (int) threadIdx.x / SP_T;
  if (i < Ng && j < Ng) {
#pragma unroll
    for (int comp = 0; comp < 3; comp++)
      slot0[comp * plane + (size_t) j * Ng + i] = tile[comp][i - i0][j - j0];
  }
}

// First c2r pass (axis i) of the three displacement components at once:
// the spectrum tile is loaded once into registers, then for a = x, y, z it
// is multiplied by n_a into LDS (the kk = 0 slot from k_slot0_pack), the
// i transform runs in LDS and the tile is stored to out[a].  Reads the
// spectrum once instead of three times.
template <class C, int L, int TK, int NT>
__global__ void __launch_bounds__(NT)
k_fft_first3(C *__restrict__ out0, C *__restrict__ out1, C *__restrict__ out2,
    const C *__restrict__ rhok, const C *__restrict__ slot0,
    const C *__restrict__ tw, int Ng, int Ngh) {
  __shared__ C lds[L * TK];
  constexpr int E = L * TK;
  constexpr int IT = (E + NT - 1) / NT;
  // 1-D grid in XCD-aware order (see k_fft_strided)
  const unsigned nx = (unsigned) (Ngh / TK);
  const unsigned t = xcd_item(blockIdx.x, gridDim.x);
  const int kk0 = (int) (t % nx) * TK;
  const int o = (int) (t / nx);             // j
  const size_t stride = (size_t) Ng * Ngh;
  const size_t base = (size_t) o * Ngh + kk0;
  const int ko = (o <= Ngh) ? o : o - Ng;
  C h[IT];
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int e = threadIdx.x + it * NT;
    if (IT * NT == E || e < E) {
      const int n = e / TK, c = e % TK;
      h[it] = rhok[base + (size_t) n * stride + c];
    }
  }
  C *outs[3] = {out0, out1, out2};
#pragma unroll 1
  for (int comp = 0; comp < 3; comp++) {
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int e = threadIdx.x + it * NT;
      if (IT * NT == E || e < E) {
        const int n = e / TK, c = e % TK;
        const int kn = (n <= Ngh) ? n : n - Ng;
        lds[n * TK + c] = spec_mul(h[it], comp, kn, ko, kk0 + c);
      }
    }
    if (kk0 == 0) {                         // uniform over the block
      const C *col = slot0 + ((size_t) comp * Ng + o) * Ng;
      __syncthreads();
      for (int n = threadIdx.x; n < L; n += NT) lds[n * TK] = col[n];
    }
    __syncthreads();
    lds_fft<C, L, TK, NT, +1, TK, 1, true>(lds, tw);
    C *out = outs[comp];
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int e = threadIdx.x + it * NT;
      if (IT * NT == E || e < E) {
        const int n = e / TK, c = e % TK;
        out[base + (size_t) n * stride + c] = lds[n * TK + c];
      }
    }
    __syncthreads();                        // before the next component's fill
  }
}

// The same pass generating its spectrum tile itself (MRG32k3a, initial
// conditions): each thread draws whole rows of the tile, row i holding the
// cells (i, j, kk0 .. kk0 + TK - 1), whose draws are consecutive (one
// jump-ahead per row), and stages them in LDS (slot c of row n at
// c ^ (n mod TK), so that the lanes writing one column of rows hit distinct
// banks); the kz = 0 column comes from slot0 as in k_fft_first3, whose
// planes k_gen_planes generated.  Saves writing and reading the spectrum.
template <class C, int L, int TK, int NT>
__global__ void __launch_bounds__(NT)
k_fft_first3_gen(C *__restrict__ out0, C *__restrict__ out1, C *__restrict__ out2,
    const C *__restrict__ slot0, const C *__restrict__ tw, int Ng, int Ngh, GenParams gp,
    const double *__restrict__ ptab, const uint32_t *__restrict__ jtab,
    const MrgState *__restrict__ seedp) {
  __shared__ C lds[L * TK];
  constexpr int E = L * TK;
  constexpr int IT = (E + NT - 1) / NT;
  const unsigned nx = (unsigned) (Ngh / TK);
  const unsigned t = xcd_item(blockIdx.x, gridDim.x);
  const int kk0 = (int) (t % nx) * TK;
  const int o = (int) (t / nx);             // j
  const size_t stride = (size_t) Ng * Ngh;
  const size_t base = (size_t) o * Ngh + kk0;
  const int ko = (o <= Ngh) ? o : o - Ng;
  {
    // work units: PR parts of TK / PR cells per row, so that every thread
    // draws when the workgroup has more threads than rows
    constexpr int PR = (NT >= 2 * L && TK >= 2) ? ((NT >= 4 * L && TK >= 4) ? 4 : 2) : 1;
    constexpr int CU = TK / PR;
    const MrgState seed = seedp ? *seedp : gp.seed_state;
    const uint64_t mult = gp.fixamp ? 1 : 2;
    const size_t Ngk = (size_t) Ngh + 1;
#pragma unroll 1
    for (int w = threadIdx.x; w < L * PR; w += NT) {
      const int n = w / PR, ca = (w % PR) * CU;
      C *row = lds + n * TK;
      const int sw = n & (TK - 1);
      const int c0 = (kk0 == 0 && ca == 0) ? 1 : 0;   // kz = 0: from slot0
      if (c0) row[sw] = C{0, 0};
      MrgState g = seed;
      mrg_jump(g, (((size_t) n * Ng + o) * Ngk + kk0 + ca + c0 - 1) * mult, jtab);
#pragma unroll 1
      for (int c = ca + c0; c < ca + CU; c++) {
        const double amp = gp.fixamp ? 1.0 : mrg_next(g);
        const double ph = mrg_next(g);
        const GenCell gc = gen_cell(n, o, kk0 + c, amp, ph, gp, ptab);
        row[c ^ sw] = C{gc.re, gc.im};
      }
    }
  }
  __syncthreads();
  C h[IT];
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int e = threadIdx.x + it * NT;
    if (IT * NT == E || e < E) {
      const int n = e / TK, c = e % TK;
      h[it] = lds[n * TK + (c ^ (n & (TK - 1)))];
    }
  }
  __syncthreads();
  C *outs[3] = {out0, out1, out2};
#pragma unroll 1
  for (int comp = 0; comp < 3; comp++) {
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int e = threadIdx.x + it * NT;
      if (IT * NT == E || e < E) {
        const int n = e / TK, c = e % TK;
        const int kn = (n <= Ngh) ? n : n - Ng;
        lds[n * TK + c] = spec_mul(h[it], comp, kn, ko, kk0 + c);
      }
    }
    if (kk0 == 0) {                         // uniform over the block
      const C *col = slot0 + ((size_t) comp * Ng + o) * Ng;
      __syncthreads();
      for (int n = threadIdx.x; n < L; n += NT) lds[n * TK] = col[n];
    }
    __syncthreads();
    lds_fft<C, L, TK, NT, +1, TK, 1, true>(lds, tw);
    C *out = outs[comp];
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int e = threadIdx.x + it * NT;
      if (IT * NT == E || e < E) {
        const int n = e / TK, c = e % TK;
        out[base + (size_t) n * stride + c] = lds[n * TK + c];
      }
    }
    __syncthreads();                        // before the next component's fill
  }
}

// ------------------------------------------------------------ row passes
// c2r along kk: in place, rows of Lh complex -> 2 Lh reals.
// twN[m] = exp(+2 pi i m / (2 Lh)), twh[m] = exp(+2 pi i m / Lh).
template <class C, int Lh, int TR, int NT>
__global__ void __launch_bounds__(NT)
k_fft_rows_c2r(C *__restrict__ buf, const C *__restrict__ twN,
    const C *__restrict__ twh) {
  using R = decltype(C{}.x);
  __shared__ C lds[Lh * TR];
  const size_t row0 = (size_t) blockIdx.x * TR;
  C *g = buf + row0 * Lh;
  constexpr int E = Lh * TR;
#pragma unroll 16
  for (int e = threadIdx.x; e < E; e += NT) lds[e] = g[e];
  __syncthreads();
  // Hermitian pre-processing, pairs (m, Lh - m) for m in [0, Lh/2].
  constexpr int NP = Lh / 2 + 1;
  constexpr int NPT = NP * TR;
  constexpr int IT = (NPT + NT - 1) / NT;
  C za[IT], zb[IT];
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int p = threadIdx.x + it * NT;
    if (p < NPT) {
      const int rr = p / NP, m = p % NP;
      C *row = lds + rr * Lh;
      if (m == 0) {
        const C s = row[0];                 // (G0, GNy)
        za[it] = C{s.x + s.y, s.x - s.y};
      }
      else {
        const int m2 = Lh - m;
        const C xm = row[m], xn = row[m2];
        const C wm = twN[m], wn = twN[m2];
        // Z[m]  = X[m]  + conj(X[m2]) + i W^m  (X[m]  - conj(X[m2]))
        // Z[m2] = X[m2] + conj(X[m])  + i W^m2 (X[m2] - conj(X[m]))
        C s1 = cadd(xm, cconj(xn)), d1 = csub(xm, cconj(xn));
        C s2 = cadd(xn, cconj(xm)), d2 = csub(xn, cconj(xm));
        C t1 = cmul(wm, d1), t2 = cmul(wn, d2);
        za[it] = C{s1.x - t1.y, s1.y + t1.x};
        zb[it] = C{s2.x - t2.y, s2.y + t2.x};
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int p = threadIdx.x + it * NT;
    if (p < NPT) {
      const int rr = p / NP, m = p % NP;
      C *row = lds + rr * Lh;
      row[m] = za[it];
      if (m != 0 && m != Lh - m) row[Lh - m] = zb[it];
    }
  }
  __syncthreads();
  (void) sizeof(R);
  lds_fft<C, Lh, TR, NT, +1, 1, Lh, false>(lds, twh);
#pragma unroll 4
  for (int e = threadIdx.x; e < E; e += NT) g[e] = lds[e];
}

// r2c along kk (forward): rows of 2 Lh reals viewed as Lh complex z[n] =
// y[2n] + i y[2n+1]; after the Lh-point forward FFT Z[m]:
//   X[m] = (Z[m] + conj(Z[Lh-m]))/2 - i/2 W^-m (Z[m] - conj(Z[Lh-m])),
// X[0] and X[Lh] are real; they are packed as X[0] + i X[Lh] in slot 0.
template <class C, int Lh, int TR, int NT>
__global__ void __launch_bounds__(NT)
k_fft_rows_r2c(C *__restrict__ buf, const C *__restrict__ twN,
    const C *__restrict__ twh) {
  using R = decltype(C{}.x);
  __shared__ C lds[Lh * TR];
  const size_t row0 = (size_t) blockIdx.x * TR;
  C *g = buf + row0 * Lh;
  constexpr int E = Lh * TR;
#pragma unroll 16
  for (int e = threadIdx.x; e < E; e += NT) lds[e] = g[e];
  __syncthreads();
  lds_fft<C, Lh, TR, NT, -1, 1, Lh, false>(lds, twh);
  constexpr int NP = Lh / 2 + 1;
  constexpr int NPT = NP * TR;
  constexpr int IT = (NPT + NT - 1) / NT;
  C xa[IT], xb[IT];
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int p = threadIdx.x + it * NT;
    if (p < NPT) {
      const int rr = p / NP, m = p % NP;
      C *row = lds + rr * Lh;
      if (m == 0) {
        const C z = row[0];
        xa[it] = C{z.x + z.y, z.x - z.y};   // (X[0], X[Lh]) packed
      }
      else {
        const int m2 = Lh - m;
        const C zm = row[m], zn = row[m2];
        C wm = twN[m], wn = twN[m2];
        wm.y = -wm.y; wn.y = -wn.y;          // W^-m
        C s1 = cadd(zm, cconj(zn)), d1 = csub(zm, cconj(zn));
        C s2 = cadd(zn, cconj(zm)), d2 = csub(zn, cconj(zm));
        C t1 = cmul(wm, d1), t2 = cmul(wn, d2);
        // X = s/2 - i/2 * t
        xa[it] = C{(R) 0.5 * (s1.x + t1.y), (R) 0.5 * (s1.y - t1.x)};
        xb[it] = C{(R) 0.5 * (s2.x + t2.y), (R) 0.5 * (s2.y - t2.x)};
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int p = threadIdx.x + it * NT;
    if (p < NPT) {
      const int rr = p / NP, m = p % NP;
      C *row = lds + rr * Lh;
      row[m] = xa[it];
      if (m != 0 && m != Lh - m) row[Lh - m] = xb[it];
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int e = threadIdx.x; e < E; e += NT) g[e] = lds[e];
}

// Threads of a pass over a tile of E elements: 8 elements per thread, up
// to 1024 threads (twice the threads of 16 per thread: the 512^3 first pass
// 1.24 -> 1.04 ms, rows 0.83 -> 0.78 ms, j 0.74 -> 0.72 ms; 1024^3 first
// pass 14.3 -> 12.1 ms); larger tiles keep 16 per thread, at most 512
// threads, whose registers would spill at 1024.
template <int E> constexpr int pass_threads() {
  return (E / 8 <= 1024) ? (E / 8 < 64 ? 64 : E / 8) : (E / 16 > 512 ? 512 : E / 16);
}

// --------------------------------------------------------- dispatching
template <int L> struct StridedCfg {
#ifdef SINGLE_PREC
  static constexpr int TKMAX = (L <= 1024) ? 16 : 8;
#else
  static constexpr int TKMAX = (L <= 1024) ? 8 : 4;
#endif
};
template <int L> struct RowCfg {
#ifdef SINGLE_PREC
  static constexpr int TR = (L <= 512) ? 16 : 8;
#else
  static constexpr int TR = (L <= 512) ? 8 : 4;
#endif
  static constexpr int E = L * TR;
  static constexpr int NT = pass_threads<E>();
};

template <int L, int TK, int SIGN>
static hipError_t launch_strided_tk(cplx *buf, const cplx *rhok,
    const cplx *rhok_ny, const cplx *tw, int Ng, int Ngh, int axis, int mode,
    int comp, hipStream_t st) {
  constexpr int E = L * TK;
  constexpr int NT = pass_threads<E>();
  const size_t plane = (size_t) Ng * Ngh;
  const size_t stride = (axis == 0) ? plane : (size_t) Ngh;
  const size_t outer = (axis == 0) ? (size_t) Ngh : plane;
  dim3 grid((unsigned) ((Ngh / TK) * (size_t) Ng));
  if (mode == 0)
    hipLaunchKernelGGL((k_fft_strided<cplx, L, TK, NT, SIGN, 0>), grid, dim3(NT), 0, st,
        buf, rhok, rhok_ny, tw, Ng, Ngh, stride, outer, comp, axis);
  else
    hipLaunchKernelGGL((k_fft_strided<cplx, L, TK, NT, SIGN, 1>), grid, dim3(NT), 0, st,
        buf, rhok, rhok_ny, tw, Ng, Ngh, stride, outer, comp, axis);
  return hipGetLastError();
}

// Column tile widths of the strided passes: the default (StridedCfg), or a
// candidate of the schedule (4, 8, 16 columns, where Ngh allows it and the
// tile fits the 160 KB of LDS).
template <int L, int TK> constexpr bool tk_fits() {
  return L * TK * (int) sizeof(cplx) <= 160 * 1024;
}

template <int L, int SIGN>
static hipError_t launch_strided_default(cplx *buf, const cplx *rhok,
    const cplx *rhok_ny, const cplx *tw, int Ng, int Ngh, int axis, int mode,
    int comp, hipStream_t st) {
  constexpr int T = StridedCfg<L>::TKMAX;
  if (Ngh % T == 0) return launch_strided_tk<L, T, SIGN>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, st);
  if constexpr (T >= 8) if (Ngh % (T / 2) == 0) return launch_strided_tk<L, T / 2, SIGN>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, st);
  if constexpr (T >= 4) if (Ngh % 4 == 0) return launch_strided_tk<L, 4, SIGN>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, st);
  if (Ngh % 2 == 0) return launch_strided_tk<L, 2, SIGN>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, st);
  return launch_strided_tk<L, 1, SIGN>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, st);
}

// Sizes whose passes have schedule candidates (smaller grids: defaults).
template <int L> constexpr bool tunable() { return L >= 256; }

template <int L, int SIGN>
static hipError_t launch_strided(cplx *buf, const cplx *rhok,
    const cplx *rhok_ny, const cplx *tw, int Ng, int Ngh, int axis, int mode,
    int comp, int tk, hipStream_t st) {
  if constexpr (tunable<L>()) {
    if constexpr (tk_fits<L, 4>())
      if (tk == 4 && Ngh % 4 == 0)
        return launch_strided_tk<L, 4, SIGN>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, st);
    if constexpr (tk_fits<L, 8>())
      if (tk == 8 && Ngh % 8 == 0)
        return launch_strided_tk<L, 8, SIGN>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, st);
    if constexpr (tk_fits<L, 16>())
      if (tk == 16 && Ngh % 16 == 0)
        return launch_strided_tk<L, 16, SIGN>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, st);
  }
  return launch_strided_default<L, SIGN>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, st);
}

template <int L, int T>
static hipError_t launch_first3_tk(cplx *o0, cplx *o1, cplx *o2, const cplx *rhok,
    const cplx *slot0, const cplx *tw, int Ng, int Ngh, hipStream_t st) {
  constexpr int E = L * T;
  constexpr int NT = pass_threads<E>();
  hipLaunchKernelGGL((k_fft_first3<cplx, L, T, NT>), dim3((unsigned) ((Ngh / T) * (size_t) Ng)),
      dim3(NT), 0, st, o0, o1, o2, rhok, slot0, tw, Ng, Ngh);
  return hipGetLastError();
}

template <int L, int T>
static hipError_t launch_first3_gen_tk(cplx *o0, cplx *o1, cplx *o2, const cplx *slot0,
    const cplx *tw, int Ng, int Ngh, const GenParams &gp, const double *ptab,
    const uint32_t *jtab, const MrgState *seedp, hipStream_t st) {
  constexpr int E = L * T;
  constexpr int NT = pass_threads<E>();    // (4 per thread: 1.33 against 1.04 ms at 512^3)
  hipLaunchKernelGGL((k_fft_first3_gen<cplx, L, T, NT>),
      dim3((unsigned) ((Ngh / T) * (size_t) Ng)), dim3(NT), 0, st, o0, o1, o2, slot0, tw,
      Ng, Ngh, gp, ptab, jtab, seedp);
  return hipGetLastError();
}

// The three-component pass keeps its tile in registers and runs three
// transforms per workgroup: the narrowest power-of-two tile whose LDS
// reaches 64 KB (at most 16 columns; measured at 512, 1024 and 1536).
template <int L> struct First3Cfg {
  static constexpr int bytes = L * (int) sizeof(cplx);
  static constexpr int TKMAX = (16 * bytes <= 65536) ? 16 : (2 * bytes >= 65536) ? 2
      : (4 * bytes >= 65536) ? 4 : 8;
};

template <int L>
static hipError_t launch_first3_default(cplx *o0, cplx *o1, cplx *o2, const cplx *rhok,
    const cplx *slot0, const cplx *tw, int Ng, int Ngh, hipStream_t st) {
  constexpr int T = First3Cfg<L>::TKMAX;
  if (Ngh % T == 0) return launch_first3_tk<L, T>(o0, o1, o2, rhok, slot0, tw, Ng, Ngh, st);
  if constexpr (T >= 8) if (Ngh % (T / 2) == 0) return launch_first3_tk<L, T / 2>(o0, o1, o2, rhok, slot0, tw, Ng, Ngh, st);
  if constexpr (T >= 4) if (Ngh % 4 == 0) return launch_first3_tk<L, 4>(o0, o1, o2, rhok, slot0, tw, Ng, Ngh, st);
  if (Ngh % 2 == 0) return launch_first3_tk<L, 2>(o0, o1, o2, rhok, slot0, tw, Ng, Ngh, st);
  return launch_first3_tk<L, 1>(o0, o1, o2, rhok, slot0, tw, Ng, Ngh, st);
}

template <int L>
static hipError_t launch_first3(cplx *o0, cplx *o1, cplx *o2, const cplx *rhok,
    const cplx *slot0, const cplx *tw, int Ng, int Ngh, int tk, hipStream_t st) {
  if constexpr (tunable<L>()) {
    if constexpr (tk_fits<L, 4>())
      if (tk == 4 && Ngh % 4 == 0)
        return launch_first3_tk<L, 4>(o0, o1, o2, rhok, slot0, tw, Ng, Ngh, st);
    if constexpr (tk_fits<L, 8>())
      if (tk == 8 && Ngh % 8 == 0)
        return launch_first3_tk<L, 8>(o0, o1, o2, rhok, slot0, tw, Ng, Ngh, st);
    if constexpr (tk_fits<L, 16>())
      if (tk == 16 && Ngh % 16 == 0)
        return launch_first3_tk<L, 16>(o0, o1, o2, rhok, slot0, tw, Ng, Ngh, st);
  }
  return launch_first3_default<L>(o0, o1, o2, rhok, slot0, tw, Ng, Ngh, st);
}

template <int L>
static hipError_t launch_first3_gen(cplx *o0, cplx *o1, cplx *o2, const cplx *slot0,
    const cplx *tw, int Ng, int Ngh, int tk, const GenParams &gp, const double *ptab,
    const uint32_t *jtab, const MrgState *seedp, hipStream_t st) {
  if constexpr (tunable<L>()) {
    if constexpr (tk_fits<L, 4>())
      if (tk == 4 && Ngh % 4 == 0)
        return launch_first3_gen_tk<L, 4>(o0, o1, o2, slot0, tw, Ng, Ngh, gp, ptab, jtab, seedp, st);
    if constexpr (tk_fits<L, 8>())
      if (tk == 8 && Ngh % 8 == 0)
        return launch_first3_gen_tk<L, 8>(o0, o1, o2, slot0, tw, Ng, Ngh, gp, ptab, jtab, seedp, st);
    if constexpr (tk_fits<L, 16>())
      if (tk == 16 && Ngh % 16 == 0)
        return launch_first3_gen_tk<L, 16>(o0, o1, o2, slot0, tw, Ng, Ngh, gp, ptab, jtab, seedp, st);
  }
  constexpr int T = First3Cfg<L>::TKMAX;
  if (Ngh % T == 0)
    return launch_first3_gen_tk<L, T>(o0, o1, o2, slot0, tw, Ng, Ngh, gp, ptab, jtab, seedp, st);
  if constexpr (T >= 8) if (Ngh % (T / 2) == 0)
    return launch_first3_gen_tk<L, T / 2>(o0, o1, o2, slot0, tw, Ng, Ngh, gp, ptab, jtab, seedp, st);
  if constexpr (T >= 4) if (Ngh % 4 == 0)
    return launch_first3_gen_tk<L, 4>(o0, o1, o2, slot0, tw, Ng, Ngh, gp, ptab, jtab, seedp, st);
  if (Ngh % 2 == 0)
    return launch_first3_gen_tk<L, 2>(o0, o1, o2, slot0, tw, Ng, Ngh, gp, ptab, jtab, seedp, st);
  return launch_first3_gen_tk<L, 1>(o0, o1, o2, slot0, tw, Ng, Ngh, gp, ptab, jtab, seedp, st);
}

template <int Lh>
static hipError_t launch_rows(cplx *buf, const cplx *twN, const cplx *twh,
    size_t nrows, int sign, hipStream_t st) {
  constexpr int TR = RowCfg<Lh>::TR, NT = RowCfg<Lh>::NT;
  if (nrows % TR) return hipErrorInvalidValue;
  dim3 grid((unsigned) (nrows / TR));
  if (sign > 0)
    hipLaunchKernelGGL((k_fft_rows_c2r<cplx, Lh, TR, NT>), grid, dim3(NT), 0, st, buf, twN, twh);
  else
    hipLaunchKernelGGL((k_fft_rows_r2c<cplx, Lh, TR, NT>), grid, dim3(NT), 0, st, buf, twN, twh);
  return hipGetLastError();
}

#define EZ_FFT_SIZES(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) \
  X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048)

bool fft_supported(int Ng) {
  switch (Ng) {
#define X(n) case n:
    EZ_FFT_SIZES(X)
#undef X
      return true;
    default: return false;
  }
}

// ------------------------------------------------------------ schedules
// Per grid size (and precision of this build): the column tile width of the
// strided passes (axis j; and axis i where the first pass is split), the
// tile width of the three-component first pass, and whether the first pass
// runs as three single-component strided passes instead (each reads the
// spectrum; smaller tiles in registers).  tk = 0 means the built-in
// default.  Schedules come from fft_autotune (timing every candidate on
// the caller's buffers) or fft_set_schedule.
namespace {
std::mutex g_sched_mu;
std::map<int, FftSchedule> g_sched;
}

FftSchedule fft_schedule(int Ng, bool *tuned) {
  std::lock_guard<std::mutex> lk(g_sched_mu);
  auto it = g_sched.find(Ng);
  if (tuned) *tuned = it != g_sched.end();
  return it != g_sched.end() ? it->second : FftSchedule{0, 0, 0};
}

void fft_set_schedule(int Ng, const FftSchedule &s) {
  std::lock_guard<std::mutex> lk(g_sched_mu);
  g_sched[Ng] = s;
}

hipError_t fft_strided_pass(cplx *buf, const cplx *rhok, const cplx *rhok_ny,
    const cplx *tw, int Ng, int axis, int sign, int mode, int comp,
    hipStream_t st, int tk) {
  const int Ngh = Ng / 2;
  if (tk < 0) tk = fft_schedule(Ng, nullptr).tk_j;
  switch (Ng) {
#define X(n) case n: return (sign > 0) \
      ? launch_strided<n, +1>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, tk, st) \
      : launch_strided<n, -1>(buf, rhok, rhok_ny, tw, Ng, Ngh, axis, mode, comp, tk, st);
    EZ_FFT_SIZES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
}

hipError_t fft_row_pass(cplx *buf, const cplx *twN, const cplx *twh, int Ng,
    int sign, hipStream_t st) {
  const size_t nrows = (size_t) Ng * Ng;
  switch (Ng / 2) {
#define X(n) case n / 2: return launch_rows<n / 2>(buf, twN, twh, nrows, sign, st);
    EZ_FFT_SIZES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
}

hipError_t fft_slot0_pack(const cplx *rhok, const cplx *rhok_ny, cplx *slot0,
    int Ng, hipStream_t st) {
  const unsigned nb = (unsigned) ((Ng + SP_T - 1) / SP_T);
  hipLaunchKernelGGL(k_slot0_pack, dim3(nb, nb), dim3(SP_T * SP_T), 0, st,
      rhok, rhok_ny, Ng, slot0);
  return hipGetLastError();
}

hipError_t fft_first_pass3(cplx *out0, cplx *out1, cplx *out2, const cplx *rhok,
    const cplx *rhok_ny, const cplx *slot0, const cplx *tw, int Ng, hipStream_t st,
    const FftSchedule *sched) {
  const int Ngh = Ng / 2;
  const FftSchedule sc = sched ? *sched : fft_schedule(Ng, nullptr);
  if (sc.split) {                       // three single-component passes
    cplx *outs[3] = {out0, out1, out2};
    for (int c = 0; c < 3; c++) {
      const hipError_t e = fft_strided_pass(outs[c], rhok, rhok_ny, tw, Ng, 0, +1, 1, c, st,
          sc.tk_i);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  switch (Ng) {
#define X(n) case n: return launch_first3<n>(out0, out1, out2, rhok, slot0, tw, Ng, Ngh, sc.tk_i, st);
    EZ_FFT_SIZES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
}

hipError_t fft_first_pass3_gen(cplx *out0, cplx *out1, cplx *out2, const cplx *slot0,
    const cplx *tw, int Ng, const GenParams &gp, const double *ptab, const uint32_t *jtab,
    const MrgState *seedp, hipStream_t st, int tk) {
  const int Ngh = Ng / 2;
  if (tk < 0) {
    const FftSchedule sc = fft_schedule(Ng, nullptr);
    if (sc.split) return hipErrorNotSupported;
    tk = sc.tk_g ? sc.tk_g : sc.tk_i;
  }
  switch (Ng) {
#define X(n) case n: return launch_first3_gen<n>(out0, out1, out2, slot0, tw, Ng, Ngh, tk, gp, ptab, jtab, seedp, st);
    EZ_FFT_SIZES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
}

// Times every candidate schedule of grid Ng on the caller's buffers (the
// spectrum rhok / rhok_ny / slot0 is read; out0..out2 are overwritten) and
// records the fastest: tile widths 4, 8 and 16 (where they fit), the first
// pass fused or split.  Each candidate: one untimed run, then the best of
// two timed runs.
hipError_t fft_autotune(int Ng, cplx *out0, cplx *out1, cplx *out2, const cplx *rhok,
    const cplx *rhok_ny, const cplx *slot0, const cplx *tw, hipStream_t st,
    FftSchedule *best, double *ms_best, const GenParams *gp, const double *ptab,
    const uint32_t *jtab) {
  hipEvent_t a, b;
  hipError_t e;
  if ((e = hipEventCreate(&a)) != hipSuccess) return e;
  if ((e = hipEventCreate(&b)) != hipSuccess) { (void) hipEventDestroy(a); return e; }
  auto time_it = [&](auto &&run, float &ms) -> hipError_t {
    hipError_t r = run();
    if (r != hipSuccess) return r;
    ms = 1e30f;
    for (int rep = 0; rep < 2; rep++) {
      if ((r = hipEventRecord(a, st)) != hipSuccess || (r = run()) != hipSuccess ||
          (r = hipEventRecord(b, st)) != hipSuccess || (r = hipEventSynchronize(b)) != hipSuccess)
        return r;
      float t = 0;
      if ((r = hipEventElapsedTime(&t, a, b)) != hipSuccess) return r;
      ms = t < ms ? t : ms;
    }
    return hipSuccess;
  };
  const int Ngh = Ng / 2;
  const int tks[4] = {0, 4, 8, 16};
  FftSchedule bs{0, 0, 0};
  float best_j = 1e30f, best_i = 1e30f;
  for (int t : tks) {
    if (t && Ngh % t) continue;
    if (t * Ng * (int) sizeof(cplx) > 160 * 1024) continue;
    float ms;
    e = time_it([&] { return fft_strided_pass(out0, nullptr, nullptr, tw, Ng, 1, +1, 0, 0, st, t); },
        ms);
    if (e != hipSuccess) break;
    if (ms < best_j) { best_j = ms; bs.tk_j = t; }
  }
  for (int split = 0; split < 2 && e == hipSuccess; split++) {
    for (int t : tks) {
      if (t && Ngh % t) continue;
      if (t * Ng * (int) sizeof(cplx) > 160 * 1024) continue;
      const FftSchedule c{0, t, split};
      float ms;
      e = time_it([&] { return fft_first_pass3(out0, out1, out2, rhok, rhok_ny, slot0, tw, Ng,
          st, &c); }, ms);
      if (e != hipSuccess) break;
      if (ms < best_i) { best_i = ms; bs.tk_i = t; bs.split = split; }
    }
  }
  // the first pass generating its spectrum (used with the fused first pass)
  float best_g = 1e30f;
  for (int t : tks) {
    if (!gp || e != hipSuccess) break;
    if (t && Ngh % t) continue;
    if (t * Ng * (int) sizeof(cplx) > 160 * 1024) continue;
    float ms;
    e = time_it([&] { return fft_first_pass3_gen(out0, out1, out2, slot0, tw, Ng, *gp, ptab,
        jtab, nullptr, st, t); }, ms);
    if (e != hipSuccess) break;
    if (ms < best_g) { best_g = ms; bs.tk_g = t; }
  }
  (void) hipEventDestroy(a);
  (void) hipEventDestroy(b);
  if (e != hipSuccess) return e;
  fft_set_schedule(Ng, bs);
  if (best) *best = bs;
  if (ms_best) { ms_best[0] = best_j; ms_best[1] = best_i; }
  return hipSuccess;
}
