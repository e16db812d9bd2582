// gen.h -- one cell of the initial-condition spectrum (perturb.c:78-359),
// shared by the spectrum generator (field.hip) and the first FFT pass that
// generates its spectrum tile itself (fft.hip).  The arithmetic keeps the
// reference's IEEE operation order (no contraction into fused multiply-adds)
// whatever the including file is compiled with.
#pragma once
#include "common.h"
#include "kernels.h"

// One cell (i, j, k) of the generated spectrum from its two draws (amp in
// (0, 1], phase draw ph01 in [0, 1)), perturb.c:78-359: whether this cell
// is written (cells of the kz = 0 and Nyquist planes outside the lower half
// are filled by their mirror), whether it also writes its mirror (ni, nj),
// and its value.
struct GenCell {
  real re, im;
  bool write, mirror;
  int ni, nj;
};

__device__ __forceinline__ GenCell gen_cell(int i, int j, int k, double amp, double ph01,
    const GenParams &gp, const double *__restrict__ ptab) {
#pragma clang fp contract(off)
  const int Ng = gp.Ng, Ngh = Ng >> 1;
  GenCell c{0, 0, true, false, 0, 0};
  double ph = ph01 * 2 * EZ_PI;
  const bool plane_k = (k == 0) || (k == Ngh);
  if (plane_k) {
    c.ni = (i == 0) ? 0 : Ng - i;
    c.nj = (j == 0) ? 0 : Ng - j;
    const bool lower = (i == c.ni && j <= Ngh) || (i != c.ni && i <= Ngh);
    if (!lower) c.write = false;      // filled by its mirror
    c.mirror = lower;
  }
  if (!c.write) return c;
  const int ki = (i <= Ngh) ? i : i - Ng;
  const int kj = (j <= Ngh) ? j : j - Ng;
  const int ksqi = ki * ki + kj * kj + k * k;
  const double ksq = (double) ksqi;
  double P = ptab[ksqi];
  double a = gp.fixamp ? 1.0 : -log(amp);
  if (gp.logpk) P = gp.fixamp ? exp(P * 0.5) : exp(P * 0.5) * sqrt(a);
  else P = gp.fixamp ? sqrt(P) : sqrt(P * a);
  P *= gp.Pfac;
  if (gp.iphase) ph += EZ_PI;
  P /= (ksq * gp.kfac);
  double sn, cs;
  sincos(ph, &sn, &cs);                 // one argument reduction for both
  c.re = (real) (-P * sn);
  c.im = (real) (P * cs);
  // perturb.c:341-359: self-conjugate modes keep only the imaginary part
  // (the reference writes (+P sin, P cos) twice, then zeroes re).
  if (c.mirror && i == c.ni && j == c.nj) c.re = 0;
  return c;
}

// The mirror of a written kz = 0 / Nyquist cell (not for self-conjugate
// modes): (-re, im) at (ni, nj, k).
__device__ __forceinline__ void gen_mirror(const GenCell &c, int i, int j, int k,
    const GenParams &gp, cplx *__restrict__ rhok, cplx *__restrict__ rhok_ny,
    real *__restrict__ ref_copy) {
  if (!c.mirror || (i == c.ni && j == c.nj)) return;
  const int Ng = gp.Ng, Ngh = Ng >> 1, Ngk = Ngh + 1;
  const cplx w = {-c.re, c.im};
  if (k < Ngh) rhok[((size_t) c.ni * Ng + c.nj) * Ngh + k] = w;
  else rhok_ny[(size_t) c.ni * Ng + c.nj] = w;
  if (ref_copy) {
    const size_t nidx = ((size_t) c.ni * Ng + c.nj) * Ngk + k;
    ref_copy[2 * nidx] = -c.re; ref_copy[2 * nidx + 1] = c.im;
  }
}

