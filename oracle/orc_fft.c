/*
 * orc_fft.c -- discrete Fourier transforms of the CPU oracle (test
 * infrastructure only).
 *
 * The reference calls FFTW3 (fftw_def.h:29-64): 3-D complex-to-real plans
 * (FFT_PLAN_C2R, dens_field.c:141-142, executed at perturb.c:361,397,398,
 * 556,592,593) and one real-to-complex plan (dens_field.c:197-201).  FFTW is
 * not available here; this file computes the same unnormalised transforms:
 *   c2r: out[x] = sum over the full Hermitian-extended spectrum of
 *        H(k) e^{+2 pi i k.x / Ng}, where the imaginary parts of the kz = 0
 *        and kz = Ng/2 terms are dropped by the last 1-D c2r (FFTW and
 *        numpy.fft.irfftn agree on this);
 *   r2c: out[k] = sum_x in[x] e^{-2 pi i k.x / Ng}, kz in [0, Ng/2].
 * Accumulation is in double for both builds; results are rounded to `real`
 * after every axis pass, as a float FFTW plan stores them.
 * A mixed-radix Stockham self-sorting FFT (any length; radices 4, 2, 3, 5
 * and a direct DFT for other prime factors) does the 1-D work.
 */
#include "orc.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef struct { double re, im; } cplx;

typedef struct {
  int n, nf;
  int fac[64];
  cplx *tw;                     /* tw[m] = exp(sign 2 pi i m / n) */
  int sign;
} plan1d;

static int plan_make(plan1d *p, int n, int sign) {
  p->n = n; p->sign = sign; p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
  }
  p->tw = malloc(sizeof(cplx) * (size_t) (n > 0 ? n : 1));
  if (!p->tw) return 1;
  for (int i = 0; i < n; i++) {
    double a = sign * 2 * ORC_PI * (double) i / n;
    p->tw[i].re = cos(a);
    p->tw[i].im = sin(a);
  }
  return 0;
}

static void plan_free(plan1d *p) { free(p->tw); p->tw = NULL; }

/* Stockham pass sequence: y[j*r + k] holds the length-L DFT (frequency j) of
   the subsequence x[k + r*m].  A radix-p stage maps (L, r) -> (L*p, r/p). */
static void fft_exec(const plan1d *p, cplx *x, cplx *work) {
  const int n = p->n;
  cplx *src = x, *dst = work;
  int L = 1;
  cplx in[64], out[64];
  for (int s = 0; s < p->nf; s++) {
    const int rad = p->fac[s];
    const int Lp = L * rad, r = n / L, rp = r / rad;
    for (int j = 0; j < L; j++) {
      for (int k = 0; k < rp; k++) {
        for (int t = 0; t < rad; t++) {
          cplx v = src[(size_t) j * r + k + (size_t) rp * t];
          cplx w = p->tw[(size_t) j * t * rp];
          in[t].re = v.re * w.re - v.im * w.im;
          in[t].im = v.re * w.im + v.im * w.re;
        }
        for (int q = 0; q < rad; q++) {
          double sr = 0, si = 0;
          for (int t = 0; t < rad; t++) {
            cplx w = p->tw[(size_t) ((q * t) % rad) * (n / rad)];
            sr += in[t].re * w.re - in[t].im * w.im;
            si += in[t].re * w.im + in[t].im * w.re;
          }
          out[q].re = sr; out[q].im = si;
        }
        for (int q = 0; q < rad; q++)
          dst[(size_t) (j + L * q) * rp + k] = out[q];
      }
    }
    cplx *tmp = src; src = dst; dst = tmp;
    L = Lp;
  }
  if (src != x) memcpy(x, src, sizeof(cplx) * (size_t) n);
}

int orc_fft_1d(int n, real *data, int sign) {
  plan1d p;
  if (plan_make(&p, n, sign)) return 1;
  for (int s = 0; s < p.nf; s++) if (p.fac[s] > 64) { plan_free(&p); return 1; }
  cplx *x = malloc(sizeof(cplx) * n * 2);
  if (!x) { plan_free(&p); return 1; }
  for (int i = 0; i < n; i++) { x[i].re = data[2 * i]; x[i].im = data[2 * i + 1]; }
  fft_exec(&p, x, x + n);
  for (int i = 0; i < n; i++) { data[2 * i] = x[i].re; data[2 * i + 1] = x[i].im; }
  free(x);
  plan_free(&p);
  return 0;
}

/* Transform along one axis of a (n0, n1, n2) complex array stored as `real`
   pairs, for every index of the other two axes. */
static int axis_pass(real *a, int n0, int n1, int n2, int axis,
    const plan1d *p, int nthread) {
  const int n = (axis == 0) ? n0 : (axis == 1) ? n1 : n2;
  const size_t stride = (axis == 0) ? (size_t) n1 * n2 : (axis == 1) ? n2 : 1;
  const size_t nb = (size_t) n0 * n1 * n2 / n;
  int fail = 0;
#pragma omp parallel num_threads(nthread)
  {
    cplx *buf = malloc(sizeof(cplx) * n * 2);
    if (!buf) {
#pragma omp atomic write
      fail = 1;
    }
    else {
#pragma omp for schedule(static)
      for (size_t b = 0; b < nb; b++) {
        size_t base;
        if (axis == 0) base = b;                          /* (i1,i2) */
        else if (axis == 1) base = (b / n2) * (size_t) n1 * n2 + b % n2;
        else base = b * n2;
        for (int i = 0; i < n; i++) {
          buf[i].re = a[2 * (base + i * stride)];
          buf[i].im = a[2 * (base + i * stride) + 1];
        }
        fft_exec(p, buf, buf + n);
        for (int i = 0; i < n; i++) {
          a[2 * (base + i * stride)] = (real) buf[i].re;
          a[2 * (base + i * stride) + 1] = (real) buf[i].im;
        }
      }
      free(buf);
    }
  }
  return fail;
}

int orc_fft_c2r_3d(int Ng, const real *in, real *out, int nthread) {
  const int Ngk = Ng / 2 + 1;
  const size_t nh = (size_t) Ng * Ng * Ngk;
  real *w = malloc(sizeof(real) * 2 * nh);
  if (!w) return 1;
  memcpy(w, in, sizeof(real) * 2 * nh);
  plan1d p;
  if (plan_make(&p, Ng, +1)) { free(w); return 1; }
  int fail = axis_pass(w, Ng, Ng, Ngk, 0, &p, nthread) ||
             axis_pass(w, Ng, Ng, Ngk, 1, &p, nthread);
  if (!fail) {
#pragma omp parallel num_threads(nthread)
    {
      cplx *buf = malloc(sizeof(cplx) * Ng * 2);
      if (!buf) {
#pragma omp atomic write
        fail = 1;
      }
      else {
#pragma omp for schedule(static)
        for (size_t row = 0; row < (size_t) Ng * Ng; row++) {
          const real *h = w + 2 * row * Ngk;
          /* Hermitian extension of the last axis; the DC and Nyquist terms
             contribute their real parts only. */
          buf[0].re = h[0]; buf[0].im = 0;
          for (int k = 1; k < Ngk; k++) {
            buf[k].re = h[2 * k]; buf[k].im = h[2 * k + 1];
            buf[Ng - k].re = h[2 * k]; buf[Ng - k].im = -h[2 * k + 1];
          }
          if ((Ng & 1) == 0) { buf[Ng / 2].re = h[Ng]; buf[Ng / 2].im = 0; }
          fft_exec(&p, buf, buf + Ng);
          real *o = out + row * Ng;
          for (int x = 0; x < Ng; x++) o[x] = (real) buf[x].re;
        }
        free(buf);
      }
    }
  }
  plan_free(&p);
  free(w);
  return fail;
}

int orc_fft_r2c_3d(int Ng, const real *in, real *out, int nthread) {
  const int Ngk = Ng / 2 + 1;
  plan1d p;
  if (plan_make(&p, Ng, -1)) return 1;
  int fail = 0;
#pragma omp parallel num_threads(nthread)
  {
    cplx *buf = malloc(sizeof(cplx) * Ng * 2);
    if (!buf) {
#pragma omp atomic write
      fail = 1;
    }
    else {
#pragma omp for schedule(static)
      for (size_t row = 0; row < (size_t) Ng * Ng; row++) {
        const real *r = in + row * Ng;
        for (int x = 0; x < Ng; x++) { buf[x].re = r[x]; buf[x].im = 0; }
        fft_exec(&p, buf, buf + Ng);
        real *o = out + 2 * row * Ngk;
        for (int k = 0; k < Ngk; k++) {
          o[2 * k] = (real) buf[k].re;
          o[2 * k + 1] = (real) buf[k].im;
        }
      }
      free(buf);
    }
  }
  if (!fail) fail = axis_pass(out, Ng, Ng, Ngk, 1, &p, nthread) ||
                    axis_pass(out, Ng, Ng, Ngk, 0, &p, nthread);
  plan_free(&p);
  return fail;
}
