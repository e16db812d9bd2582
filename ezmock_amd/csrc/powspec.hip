// powspec.hip -- power-spectrum multipoles of a device-resident catalogue.
//
// The reference fits its mocks with an external FFT estimator (interlaced
// CIC assignment, monopole and quadrupole; example/fit/conf/powspec.conf and
// the clustering step of example/fit/fit_EZmock.ipynb).  This is the on-GPU
// counterpart for the calibration loop:
//   k_pk_paint  CIC assignment of the tracers (optionally shifted by half a
//               cell for interlacing), redshift-space z = z + f vz wrapped
//               once into [0, L) as save_res.c:147-158 does;
//   r2c of each grid (fft.hip);
//   k_pk_bin    |delta_k|^2 per mode, CIC window compensation, interlacing
//               phase, Legendre weights L0 and L2 of mu = k_z / k, binned
//               in LDS (double) and reduced with one device atomic per block
//               and bin.
#include "common.h"
#include "kernels.h"

__global__ void k_pk_paint(const real *__restrict__ x, const real *__restrict__ y,
    const real *__restrict__ z, const real *__restrict__ vz, size_t n, int Ng,
    double Lbox, double rsd_fac, double shift, real *__restrict__ grid) {
  const size_t t = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double p[3] = {(double) x[t], (double) y[t], (double) z[t]};
  if (rsd_fac != 0) {
    double zr = p[2] + rsd_fac * (double) vz[t];
    if (zr >= Lbox) zr -= Lbox;
    if (zr < 0) zr += Lbox;
    p[2] = zr;
  }
  const double igs = (double) Ng / Lbox;
  int c0[3], c1[3];
  real w0[3], w1[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double u = p[a] * igs + shift;
    const double f = floor(u);
    const double d = u - f;
    int i = (int) f % Ng;
    if (i < 0) i += Ng;
    c0[a] = i;
    c1[a] = (i + 1 == Ng) ? 0 : i + 1;
    w1[a] = (real) d;
    w0[a] = (real) (1 - d);
  }
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const size_t ix = (size_t) (a ? c1[0] : c0[0]);
    const real wx = a ? w1[0] : w0[0];
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const size_t iy = (size_t) (b ? c1[1] : c0[1]);
      const real wxy = wx * (b ? w1[1] : w0[1]);
#pragma unroll
      for (int d = 0; d < 2; d++) {
        const size_t iz = (size_t) (d ? c1[2] : c0[2]);
        atomicAdd(&grid[(ix * Ng + iy) * Ng + iz], wxy * (d ? w1[2] : w0[2]));
      }
    }
  }
}

hipError_t launch_pk_paint(const real *x, const real *y, const real *z,
    const real *vz, size_t n, int Ng, double Lbox, double rsd_fac,
    double shift, real *grid, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pk_paint, dim3(ez_blocks(n, 256)), dim3(256), 0, st, x, y,
      z, vz, n, Ng, Lbox, rsd_fac, shift, grid);
  return hipGetLastError();
}

__device__ __forceinline__ double sinc_pi(int n, int Ng) {
  if (n == 0) return 1.0;
  const double a = EZ_PI * (double) n / (double) Ng;
  return sin(a) / a;
}

#define PK_MAX_BIN 1024
#define PK_TPB 256

// acc[4 * b + {0,1,2,3}] += {w, w k, w P, w P L2(mu)}
__global__ void __launch_bounds__(PK_TPB)
k_pk_bin(const cplx *__restrict__ X1, const cplx *__restrict__ N1,
    const cplx *__restrict__ X2, const cplx *__restrict__ N2, int Ng,
    double kf, double dk, int nbin, double pnorm, double *__restrict__ acc) {
  __shared__ double h[4 * PK_MAX_BIN];
  for (int i = threadIdx.x; i < 4 * nbin; i += PK_TPB) h[i] = 0;
  __syncthreads();
  const int Ngh = Ng >> 1, Ngk = Ngh + 1;
  const size_t nmode = (size_t) Ng * Ng * Ngk;
  const double kmax = dk * nbin;
  for (size_t m = (size_t) blockIdx.x * PK_TPB + threadIdx.x; m < nmode;
       m += (size_t) gridDim.x * PK_TPB) {
    const int kk = (int) (m % Ngk);
    const size_t ij = m / Ngk;
    const int j = (int) (ij % Ng), i = (int) (ij / Ng);
    const int ni = (i <= Ngh) ? i : i - Ng, nj = (j <= Ngh) ? j : j - Ng;
    const double kx = kf * ni, ky = kf * nj, kz = kf * kk;
    const double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 == 0) continue;
    const double k = sqrt(k2);
    const int b = (int) (k / dk);
    if (k >= kmax || b >= nbin) continue;
    cplx a = (kk < Ngh) ? X1[ij * Ngh + kk] : N1[ij];
    double re = a.x, im = a.y;
    if (X2) {
      const cplx c = (kk < Ngh) ? X2[ij * Ngh + kk] : N2[ij];
      const double th = EZ_PI * (double) (ni + nj + kk) / (double) Ng;
      const double cs = cos(th), sn = sin(th);
      re = 0.5 * (re + (double) c.x * cs - (double) c.y * sn);
      im = 0.5 * (im + (double) c.x * sn + (double) c.y * cs);
    }
    const double wx = sinc_pi(ni, Ng), wy = sinc_pi(nj, Ng), wz = sinc_pi(kk, Ng);
    const double W = wx * wx * wy * wy * wz * wz;
    const double P = (re * re + im * im) * pnorm / (W * W);
    const double mu = kz / k;
    const double L2 = 1.5 * mu * mu - 0.5;
    const double w = (kk == 0 || kk == Ngh) ? 1.0 : 2.0;
    atomicAdd(&h[4 * b], w);
    atomicAdd(&h[4 * b + 1], w * k);
    atomicAdd(&h[4 * b + 2], w * P);
    atomicAdd(&h[4 * b + 3], w * P * L2);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 4 * nbin; q += PK_TPB)
    if (h[q] != 0) atomicAdd(&acc[q], h[q]);
}

hipError_t launch_pk_bin(const cplx *X1, const cplx *N1, const cplx *X2,
    const cplx *N2, int Ng, double kf, double dk, int nbin, double pnorm,
    double *acc, hipStream_t st) {
  if (nbin <= 0 || nbin > PK_MAX_BIN) return hipErrorInvalidValue;
  const size_t nmode = (size_t) Ng * Ng * (Ng / 2 + 1);
  unsigned nb = ez_blocks(nmode, PK_TPB * 16);
  if (nb > 1024) nb = 1024;
  if (nb == 0) nb = 1;
  hipLaunchKernelGGL(k_pk_bin, dim3(nb), dim3(PK_TPB), 0, st, X1, N1, X2, N2, Ng,
      kf, dk, nbin, pnorm, acc);
  return hipGetLastError();
}
