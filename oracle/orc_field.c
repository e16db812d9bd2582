/*
 * orc_field.c -- set-up and density-field stages of the CPU oracle
 * (test infrastructure only).
 *
 * Follows, stage by stage:
 *   EZmock_init / EZmock_destroy ........... EZmock.c:41-142
 *   EZmock_set_cosmology, cosmo_growth ...... cosmology.c:53-157
 *   EZmock_setup_linear_pk, pk_interp ....... linear_pk.c:36-202
 *   EZmock_ZA_disp* (random initial field) .. perturb.c:78-399
 *   EZmock_ZA_disp_wn* (white-noise input) .. perturb.c:456-604
 *   density_field_cic ....................... dens_field.c:233-310
 *   EZmock_create_dens_field ................ dens_field.c:337-432
 *   EZmock_errmsg ........................... errmsg.c:25-70
 */
#include "orc_internal.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

/* --------------------------------------------------------- lifecycle --- */

/* EZmock.c:41-124 */
orc_ez *orc_init(double Lbox, int Ngrid, int randgen, uint64_t seed,
    int nthread, int *err) {
  if (!err || *err != ORC_SUCCESS) return NULL;
  if (Lbox <= 0) { *err = ORC_ERR_ARG_LBOX; return NULL; }
  if (Ngrid <= 0 || Ngrid > ORC_MAX_GRID) { *err = ORC_ERR_ARG_NGRID; return NULL; }
  if (randgen != ORC_RNG_MRG32K3A && randgen != ORC_RNG_MT19937) {
    *err = ORC_ERR_ARG_RNG; return NULL;
  }
  if (seed == 0) { *err = ORC_ERR_ARG_SEED; return NULL; }
  if (nthread < 0) { *err = ORC_ERR_ARG_THREAD; return NULL; }
  orc_ez *ez = calloc(1, sizeof(orc_ez));
  if (!ez) { *err = ORC_ERR_MEMORY; return NULL; }
  ez->nthread = nthread ? nthread : omp_get_max_threads();
  ez->Ng = Ngrid;
  ez->Lbox = Lbox;
  ez->rngtype = randgen;
  ez->seed = seed;
  ez->growth2 = ez->vfac = HUGE_VAL;
  ez->bao = ez->rho_c = ez->rho_exp = ez->pdf_base = ez->sigma_v = HUGE_VAL;
  return ez;
}

static void mesh_free(orc_ez *ez) {
  if (!ez->psi_ref) for (int i = 0; i < 3; i++) free(ez->psi[i]);
  for (int i = 0; i < 3; i++) ez->psi[i] = NULL;
  if (ez->rho && ez->rho != ez->rhok) free(ez->rho);
  free(ez->rhok); free(ez->rhok2); free(ez->rhot); free(ez->spec);
  free(ez->ntbuf); free(ez->npbuf);
  ez->rho = ez->rhok = ez->rhok2 = ez->rhot = ez->spec = NULL;
  ez->ntbuf = NULL; ez->npbuf = NULL;
}

/* EZmock.c:133-142 */
void orc_destroy(orc_ez *ez) {
  if (!ez) return;
  mesh_free(ez);
  free(ez->pk_k); free(ez->pk_P); free(ez->pk_P2);
  free(ez->pdf_bins);
  free(ez);
}

void orc_free(void *p) { free(p); }

/* errmsg.c:25-70 */
const char *orc_errmsg(int err) {
  static const char *msg[] = {
    "no error", "failed to allocate memory", "error code is not set",
    "EZmock interface uninitialized", "EZmock cosmology uninitialized",
    "EZmock power spectrum uninitialized",
    "EZmock density fields uninitialized", "invalid box size",
    "invalid grid size", "invalid random number generator",
    "invalid random seed", "invalid number of threads", "null input array",
    "invalid number of bins", "EZmock calibration parameters not set",
    "too few expected number of tracers",
    "too many expected number of tracers",
    "cannot return the number of tracers", "invalid file name",
    "invalid RSD factor", "cannot obtain cosmic growth parameters",
    "infinite value detected", "non-positive value detected",
    "non-ascending k array detected", "k range insufficient for density field",
    "too few k bins available for the field",
    "invalid P(k) after BAO modification",
    "failed to setup the random generator",
    "random generator reseting failed",
    "random generator jumping ahead failed",
    "invalid cosmological parameters", "invalid density threshold parameter",
    "invalid exponential bias parameter",
    "invalid base number for PDF mapping",
    "invalid parameter for random motions",
    "the density field has been overwritten",
    "not enough cells to host all tracers",
    "PDF implies too many tracers per cell", "too many particles per cell",
    "failed to initialise FFTW with OpenMP",
    "failed to write to the output file", "unknown error"
  };
  if (err < 0 || err > ORC_ERR_UNKNOWN) return "unregistered error code";
  return msg[err];
}

/* --------------------------------------------------------- cosmology --- */

/* cosmology.c:53-95: D(a)/D(ainit) and f(a) in flat wCDM via 2F1. */
static int growth(double a, double ainit, double Om0, double w, double *D,
    double *f) {
  if (a <= 0 || a > 1 || ainit <= 0 || ainit > 1 || Om0 <= 0 || Om0 > 1 ||
      w >= -1.0 / 3.0) return ORC_ERR_PAR_COSMO;
  const double iw = 1 / w;
  const double c1 = 0.5 * (1 - iw);
  const double c2 = -0x1.5555555555555p-2 * iw;
  const double c3 = 1 - 0x1.aaaaaaaaaaaabp-1 * iw;
  const double x = (1 - 1 / Om0) * pow(a, -3 * w);
  double h1, h2;
  if (orc_hyp2f1(c1, c2, c3, x, 1e-12, 10000, &h1) ||
      orc_hyp2f1(c1 + 1, c2 + 1, c3 + 1, x, 1e-12, 10000, &h2))
    return ORC_ERR_COS_GROWTH;
  *f = 1 + 3 * (w - 1) * x / (6 * w - 5) * h2 / h1;
  if (a == ainit) { *D = 1; return 0; }
  const double x0 = (1 - 1 / Om0) * pow(ainit, -3 * w);
  if (orc_hyp2f1(c1, c2, c3, x0, 1e-12, 10000, &h2)) return ORC_ERR_COS_GROWTH;
  *D = (a * h1) / (ainit * h2);
  return 0;
}

/* cosmology.c:121-157 */
int orc_set_cosmology(orc_ez *ez, double pk_norm, double fHa, bool eval,
    double z, double z_pk, double Omega_m, double Omega_nu, double w,
    int *err) {
  if (!err) return ORC_ERR_ARG_ECODE;
  if (*err != ORC_SUCCESS) return *err;
  if (!ez) return ORC_ERR_ARG_EZ;           /* *err untouched, as upstream */
  if (!eval) {
    if (pk_norm <= 0 || fHa <= 0) return (*err = ORC_ERR_PAR_COSMO);
    ez->growth2 = pk_norm;
    ez->vfac = fHa;
    return ORC_SUCCESS;
  }
  if (z < 0 || z_pk < 0 || Omega_m <= 0 || Omega_m > 1 || Omega_nu < 0 ||
      Omega_nu >= 1 || w >= -1.0 / 3.0) return (*err = ORC_ERR_PAR_COSMO);
  double a = 1 / (1 + z), a_pk = 1 / (1 + z_pk), D, f;
  int e = growth(a, a_pk, Omega_m, w, &D, &f);
  if (e) return (*err = e);
  const double Omn = Omega_m + Omega_nu;
  const double H = 100 * sqrt((1 - Omn) * pow(a, -3 * (1 + w)) +
      Omn * pow(a, -3));
  ez->growth2 = D * D;
  ez->vfac = f * H * a;
  return ORC_SUCCESS;
}

/* -------------------------------------------------------- linear P(k) --- */

/* linear_pk.c:75-183 */
int orc_setup_linear_pk(orc_ez *ez, const double *k, int n, const double *Pk,
    const double *Pnw, double mBAO, bool logint, int *err) {
  if (!err) return ORC_ERR_ARG_ECODE;
  if (*err != ORC_SUCCESS) return *err;
  if (!ez) return (*err = ORC_ERR_ARG_EZ);
  if (!k || !Pk || (mBAO != 0 && !Pnw)) return (*err = ORC_ERR_ARG_NULL);
  if (n <= 3 || n > ORC_MAX_PK_NBIN) return (*err = ORC_ERR_ARG_NBIN);
  if (ez->growth2 == HUGE_VAL) return (*err = ORC_ERR_ARG_COSMO);
  for (int i = 0; i < n; i++) {
    if (!isfinite(k[i]) || !isfinite(Pk[i])) return (*err = ORC_ERR_PK_FINITE);
    if (k[i] <= 0 || Pk[i] <= 0) return (*err = ORC_ERR_PK_NONPOS);
    if (i && k[i] <= k[i - 1]) return (*err = ORC_ERR_PK_NONASC);
  }
  if (mBAO != 0) {
    for (int i = 0; i < n; i++) {
      if (!isfinite(Pnw[i])) return (*err = ORC_ERR_PK_FINITE);
      if (Pnw[i] <= 0) return (*err = ORC_ERR_PK_NONPOS);
    }
  }
  const double kmin = 2 * ORC_PI / ez->Lbox;
  const double kmax = ORC_PI * ez->Ng / ez->Lbox * sqrt(3);
  if (k[0] > kmin - ORC_DOUBLE_TOL || k[n - 1] < kmax + ORC_DOUBLE_TOL)
    return (*err = ORC_ERR_PK_RANGE);
  int lo, hi;
  for (lo = 0; lo < n; lo++) if (k[lo] > kmin) break;
  for (hi = lo; hi < n; hi++) if (k[hi] > kmax) break;
  lo = (lo < 3) ? 0 : lo - 3;
  hi = (hi + 3 >= n) ? n : hi + 3;
  const int nb = hi - lo;
  if (nb < 3) return (*err = ORC_ERR_PK_NBIN);
  double *kk = malloc(sizeof(double) * nb), *PP = malloc(sizeof(double) * nb),
         *P2 = malloc(sizeof(double) * nb * 2);
  if (!kk || !PP || !P2) {
    free(kk); free(PP); free(P2); return (*err = ORC_ERR_MEMORY);
  }
  for (int i = 0; i < nb; i++) {
    const int j = i + lo;
    if (mBAO == 0) PP[i] = Pk[j] * ez->growth2;
    else {
      PP[i] = (k[j] <= ORC_MAX_K_BAO) ?
          (Pk[j] - Pnw[j]) * exp(k[j] * k[j] * mBAO) + Pnw[j] : Pk[j];
      if (!isfinite(PP[i]) || PP[i] <= 0) {
        free(kk); free(PP); free(P2); return (*err = ORC_ERR_PK_MODBAO);
      }
      PP[i] *= ez->growth2;
    }
  }
  for (int i = 0; i < nb; i++) {
    if (logint) { kk[i] = log(k[i + lo]); PP[i] = log(PP[i]); }
    else kk[i] = k[i + lo];
  }
  orc_cspline_ypp(kk, PP, nb, P2);
  free(ez->pk_k); free(ez->pk_P); free(ez->pk_P2);
  ez->pk_k = kk; ez->pk_P = PP; ez->pk_P2 = P2;
  ez->pk_n = nb;
  ez->pk_log = logint;
  ez->pk_kmin = k[lo];
  ez->pk_kmax = k[hi - 1];
  ez->pk_set = 1;
  ez->bao = mBAO;
  return ORC_SUCCESS;
}

/* linear_pk.c:36-51: index i with x[i] <= v < x[i+1]; n-1 if v == x[n-1];
   -1 when v is outside the table. */
static int bin_search(const double *x, int n, double v) {
  int l = 0, u = n - 1;
  while (l <= u) {
    int i = (l + u) >> 1;
    if (i >= n - 1) return (x[n - 1] == v) ? n - 1 : -1;
    if (x[i + 1] <= v) l = i + 1;
    else if (x[i] > v) u = i - 1;
    else return i;
  }
  return -1;
}

/* linear_pk.c:194-202 */
double orc_pk_eval(const orc_ez *ez, double k) {
  const int n = ez->pk_n;
  const int i = bin_search(ez->pk_k, n, k);
  if (i == n - 1) return ez->pk_P[i];
  if (i < 0) {
    if (k >= ez->pk_k[n - 1]) return ez->pk_P[n - 1];
    if (k <= ez->pk_k[0]) return ez->pk_P[0];
  }
  return orc_cspline_eval(ez->pk_k, ez->pk_P, ez->pk_P2, k, (size_t) i);
}

double orc_pk_interp(const orc_ez *ez, double k) {
  return (ez && ez->pk_set) ? orc_pk_eval(ez, k) : HUGE_VAL;
}

size_t orc_pk_nbin(const orc_ez *ez) { return ez->pk_set ? ez->pk_n : 0; }

int orc_pk_table(const orc_ez *ez, double *k, double *P, double *P2) {
  if (!ez->pk_set) return 1;
  memcpy(k, ez->pk_k, sizeof(double) * ez->pk_n);
  memcpy(P, ez->pk_P, sizeof(double) * ez->pk_n);
  memcpy(P2, ez->pk_P2, sizeof(double) * ez->pk_n);
  return 0;
}

/* -------------------------------------------------- initial condition --- */

/* Shared part of the per-cell spectrum: sqrt(P(k) * (-ln u_amp)) * L^-1.5
   / (|n|^2 k_f), perturb.c:171-201 (and the 0<k<Ny and Nyquist copies). */
static inline double mode_amp(const orc_ez *ez, double ksq, double amp,
    bool fixamp, double kfac, double logkfac, double Pfac) {
  double kmod = ez->pk_log ? log(ksq) * 0.5 + logkfac : kfac * sqrt(ksq);
  double P = orc_pk_eval(ez, kmod);
  if (!fixamp) amp = -log(amp);
  if (ez->pk_log) P = fixamp ? exp(P * 0.5) : exp(P * 0.5) * sqrt(amp);
  else P = fixamp ? sqrt(P) : sqrt(P * amp);
  P *= Pfac;
  P /= (ksq * kfac);
  return P;
}

/* perturb.c:78-359: rhok (the i*delta/(k_f |n|^2) spectrum) and rhok2 =
   rhok * n_x on the half-complex grid, with the reference's OpenMP row
   partition and per-thread stream jumps (the draws are thread-count
   independent: cell idx >= 1 uses draws 2(idx-1), 2(idx-1)+1, or idx-1
   with fixed amplitudes). */
static int gen_rhok(orc_ez *ez, bool fixamp, bool iphase) {
  const int Ng = ez->Ng, Ngh = Ng >> 1, Ngk = Ngh + 1;
  const size_t rows = (size_t) Ng * Ng;
  const size_t pnum = rows / ez->nthread;
  const int rem = rows % ez->nthread;
  const double kfac = ORC_PI * 2 / ez->Lbox;
  const double logkfac = log(kfac);
  const double Pfac = pow(ez->Lbox, -1.5);
  real *rk = ez->rhok, *rk2 = ez->rhok2;
  int fail = 0;
#pragma omp parallel num_threads(ez->nthread)
  {
    const int tid = omp_get_thread_num();
    const size_t pcnt = (tid < rem) ? pnum + 1 : pnum;
    const size_t istart = (tid < rem) ? pcnt * tid : pnum * tid + rem;
    const size_t iend = istart + pcnt;
    orc_rng_t rng;
    uint64_t step = (tid == 0) ? 0 : (uint64_t) (istart * Ngk - 1);
    if (!fixamp) step <<= 1;
    if (orc_rng_reset(&rng, ez->rngtype, ez->seed, step)) {
#pragma omp atomic write
      fail = 1;
    }
#pragma omp barrier
    if (!fail) {
      for (size_t ij = istart; ij < iend; ij++) {
        const int i = ij / Ng, j = ij % Ng;
        const double ki = (i <= Ngh) ? i : i - Ng;
        const int ni = (i == 0) ? 0 : Ng - i;
        const double kni = (ni <= Ngh) ? ni : ni - Ng;
        const double kj = (j <= Ngh) ? j : j - Ng;
        const int nj = (j == 0) ? 0 : Ng - j;
        const double kij = ki * ki + kj * kj;
        const size_t idx0 = ij * Ngk;
        const bool lower = (i == ni && j <= Ngh) || (i != ni && i <= Ngh);
        /* kz = 0 plane (and kz = Ny below): Hermitian pairs written together
           from the lower half plane. */
        if (idx0 == 0) {
          rk[0] = rk[1] = 0;
          rk2[0] = rk2[1] = 0;
        }
        else {
          double amp = fixamp ? 1 : orc_rng_double_pos(&rng);
          double ph = orc_rng_double(&rng) * 2 * ORC_PI;
          if (lower) {
            const size_t idx = idx0, nidx = ((size_t) ni * Ng + nj) * Ngk;
            double P = mode_amp(ez, kij, amp, fixamp, kfac, logkfac, Pfac);
            if (iphase) ph += ORC_PI;
            rk[2 * idx] = -P * sin(ph);
            rk[2 * idx + 1] = P * cos(ph);
            rk2[2 * idx] = rk[2 * idx] * ki;
            rk2[2 * idx + 1] = rk[2 * idx + 1] * ki;
            rk[2 * nidx] = -rk[2 * idx];
            rk[2 * nidx + 1] = rk[2 * idx + 1];
            rk2[2 * nidx] = rk[2 * nidx] * kni;
            rk2[2 * nidx + 1] = rk[2 * nidx + 1] * kni;
          }
        }
        for (int k = 1; k <= ((Ng - 1) >> 1); k++) {
          double amp = fixamp ? 1 : orc_rng_double_pos(&rng);
          double ph = orc_rng_double(&rng) * 2 * ORC_PI;
          const size_t idx = idx0 + k;
          const double ksq = kij + k * k;
          double P = mode_amp(ez, ksq, amp, fixamp, kfac, logkfac, Pfac);
          if (iphase) ph += ORC_PI;
          rk[2 * idx] = -P * sin(ph);
          rk[2 * idx + 1] = P * cos(ph);
          rk2[2 * idx] = rk[2 * idx] * ki;
          rk2[2 * idx + 1] = rk[2 * idx + 1] * ki;
        }
        if ((Ng & 1) == 0) {
          double amp = fixamp ? 1 : orc_rng_double_pos(&rng);
          double ph = orc_rng_double(&rng) * 2 * ORC_PI;
          if (lower) {
            const int k = Ngh;
            const size_t idx = idx0 + k;
            const size_t nidx = ((size_t) ni * Ng + nj) * Ngk + k;
            const double ksq = kij + k * k;
            double P = mode_amp(ez, ksq, amp, fixamp, kfac, logkfac, Pfac);
            if (iphase) ph += ORC_PI;
            rk[2 * idx] = -P * sin(ph);
            rk[2 * idx + 1] = P * cos(ph);
            rk2[2 * idx] = rk[2 * idx] * ki;
            rk2[2 * idx + 1] = rk[2 * idx + 1] * ki;
            rk[2 * nidx] = -rk[2 * idx];
            rk[2 * nidx + 1] = rk[2 * idx + 1];
            rk2[2 * nidx] = rk[2 * nidx] * kni;
            rk2[2 * nidx + 1] = rk[2 * nidx + 1] * kni;
          }
        }
      }
    }
  }
  if (fail) return ORC_ERR_RNG_JUMP;
  /* perturb.c:341-359: real parts vanish at the 7 self-conjugate modes. */
  if ((Ng & 1) == 0) {
    const size_t pts[7] = {
      (size_t) Ngh, (size_t) Ngh * Ngk, (size_t) Ngh * (Ngk + 1),
      (size_t) Ngh * Ng * Ngk, (size_t) Ngh * ((size_t) Ng * Ngk + 1),
      (size_t) Ngh * (Ng + 1) * Ngk, ((size_t) Ngh * Ng + Ngh) * Ngk + Ngh };
    for (int p = 0; p < 7; p++) rk[2 * pts[p]] = rk2[2 * pts[p]] = 0;
  }
  return 0;
}

/* perturb.c:361-398: psi_x = c2r(rhok2); rhok2 = rhok*n_y, rhok *= n_z;
   psi_y = c2r(rhok2); psi_z = c2r(rhok). */
static int spectra_to_psi(orc_ez *ez) {
  const int Ng = ez->Ng, Ngh = Ng >> 1, Ngk = Ngh + 1;
  double t0 = omp_get_wtime();
  if (orc_fft_c2r_3d(Ng, ez->rhok2, ez->psi[0], ez->nthread)) return ORC_ERR_MEMORY;
  ez->tstage[ORC_ST_FFT] += omp_get_wtime() - t0;
  t0 = omp_get_wtime();
  const size_t rows = (size_t) Ng * Ng;
  real *rk = ez->rhok, *rk2 = ez->rhok2;
#pragma omp parallel for num_threads(ez->nthread) schedule(static)
  for (size_t ij = 0; ij < rows; ij++) {
    const int j = ij % Ng;
    const double kj = (j <= Ngh) ? j : j - Ng;
    for (int k = 0; k < Ngk; k++) {
      const size_t idx = ij * Ngk + k;
      rk2[2 * idx] = rk[2 * idx] * kj;
      rk2[2 * idx + 1] = rk[2 * idx + 1] * kj;
      rk[2 * idx] *= k;
      rk[2 * idx + 1] *= k;
    }
  }
  ez->tstage[ORC_ST_NOISE] += omp_get_wtime() - t0;
  t0 = omp_get_wtime();
  if (orc_fft_c2r_3d(Ng, ez->rhok2, ez->psi[1], ez->nthread) ||
      orc_fft_c2r_3d(Ng, ez->rhok, ez->psi[2], ez->nthread))
    return ORC_ERR_MEMORY;
  ez->tstage[ORC_ST_FFT] += omp_get_wtime() - t0;
  return 0;
}

/* perturb.c:456-554: white-noise input; rhok2 holds r2c(delta). */
static int gen_rhok_wn(orc_ez *ez) {
  const int Ng = ez->Ng, Ngk = (Ng >> 1) + 1;
  const size_t rows = (size_t) Ng * Ng;
  const double kfac = ORC_PI * 2 / ez->Lbox;
  const double logkfac = log(kfac);
  const double Pfac = pow(ez->Lbox, -1.5);
  real *rk = ez->rhok, *rk2 = ez->rhok2;
#pragma omp parallel for num_threads(ez->nthread) schedule(static)
  for (size_t ij = 0; ij < rows; ij++) {
    const int i = ij / Ng, j = ij % Ng;
    const double ki = (i <= (Ng >> 1)) ? i : i - Ng;
    const double kj = (j <= (Ng >> 1)) ? j : j - Ng;
    const double kij = ki * ki + kj * kj;
    for (int k = 0; k < Ngk; k++) {
      const size_t idx = ij * Ngk + k;
      if (idx == 0) { rk[0] = rk[1] = rk2[0] = rk2[1] = 0; continue; }
      const double ksq = kij + k * k;
      double kmod = ez->pk_log ? log(ksq) * 0.5 + logkfac : kfac * sqrt(ksq);
      double P = orc_pk_eval(ez, kmod);
      P = ez->pk_log ? exp(P * 0.5) : sqrt(P);
      P *= Pfac / (ksq * kfac);
      rk[2 * idx] = -P * rk2[2 * idx + 1];
      rk[2 * idx + 1] = P * rk2[2 * idx];
      rk2[2 * idx] = rk[2 * idx] * ki;
      rk2[2 * idx + 1] = rk[2 * idx + 1] * ki;
    }
  }
  return 0;
}

/* --------------------------------------------------------------- CIC --- */

/* dens_field.c:233-310 and pop_tracer.c:593-675 (dim >= 0 weights every
   contribution by psi[dim]). */
void orc_cic_paint(orc_ez *ez, int dim, real *out) {
  const int Ng = ez->Ng;
  const real igs = ez->Ng / ez->Lbox;
  const size_t rows = (size_t) Ng * Ng;
  const size_t pnum = rows / ez->nthread;
  const int rem = rows % ez->nthread;
#pragma omp parallel num_threads(ez->nthread)
  {
    const int tid = omp_get_thread_num();
    const size_t pcnt = (tid < rem) ? pnum + 1 : pnum;
    const size_t istart = (tid < rem) ? pcnt * tid : pnum * tid + rem;
    const size_t iend = istart + pcnt;
    for (size_t ij = istart; ij < iend; ij++) {
      const int i = ij / Ng, j = ij % Ng;
      for (int k = 0; k < Ng; k++) {
        const size_t idx = ij * Ng + k;
        real x = ez->psi[0][idx] * igs + i;
        real y = ez->psi[1][idx] * igs + j;
        real z = ez->psi[2][idx] * igs + k;
        int x0 = (int) floor(x), y0 = (int) floor(y), z0 = (int) floor(z);
        real wx1 = x - x0, wx0 = 1 - wx1;
        real wy1 = y - y0, wy0 = 1 - wy1;
        real wz1 = z - z0, wz0 = 1 - wz1;
        if (dim >= 0) { wx0 *= ez->psi[dim][idx]; wx1 *= ez->psi[dim][idx]; }
        if (x0 < 0) x0 += Ng; else if (x0 >= Ng) x0 -= Ng;
        if (y0 < 0) y0 += Ng; else if (y0 >= Ng) y0 -= Ng;
        if (z0 < 0) z0 += Ng; else if (z0 >= Ng) z0 -= Ng;
        const int x1 = (x0 == Ng - 1) ? 0 : x0 + 1;
        const int y1 = (y0 == Ng - 1) ? 0 : y0 + 1;
        const int z1 = (z0 == Ng - 1) ? 0 : z0 + 1;
#define ORC_ADD(a, b, c, w) { real v_ = (w); \
  _Pragma("omp atomic") out[ORC_IDX(Ng, a, b, c)] += v_; }
        ORC_ADD(x0, y0, z0, wx0 * wy0 * wz0);
        ORC_ADD(x0, y0, z1, wx0 * wy0 * wz1);
        ORC_ADD(x0, y1, z0, wx0 * wy1 * wz0);
        ORC_ADD(x0, y1, z1, wx0 * wy1 * wz1);
        ORC_ADD(x1, y0, z0, wx1 * wy0 * wz0);
        ORC_ADD(x1, y0, z1, wx1 * wy0 * wz1);
        ORC_ADD(x1, y1, z0, wx1 * wy1 * wz0);
        ORC_ADD(x1, y1, z1, wx1 * wy1 * wz1);
#undef ORC_ADD
      }
    }
  }
}

/* ------------------------------------------------ create_dens_field --- */

/* dens_field.c:337-432 */
int orc_create_dens_field(orc_ez *ez, real *psi[3], bool deepcopy,
    real *delta, bool fixamp, bool iphase, int *err) {
  if (!err) return ORC_ERR_ARG_ECODE;
  if (*err != ORC_SUCCESS) return *err;
  if (!ez) return (*err = ORC_ERR_ARG_EZ);
  const int Ng = ez->Ng;
  const size_t num = (size_t) Ng * Ng * Ng;
  const size_t size = num * sizeof(real);
  if (ez->psi_ref) ez->psi[0] = ez->psi[1] = ez->psi[2] = NULL;
  if (psi && psi[0] && psi[1] && psi[2]) {
    if (!ez->rho || ez->rho == ez->rhok) {
      if (!(ez->rho = malloc(size))) return (*err = ORC_ERR_MEMORY);
    }
    if (deepcopy) {
      for (int i = 0; i < 3; i++) {
        if (!ez->psi[i] && !(ez->psi[i] = malloc(size)))
          return (*err = ORC_ERR_MEMORY);
        memcpy(ez->psi[i], psi[i], size);
      }
      ez->psi_ref = false;
    }
    else {
      for (int i = 0; i < 3; i++) { free(ez->psi[i]); ez->psi[i] = psi[i]; }
      ez->psi_ref = true;
    }
  }
  else {
    if (!ez->pk_set) return (*err = ORC_ERR_ARG_PK);
    const double kmin = 2 * ORC_PI / ez->Lbox;
    const double kmax = ORC_PI * ez->Ng / ez->Lbox * sqrt(3);
    if (ez->pk_kmin > kmin - ORC_DOUBLE_TOL || ez->pk_kmax < kmax + ORC_DOUBLE_TOL)
      return (*err = ORC_ERR_PK_RANGE);
    const int Ngk = (Ng >> 1) + 1;
    const size_t csize = (size_t) Ng * Ng * Ngk * 2 * sizeof(real);
    for (int i = 0; i < 3; i++)
      if (!ez->psi[i] && !(ez->psi[i] = malloc(size))) return (*err = ORC_ERR_MEMORY);
    if (ez->rho && ez->rho != ez->rhok) { free(ez->rho); ez->rho = NULL; }
    if ((!ez->rhok && !(ez->rhok = malloc(csize))) ||
        (!ez->rhok2 && !(ez->rhok2 = malloc(csize))))
      return (*err = ORC_ERR_MEMORY);
    ez->psi_ref = false;
    double t0 = omp_get_wtime();
    int e;
    if (delta) {
      if (orc_fft_r2c_3d(Ng, delta, ez->rhok2, ez->nthread)) return (*err = ORC_ERR_MEMORY);
      e = gen_rhok_wn(ez);
    }
    else e = gen_rhok(ez, fixamp, iphase);
    if (e) return (*err = e);
    ez->tstage[ORC_ST_NOISE] += omp_get_wtime() - t0;
    if (ez->keep_spec) {
      free(ez->spec);
      if (!(ez->spec = malloc(csize))) return (*err = ORC_ERR_MEMORY);
      memcpy(ez->spec, ez->rhok, csize);
    }
    if ((e = spectra_to_psi(ez))) return (*err = e);
    if (delta) {
      /* perturb.c:595-603 */
      const real norm = 1 / sqrt((double) num);
      for (int j = 0; j < 3; j++)
        for (size_t i = 0; i < num; i++) ez->psi[j][i] *= norm;
    }
    ez->rho = ez->rhok;
  }
  double t0 = omp_get_wtime();
  memset(ez->rho, 0, size);
  orc_cic_paint(ez, -1, ez->rho);
  ez->tstage[ORC_ST_CIC] += omp_get_wtime() - t0;
  ez->fixamp = fixamp;
  ez->iphase = iphase;
  ez->have_rho = 1;
  return ORC_SUCCESS;
}
