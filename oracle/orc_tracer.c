/*
 * orc_tracer.c -- tracer-population stages of the CPU oracle (test
 * infrastructure only).
 *
 * Follows pop_tracer.c stage by stage:
 *   gauss_rand ............ :70-79     eval_pdf .............. :94-133
 *   bias_model ............ :149-348   n_tracer_in_cell ...... :359-583
 *   disp_cic .............. :593-683   assign_vel ............ :697-753
 *   generate_tracers ...... :773-1027  add_vel_scatter ....... :1043-1113
 *   EZmock_populate_tracer  :1140-1219
 * plus the redshift-space shift of save_res.c:147-158.
 *
 * Known reference quirks kept or replaced (each cited where it matters):
 *  - a thread chunk without any valid cell leaves ifirst[tid] = 0 in the
 *    OpenMP bias_model (pop_tracer.c:196-211); the oracle uses the intended
 *    thread-count-independent pairing instead (valid cell of rank r uses the
 *    Box-Muller pair r/2 drawn at 2N + 2*(r/2)).
 *  - the PDF walk reads cidx[cnt-1] with cnt = 0 when the first PDF bin is
 *    empty and its histogram bin is not (pop_tracer.c:433); the oracle takes
 *    that case as "no ties".
 */
#include "orc_internal.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

/* pop_tracer.c:70-79 */
static inline void gauss_pair(orc_rng_t *r, double sigma, double *g1,
    double *g2) {
  double rr = -2 * log(orc_rng_double_pos(r));
  double th = orc_rng_double(r) * 2 * ORC_PI;
  *g1 = sigma * sqrt(rr) * cos(th);
  *g2 = sigma * sqrt(rr) * sin(th);
}

/* Histogram bin of a tracer density: imax - ilogb(v), clamped to [0, 63]
   (pop_tracer.c:265-268); the subtraction wraps as 32-bit int arithmetic. */
static inline int hist_bin(real v, int imax) {
  int d = (int) ((unsigned) imax - (unsigned) ilogb(v));
  if (d < 0) d = 0;
  else if (d >= ORC_NUM_LOGB_BIN) d = ORC_NUM_LOGB_BIN - 1;
  return d;
}

/* pop_tracer.c:94-133 */
int orc_eval_pdf(size_t nexp, real pbase, size_t Ncells, size_t *nmax,
    size_t *ncell, size_t *ntot, size_t *num, size_t nnum) {
  size_t n;
  double A = 0, sum = 0, power = pbase;
  for (n = 1; n <= nexp; n++) {
    sum += power * n;
    A = nexp / sum;
    power *= pbase;
    if (A * power < 0.5) break;
  }
  if (n > ORC_MAX_NUM_TRACER_PER_CELL) return ORC_ERR_RHO_PBASE;
  size_t nc = 0, nobj = 0;
  for (size_t i = 0; i < n; i++) {
    size_t v = round(A * pow(pbase, n - i));
    if (num && i < nnum) num[i] = v;
    nc += v;
    nobj += v * (n - i);
  }
  if (nc > Ncells || nobj == 0) return ORC_ERR_PAR_PBASE;
  *nmax = n; *ncell = nc; *ntot = nobj;
  return 0;
}

/* pop_tracer.c:149-348 (OpenMP branch semantics). */
static int bias_model(orc_ez *ez, real rho_c, real rho_sat, real rho_exp,
    real lambda, size_t ncell) {
  const size_t num = (size_t) ez->Ng * ez->Ng * ez->Ng;
  const real thres = (rho_c == 0) ? ORC_REAL_MIN : rho_c;
  const double iexp = -1 / rho_exp;       /* evaluated in `real` */
  const real rho_max = 1 + lambda * 5;
  const int imax = ilogb(rho_max);
  ez->hist_imax = imax;
  memset(ez->hist, 0, sizeof(ez->hist));
  const int nt = ez->nthread;
  size_t *cnt = calloc(nt + 1, sizeof(size_t));
  if (!cnt) return ORC_ERR_MEMORY;
  const size_t pnum = num / nt;
  const int rem = num % nt;
  const real *rho = ez->rho;
  real *rhot = ez->rhot;
#pragma omp parallel num_threads(nt)
  {
    const int tid = omp_get_thread_num();
    const size_t pcnt = (tid < rem) ? pnum + 1 : pnum;
    const size_t is = (tid < rem) ? pcnt * tid : pnum * tid + rem;
    size_t c = 0;
    for (size_t i = is; i < is + pcnt; i++) {
      if (rho[i] < thres) rhot[i] = 0;
      else { rhot[i] = (rho[i] > rho_sat) ? rho_sat : rho[i]; c++; }
    }
    cnt[tid + 1] = c;
  }
  for (int t = 0; t < nt; t++) cnt[t + 1] += cnt[t];
  ez->pdf_nvalid = cnt[nt];
  if (cnt[nt] < ncell) { free(cnt); return ORC_ERR_RHO_LESS; }
  size_t *hist = ez->hist;
#pragma omp parallel num_threads(nt)
  {
    const int tid = omp_get_thread_num();
    const size_t pcnt = (tid < rem) ? pnum + 1 : pnum;
    const size_t is = (tid < rem) ? pcnt * tid : pnum * tid + rem;
    size_t rank = cnt[tid];
    size_t lh[ORC_NUM_LOGB_BIN] = {0};
    orc_rng_t rng;
    orc_rng_reset(&rng, ez->rngtype, ez->seed, (num << 1) + ((rank >> 1) << 1));
    double g[2] = {0, 0};
    if (rank & 1) {   /* second member of a pair that started earlier */
      gauss_pair(&rng, lambda, g, g + 1);
      for (int q = 0; q < 2; q++) g[q] = (g[q] >= 0) ? 1 + g[q] : exp(g[q]);
    }
    for (size_t i = is; i < is + pcnt; i++) {
      if (rhot[i] == 0) continue;
      if ((rank & 1) == 0) {
        gauss_pair(&rng, lambda, g, g + 1);
        for (int q = 0; q < 2; q++) g[q] = (g[q] >= 0) ? 1 + g[q] : exp(g[q]);
      }
      double tmp = (1 - exp((double) rhot[i] * iexp)) * g[rank & 1];
      rhot[i] = tmp;
      lh[hist_bin(rhot[i], imax)]++;
      rank++;
    }
    for (int b = 0; b < ORC_NUM_LOGB_BIN; b++) {
#pragma omp atomic
      hist[b] += lh[b];
    }
  }
  free(cnt);
  return 0;
}

static const real *g_sort_key;
static int cmp_desc(const void *a, const void *b) {
  real va = g_sort_key[*(const size_t *) a], vb = g_sort_key[*(const size_t *) b];
  if (va > vb) return -1;
  if (va < vb) return 1;
  size_t ia = *(const size_t *) a, ib = *(const size_t *) b;
  return (ia > ib) - (ia < ib);
}

/* pop_tracer.c:361-514 restated on a given rhot field.  hist/imax are
   recomputed from rhot (bins of pop_tracer.c:265-268 with ilogb_max 5). */
static int pdf_walk_core(const real *rhot, size_t num, const size_t *hist,
    int ilogb_max, size_t ncell, size_t *pdf, int nmax, int16_t *nt,
    size_t *ncell_out, size_t *ntot_out) {
  int imax;
  size_t cnt = 0;
  for (imax = 0; imax < ORC_NUM_LOGB_BIN; imax++) {
    cnt += hist[imax];
    if (cnt >= ncell) break;
  }
  if (imax == ORC_NUM_LOGB_BIN) imax = ORC_NUM_LOGB_BIN - 1;
  const real thres = (imax == ORC_NUM_LOGB_BIN - 1) ? ORC_REAL_MIN :
      (real) pow(2, ilogb_max - imax);
  size_t *cidx = calloc(cnt ? cnt : 1, sizeof(size_t));
  size_t *off = calloc(imax + 1, sizeof(size_t));
  if (!cidx || !off) { free(cidx); free(off); return ORC_ERR_MEMORY; }
  for (int i = 1; i <= imax; i++) off[i] = off[i - 1] + hist[i - 1];
  for (size_t i = 0; i < num; i++) {
    if (rhot[i] < thres) continue;
    int b = (int) ((unsigned) ilogb_max - (unsigned) ilogb(rhot[i]));
    if (b < 0) b = 0;
    else if (b > imax) b = imax;
    if (off[b] < cnt) cidx[off[b]++] = i;
  }
  free(off);
  size_t *cpdf = malloc(sizeof(size_t) * nmax);
  if (!cpdf) { free(cidx); return ORC_ERR_MEMORY; }
  memcpy(cpdf, pdf, sizeof(size_t) * nmax);
  size_t ntot = 0;
  for (int i = 0; i < nmax; i++) ntot += pdf[i] * (size_t) (nmax - i);
  cnt = 0;
  size_t used = 0;
  int ipdf = 0;
  g_sort_key = rhot;
  for (int i = 0; i <= imax; i++) {
    size_t nrest = hist[i] - used;
    if (nrest <= cpdf[ipdf]) {
      cpdf[ipdf] -= nrest;
      if (cpdf[ipdf] == 0) ipdf++;
      cnt += nrest;
      used = 0;
    }
    else {
      /* qselect (qselect.c:336-370) only fixes the selected set; a full
         descending sort of the remainder gives the same set. */
      qsort(cidx + cnt, nrest, sizeof(size_t), cmp_desc);
      size_t ecnt = cnt + cpdf[ipdf];
      if (ecnt > 0) {
        real mn = rhot[cidx[ecnt - 1]];
        for (size_t k = cnt; k + 1 < ecnt; k++)
          if (mn > rhot[cidx[k]]) mn = rhot[cidx[k]];
        for (size_t k = ecnt; k < cnt + nrest; k++) {
          if (rhot[cidx[k]] == mn) {
            size_t b = cidx[ecnt];
            cidx[ecnt++] = cidx[k];
            cidx[k] = b;
            pdf[ipdf]++;
            ncell++;
            ntot += nmax - ipdf;
          }
        }
      }
      used += ecnt - cnt;
      cnt = ecnt;
      ipdf++;
      if (used < hist[i]) i--;
      else used -= hist[i];
    }
    if (ipdf == nmax) break;
  }
  free(cpdf);
  if (ipdf != nmax || cnt != ncell) { free(cidx); return ORC_ERR_UNKNOWN; }
  memset(nt, 0, sizeof(int16_t) * num);
  size_t e = 0;
  for (int i = 0; i < nmax; i++) {
    size_t s = e;
    e += pdf[i];
    for (size_t j = s; j < e; j++) nt[cidx[j]] = (int16_t) (nmax - i);
  }
  free(cidx);
  *ncell_out = ncell;
  *ntot_out = ntot;
  return 0;
}

int orc_pdf_walk(const real *rhot, size_t num, size_t ncell_in,
    const size_t *pdf_in, int nmax, int16_t *nt, size_t *pdf_out,
    size_t *ncell_out, size_t *ntot_out) {
  size_t hist[ORC_NUM_LOGB_BIN] = {0};
  for (size_t i = 0; i < num; i++)
    if (rhot[i] != 0) hist[hist_bin(rhot[i], 5)]++;
  memcpy(pdf_out, pdf_in, sizeof(size_t) * nmax);
  return pdf_walk_core(rhot, num, hist, 5, ncell_in, pdf_out, nmax, nt,
      ncell_out, ntot_out);
}

/* pop_tracer.c:517-582: NGP particle counts in tracer-hosting cells. */
static inline void ngp_cell(const orc_ez *ez, size_t idx, int i, int j, int k,
    real igs, real *p, int *c) {
  const int Ng = ez->Ng;
  p[0] = ez->psi[0][idx] * igs + i;
  p[1] = ez->psi[1][idx] * igs + j;
  p[2] = ez->psi[2][idx] * igs + k;
  for (int d = 0; d < 3; d++) {
    if (p[d] < 0) p[d] += Ng;
    else if (p[d] >= Ng) p[d] -= Ng;
    c[d] = (int) (p[d] + 0.5);
    if (c[d] >= Ng) c[d] = 0;
  }
}

static int count_particles(orc_ez *ez) {
  const int Ng = ez->Ng;
  const real igs = ez->Ng / ez->Lbox;
  for (int i = 0; i < Ng; i++)
    for (int j = 0; j < Ng; j++)
      for (int k = 0; k < Ng; k++) {
        const size_t idx = ORC_IDX(Ng, i, j, k);
        real p[3]; int c[3];
        ngp_cell(ez, idx, i, j, k, igs, p, c);
        const size_t pidx = ORC_IDX(Ng, c[0], c[1], c[2]);
        if (ez->ntbuf[pidx] == 0) continue;
        if (ez->npbuf[pidx] == UINT16_MAX) return ORC_ERR_RHO_NPART;
        ez->npbuf[pidx]++;
      }
  return 0;
}

/* pop_tracer.c:697-753 */
static void assign_vel(const orc_ez *ez, const real *psi, size_t n,
    const real *x, const real *y, const real *z, real *v) {
  const int Ng = ez->Ng;
  const real igs = ez->Ng / ez->Lbox;
  const real *rho = ez->rho;
#pragma omp parallel for num_threads(ez->nthread) schedule(static)
  for (size_t i = 0; i < n; i++) {
    real xx = x[i] * igs, yy = y[i] * igs, zz = z[i] * igs;
    int x0 = (int) xx, y0 = (int) yy, z0 = (int) zz;
    real wx1 = xx - x0, wx0 = 1 - wx1;
    real wy1 = yy - y0, wy0 = 1 - wy1;
    real wz1 = zz - z0, wz0 = 1 - wz1;
    if (x0 == Ng) x0 = 0;
    if (y0 == Ng) y0 = 0;
    if (z0 == Ng) z0 = 0;
    const int x1 = (x0 == Ng - 1) ? 0 : x0 + 1;
    const int y1 = (y0 == Ng - 1) ? 0 : y0 + 1;
    const int z1 = (z0 == Ng - 1) ? 0 : z0 + 1;
    const real r000 = rho[ORC_IDX(Ng, x0, y0, z0)] * wx0 * wy0 * wz0;
    const real r001 = rho[ORC_IDX(Ng, x0, y0, z1)] * wx0 * wy0 * wz1;
    const real r010 = rho[ORC_IDX(Ng, x0, y1, z0)] * wx0 * wy1 * wz0;
    const real r011 = rho[ORC_IDX(Ng, x0, y1, z1)] * wx0 * wy1 * wz1;
    const real r100 = rho[ORC_IDX(Ng, x1, y0, z0)] * wx1 * wy0 * wz0;
    const real r101 = rho[ORC_IDX(Ng, x1, y0, z1)] * wx1 * wy0 * wz1;
    const real r110 = rho[ORC_IDX(Ng, x1, y1, z0)] * wx1 * wy1 * wz0;
    const real r111 = rho[ORC_IDX(Ng, x1, y1, z1)] * wx1 * wy1 * wz1;
    const real rs = r000 + r001 + r010 + r011 + r100 + r101 + r110 + r111;
    v[i] = psi[ORC_IDX(Ng, x0, y0, z0)] * r000
         + psi[ORC_IDX(Ng, x0, y0, z1)] * r001
         + psi[ORC_IDX(Ng, x0, y1, z0)] * r010
         + psi[ORC_IDX(Ng, x0, y1, z1)] * r011
         + psi[ORC_IDX(Ng, x1, y0, z0)] * r100
         + psi[ORC_IDX(Ng, x1, y0, z1)] * r101
         + psi[ORC_IDX(Ng, x1, y1, z0)] * r110
         + psi[ORC_IDX(Ng, x1, y1, z1)] * r111;
    v[i] *= ez->vfac / rs;
  }
}

/* pop_tracer.c:773-1027 */
static int generate_tracers(orc_ez *ez, size_t ntot, bool att_part,
    real *pos[3], real *vel[3]) {
  const int Ng = ez->Ng;
  const size_t num = (size_t) Ng * Ng * Ng;
  const real gs = ez->Lbox / ez->Ng;
  int16_t *nt = ez->ntbuf;
  size_t iv = 0;
  double t0 = omp_get_wtime();
  if (att_part) {
    /* pop_tracer.c:781, 804-879: serial Bentley selection on stream 4N. */
    orc_rng_t rng;
    orc_rng_reset(&rng, ez->rngtype, ez->seed, num << 2);
    const real igs = ez->Ng / ez->Lbox;
    size_t c = 0;
    for (int i = 0; i < Ng; i++)
      for (int j = 0; j < Ng; j++)
        for (int k = 0; k < Ng; k++) {
          const size_t idx = ORC_IDX(Ng, i, j, k);
          real p[3]; int cc[3];
          ngp_cell(ez, idx, i, j, k, igs, p, cc);
          const size_t pidx = ORC_IDX(Ng, cc[0], cc[1], cc[2]);
          if (nt[pidx] == 0) continue;
          if (nt[pidx] < ez->npbuf[pidx]) {
            double ran = orc_rng_double(&rng);
            if (ran * ez->npbuf[pidx] >= nt[pidx]) {
              ez->npbuf[pidx]--;
              continue;
            }
          }
          nt[pidx]--;
          ez->npbuf[pidx]--;
          for (int d = 0; d < 3; d++) {
            pos[d][c] = p[d] * gs;
            vel[d][c] = ez->psi[d][idx] * ez->vfac;
          }
          c++;
        }
    iv = c;
    ez->iv = iv;
    if (iv == ntot) {
      ez->tstage[ORC_ST_POS] += omp_get_wtime() - t0;
      return 0;
    }
  }
  /* pop_tracer.c:881-1007: tracers in cell order; tracer t (not counting
     those attached to particles) uses draws 5N + 3t, 5N + 3t + 1, +2. */
  const size_t rows = (size_t) Ng * Ng;
  const int nth = ez->nthread;
  size_t *rc = calloc(nth + 1, sizeof(size_t));
  if (!rc) return ORC_ERR_MEMORY;
  const size_t pnum = rows / nth;
  const int rem = rows % nth;
  int bad = 0;
#pragma omp parallel num_threads(nth)
  {
    const int tid = omp_get_thread_num();
    const size_t pcnt = (tid < rem) ? pnum + 1 : pnum;
    const size_t is = (tid < rem) ? pcnt * tid : pnum * tid + rem;
    size_t c = 0;
    for (size_t ij = is; ij < is + pcnt; ij++)
      for (int k = 0; k < Ng; k++)
        if (nt[ij * Ng + k] > 0) c += nt[ij * Ng + k];
    rc[tid + 1] = c;
#pragma omp barrier
#pragma omp single
    for (int t = 0; t < nth; t++) rc[t + 1] += rc[t];
    size_t off = rc[tid];
    orc_rng_t rng;
    orc_rng_reset(&rng, ez->rngtype, ez->seed, num * 5 + off * 3);
    off += iv;
    for (size_t ij = is; ij < is + pcnt; ij++) {
      const int i = ij / Ng, j = ij % Ng;
      for (int k = 0; k < Ng; k++) {
        const size_t idx = ij * Ng + k;
        for (int n = 0; n < nt[idx]; n++) {
          double r = orc_rng_double(&rng) - 0.5;
          real xx = (i + r) * gs;
          r = orc_rng_double(&rng) - 0.5;
          real yy = (j + r) * gs;
          r = orc_rng_double(&rng) - 0.5;
          real zz = (k + r) * gs;
          if (xx < 0) xx += ez->Lbox;
          if (yy < 0) yy += ez->Lbox;
          if (zz < 0) zz += ez->Lbox;
          if (off < ntot) {
            pos[0][off] = xx; pos[1][off] = yy; pos[2][off] = zz;
          }
          off++;
        }
      }
    }
    if (off != rc[tid + 1] + iv) {
#pragma omp atomic write
      bad = 1;
    }
  }
  if (rc[nth] + iv != ntot) bad = 1;
  free(rc);
  ez->tstage[ORC_ST_POS] += omp_get_wtime() - t0;
  if (bad) return ORC_ERR_UNKNOWN;
  /* pop_tracer.c:1012-1019 with disp_cic :593-683 */
  t0 = omp_get_wtime();
  /* The reference paints into the rhot buffer; a separate one keeps the
     tracer density available to the tests. */
  real *buf = malloc(num * sizeof(real));
  if (!buf) return ORC_ERR_MEMORY;
  for (int d = 0; d < 3; d++) {
    memset(buf, 0, num * sizeof(real));
    orc_cic_paint(ez, d, buf);
#pragma omp parallel for num_threads(ez->nthread) schedule(static)
    for (size_t i = 0; i < num; i++)
      if (ez->rho[i] > ORC_REAL_TOL) buf[i] /= ez->rho[i];
    assign_vel(ez, buf, ntot - iv, pos[0] + iv, pos[1] + iv, pos[2] + iv,
        vel[d] + iv);
  }
  free(buf);
  ez->tstage[ORC_ST_VEL] += omp_get_wtime() - t0;
  return 0;
}

/* pop_tracer.c:1043-1113: tracer pair (2q, 2q+1) uses three Box-Muller pairs
   from draws 8N + 6q .. 8N + 6q + 5 in the order (vx_a, vy_a), (vz_a, vx_b),
   (vy_b, vz_b). */
static void add_vel_scatter(orc_ez *ez, real *vx, real *vy, real *vz,
    size_t n, real sigma_v) {
  const size_t num = (size_t) ez->Ng * ez->Ng * ez->Ng;
  const size_t npair = (n + 1) / 2;
#pragma omp parallel num_threads(ez->nthread)
  {
    const int nth = omp_get_num_threads(), tid = omp_get_thread_num();
    const size_t q0 = npair * tid / nth, q1 = npair * (tid + 1) / nth;
    orc_rng_t rng;
    orc_rng_reset(&rng, ez->rngtype, ez->seed, (num << 3) + q0 * 6);
    for (size_t q = q0; q < q1; q++) {
      size_t i = 2 * q;
      double g1, g2;
      gauss_pair(&rng, sigma_v, &g1, &g2);
      vx[i] += g1;
      vy[i] += g2;
      gauss_pair(&rng, sigma_v, &g1, &g2);
      vz[i++] += g1;
      if (i == n) break;
      vx[i] += g2;
      gauss_pair(&rng, sigma_v, &g1, &g2);
      vy[i] += g1;
      vz[i] += g2;
    }
  }
}

/* pop_tracer.c:1140-1219 */
int orc_populate_tracer(orc_ez *ez, const real *params, size_t nexp,
    bool att_part, size_t *ntracer, real **x, real **y, real **z,
    real **vx, real **vy, real **vz, int *err) {
  if (!err) return ORC_ERR_ARG_ECODE;
  if (*err != ORC_SUCCESS) return *err;
  if (!ez) return (*err = ORC_ERR_ARG_EZ);
  if (!params) return (*err = ORC_ERR_ARG_PARAMS);
  if (nexp < ORC_MIN_NUM_TRACER) return (*err = ORC_ERR_ARG_LOWEXP);
  if (!ntracer) return (*err = ORC_ERR_ARG_NDATA);
  const size_t num = (size_t) ez->Ng * ez->Ng * ez->Ng;
  if (nexp >= num) return (*err = ORC_ERR_ARG_HIEXP);
  if (!ez->psi[0] || !ez->psi[1] || !ez->psi[2] || !ez->rho)
    return (*err = ORC_ERR_ARG_MESH);
  if (ez->vfac == HUGE_VAL) return (*err = ORC_ERR_ARG_COSMO);
  const real rho_c = params[0], rho_exp = params[1], pbase = params[2],
      sigma_v = params[3];
  if (!isfinite(rho_c) || rho_c < 0) return (*err = ORC_ERR_PAR_RCUT);
  if (!isfinite(rho_exp) || rho_exp <= 0) return (*err = ORC_ERR_PAR_REXP);
  if (!isfinite(pbase) || pbase <= 0 || pbase >= 1)
    return (*err = ORC_ERR_PAR_PBASE);
  if (!isfinite(sigma_v) || sigma_v < 0) return (*err = ORC_ERR_PAR_SIGV);

  size_t nmax, ncell, ntot;
  size_t pdf[ORC_MAX_NUM_TRACER_PER_CELL + 1];
  int e = orc_eval_pdf(nexp, pbase, num, &nmax, &ncell, &ntot, pdf,
      ORC_MAX_NUM_TRACER_PER_CELL + 1);
  if (e) return (*err = e);
  ez->pdf_ncell0 = ncell;
  ez->pdf_ntot0 = ntot;

  if (!ez->rhot && !(ez->rhot = malloc(num * sizeof(real))))
    return (*err = ORC_ERR_MEMORY);
  double t0 = omp_get_wtime();
  e = bias_model(ez, rho_c, ORC_RHO_SAT, rho_exp, ORC_GAUSS_WIDTH, ncell);
  ez->tstage[ORC_ST_BIAS] += omp_get_wtime() - t0;
  if (e) return (*err = e);

  t0 = omp_get_wtime();
  free(ez->ntbuf);
  if (!(ez->ntbuf = malloc(num * sizeof(int16_t)))) return (*err = ORC_ERR_MEMORY);
  e = pdf_walk_core(ez->rhot, num, ez->hist, ez->hist_imax, ncell, pdf,
      (int) nmax, ez->ntbuf, &ncell, &ntot);
  ez->tstage[ORC_ST_WALK] += omp_get_wtime() - t0;
  if (e) return (*err = e);
  free(ez->pdf_bins);
  ez->pdf_bins = malloc(sizeof(size_t) * nmax);
  if (ez->pdf_bins) memcpy(ez->pdf_bins, pdf, sizeof(size_t) * nmax);
  ez->pdf_nmax = nmax; ez->pdf_ncell = ncell; ez->pdf_ntot = ntot;
  ez->iv = 0;
  if (att_part) {
    free(ez->npbuf);
    if (!(ez->npbuf = calloc(num, sizeof(uint16_t)))) return (*err = ORC_ERR_MEMORY);
    if ((e = count_particles(ez))) return (*err = e);
  }
  int16_t *ntcopy = malloc(num * sizeof(int16_t));
  if (!ntcopy) return (*err = ORC_ERR_MEMORY);
  memcpy(ntcopy, ez->ntbuf, num * sizeof(int16_t));

  real *cat[6];
  for (int i = 0; i < 6; i++) {
    if (!(cat[i] = malloc((ntot ? ntot : 1) * sizeof(real)))) {
      for (int j = 0; j < i; j++) free(cat[j]);
      free(ntcopy);
      return (*err = ORC_ERR_MEMORY);
    }
  }
  e = generate_tracers(ez, ntot, att_part, cat, cat + 3);
  /* keep the per-cell counts (generate_tracers consumes them for att_part) */
  memcpy(ez->ntbuf, ntcopy, num * sizeof(int16_t));
  free(ntcopy);
  if (e) {
    for (int i = 0; i < 6; i++) free(cat[i]);
    return (*err = e);
  }
  t0 = omp_get_wtime();
  add_vel_scatter(ez, cat[3], cat[4], cat[5], ntot, sigma_v);
  ez->tstage[ORC_ST_SIGV] += omp_get_wtime() - t0;
  *ntracer = ntot;
  *x = cat[0]; *y = cat[1]; *z = cat[2];
  *vx = cat[3]; *vy = cat[4]; *vz = cat[5];
  ez->rho_c = rho_c; ez->rho_exp = rho_exp;
  ez->pdf_base = pbase; ez->sigma_v = sigma_v;
  return ORC_SUCCESS;
}

/* save_res.c:147-158 */
double orc_rsd_z(const orc_ez *ez, real z, real vz, double rsd_fac) {
  double zr = z + rsd_fac * vz;
  if (zr >= ez->Lbox) zr -= ez->Lbox;
  if (zr < 0) zr += ez->Lbox;
  return zr;
}

/* ------------------------------------------------------- test access --- */
int orc_get_info(const orc_ez *ez, double *o) {
  o[0] = ez->Ng; o[1] = ez->Lbox; o[2] = ez->growth2; o[3] = ez->vfac;
  o[4] = ez->nthread; o[5] = (double) ez->pdf_nvalid; o[6] = ez->hist_imax;
  o[7] = (double) ez->iv;
  return 0;
}

/* which: 0-2 psi, 3 rho, 4 rhot (tracer density after the bias model),
   5 nt (int16 per cell), 6 spectrum rhok as generated (interleaved complex),
   7 histogram (64 size_t) */
int orc_copy_field(const orc_ez *ez, int which, void *dst) {
  const size_t num = (size_t) ez->Ng * ez->Ng * ez->Ng;
  const size_t nh = (size_t) ez->Ng * ez->Ng * (ez->Ng / 2 + 1);
  switch (which) {
    case 0: case 1: case 2:
      if (!ez->psi[which]) return 1;
      memcpy(dst, ez->psi[which], num * sizeof(real)); return 0;
    case 3: if (!ez->rho) return 1; memcpy(dst, ez->rho, num * sizeof(real)); return 0;
    case 4: if (!ez->rhot) return 1; memcpy(dst, ez->rhot, num * sizeof(real)); return 0;
    case 5: if (!ez->ntbuf) return 1; memcpy(dst, ez->ntbuf, num * sizeof(int16_t)); return 0;
    case 6: if (!ez->spec) return 1; memcpy(dst, ez->spec, nh * 2 * sizeof(real)); return 0;
    case 7: memcpy(dst, ez->hist, sizeof(ez->hist)); return 0;
    default: return 1;
  }
}

int orc_pdf_info(const orc_ez *ez, size_t *out, size_t n) {
  if (n < 5) return 1;
  out[0] = ez->pdf_nmax; out[1] = ez->pdf_ncell; out[2] = ez->pdf_ntot;
  out[3] = ez->pdf_ncell0; out[4] = ez->pdf_ntot0;
  for (size_t i = 0; i < ez->pdf_nmax && 5 + i < n; i++) out[5 + i] = ez->pdf_bins[i];
  return 0;
}

double orc_stage_time(const orc_ez *ez, int stage) {
  if (stage < 0 || stage >= ORC_ST_N) return -1;
  return ez->tstage[stage];
}

void orc_set_keep_spec(orc_ez *ez, int keep) { ez->keep_spec = keep; }
