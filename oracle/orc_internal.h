/*
 * orc_internal.h -- state of one oracle instance (test infrastructure only).
 * Mirrors the reference's EZMOCK_CONF / RNG / COSMO / PK / PAR / MESH / PDF
 * structures (structs.h:43-105) in one struct.
 */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include "orc.h"

#define ORC_IDX(Ng, i, j, k) ((((size_t) (i) * (Ng)) + (j)) * (Ng) + (k))

enum { ORC_ST_NOISE = 0, ORC_ST_FFT, ORC_ST_CIC, ORC_ST_BIAS, ORC_ST_WALK,
       ORC_ST_POS, ORC_ST_VEL, ORC_ST_SIGV, ORC_ST_N };

struct orc_ez {
  /* EZMOCK_CONF, EZMOCK_RNG */
  int nthread, Ng;
  double Lbox;
  int rngtype;
  uint64_t seed;
  /* EZMOCK_COSMO */
  double growth2, vfac;
  /* EZMOCK_PK */
  int pk_set, pk_n;
  bool pk_log;
  double pk_kmin, pk_kmax;
  double *pk_k, *pk_P, *pk_P2;
  /* EZMOCK_PAR */
  double bao, rho_c, rho_exp, pdf_base, sigma_v;
  /* EZMOCK_MESH */
  bool psi_ref, fixamp, iphase;
  int have_rho;
  real *psi[3];
  real *rhok, *rhok2;           /* interleaved complex, Ng*Ng*(Ng/2+1) */
  real *rho, *rhot;
  int16_t *ntbuf;               /* tracers per cell (copy kept for tests) */
  uint16_t *npbuf;              /* particles per cell (att_part) */
  real *spec;                   /* rhok right after generation (tests) */
  int keep_spec;
  /* last EZMOCK_PDF */
  size_t pdf_nmax, pdf_ncell, pdf_ntot, pdf_ncell0, pdf_ntot0, pdf_nvalid;
  size_t *pdf_bins;
  size_t hist[ORC_NUM_LOGB_BIN];
  int hist_imax;
  size_t iv;                    /* tracers attached to particles */
  double tstage[ORC_ST_N];
};

double orc_pk_eval(const orc_ez *ez, double k);
void orc_cic_paint(orc_ez *ez, int dim, real *out);

#endif
