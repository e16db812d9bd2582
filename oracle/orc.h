/*
 * orc.h -- CPU restatement ("oracle") of the EZmock hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker (or the timed CPU baseline), never as the
 * thing that produces results for a user.
 *
 * The functions restate the behaviour of the reference C library
 * (the .c files of libEZmock/EZmock in cheng-zhao/EZmock) from its source, file:line
 * cited at each function.  Compiled twice: with -DSINGLE_PREC (real = float,
 * the reference default, options.mk:4) and without (real = double).
 *
 * Parity status of each part (see the test that pins it):
 *   rng (MRG32k3a core, jumps) ... pinned: RngStreams KAT + A^(2^127) matrix
 *                                   constants + brute-force stepping
 *   rng (MT19937 core) ........... pinned: C++ std::mt19937 KAT
 *   rng (seed map, int->double) .. parity unpinned (prand submodule absent)
 *   hyp2f1 / cspline ............. pinned against the reference's own
 *                                   math/hypergeom.c and math/cspline.c
 *                                   compiled into oracle/_ref, and scipy
 *   cosmology / linear P(k) ...... pinned: closed forms (EdS), scipy spline
 *   FFT (c2r/r2c) ................ pinned: numpy irfftn/rfftn fixtures
 *   noise / spectra .............. pinned: closed form per cell, Hermitian
 *                                   invariants
 *   CIC .......................... pinned: sum = N, brute force
 *   eval_pdf ..................... pinned: closed form
 *   bias / PDF walk .............. pinned: brute-force sorted formulation,
 *                                   planted ties
 *   positions / velocities ....... pinned: invariants, brute force
 */
#ifndef ORC_H
#define ORC_H

#include <stdint.h>
#include <stdbool.h>
#include <stddef.h>

#ifdef SINGLE_PREC
typedef float real;
#define ORC_REAL_MIN 1.17549435082228750797e-38F
#define ORC_REAL_TOL 1e-5
#else
typedef double real;
#define ORC_REAL_MIN 2.2250738585072013830902e-308
#define ORC_REAL_TOL 1e-10
#endif

#define ORC_PI 3.14159265358979323846
#define ORC_DOUBLE_TOL 1e-8
#define ORC_DOUBLE_EPSILON 1e-16
#define ORC_MAX_GRID 65536
#define ORC_MAX_PK_NBIN 1073741824
#define ORC_MIN_NUM_TRACER 128
#define ORC_MAX_NUM_TRACER_PER_CELL 1024
#define ORC_MAX_K_BAO 0.5
#define ORC_NUM_LOGB_BIN 64
#define ORC_RHO_SAT 100
#define ORC_GAUSS_WIDTH 10

/* Error codes: same numbering as errmsg.h:33-88 of the reference, OpenMP
   build (EZMOCK_ERR_FFT_INIT present). */
enum {
  ORC_SUCCESS = 0, ORC_ERR_MEMORY, ORC_ERR_ARG_ECODE, ORC_ERR_ARG_EZ,
  ORC_ERR_ARG_COSMO, ORC_ERR_ARG_PK, ORC_ERR_ARG_MESH, ORC_ERR_ARG_LBOX,
  ORC_ERR_ARG_NGRID, ORC_ERR_ARG_RNG, ORC_ERR_ARG_SEED, ORC_ERR_ARG_THREAD,
  ORC_ERR_ARG_NULL, ORC_ERR_ARG_NBIN, ORC_ERR_ARG_PARAMS, ORC_ERR_ARG_LOWEXP,
  ORC_ERR_ARG_HIEXP, ORC_ERR_ARG_NDATA, ORC_ERR_ARG_FNAME, ORC_ERR_ARG_RSD,
  ORC_ERR_COS_GROWTH, ORC_ERR_PK_FINITE, ORC_ERR_PK_NONPOS, ORC_ERR_PK_NONASC,
  ORC_ERR_PK_RANGE, ORC_ERR_PK_NBIN, ORC_ERR_PK_MODBAO, ORC_ERR_RNG_SET,
  ORC_ERR_RNG_RESET, ORC_ERR_RNG_JUMP, ORC_ERR_PAR_COSMO, ORC_ERR_PAR_RCUT,
  ORC_ERR_PAR_REXP, ORC_ERR_PAR_PBASE, ORC_ERR_PAR_SIGV, ORC_ERR_RHO_USED,
  ORC_ERR_RHO_LESS, ORC_ERR_RHO_PBASE, ORC_ERR_RHO_NPART, ORC_ERR_FFT_INIT,
  ORC_ERR_FILE, ORC_ERR_UNKNOWN
};

/* ---------------------------------------------------------------- RNG --- */
#define ORC_RNG_MRG32K3A 0
#define ORC_RNG_MT19937 1

typedef struct {
  int type;
  uint64_t s[6];                /* MRG32k3a: x1[n-3..n-1], x2[n-3..n-1] */
  uint32_t mt[624];             /* MT19937 */
  int mti;
} orc_rng_t;

/* Seed the generator and skip `step` draws (one draw = one get_double). */
int orc_rng_reset(orc_rng_t *r, int type, uint64_t seed, uint64_t step);
double orc_rng_double(orc_rng_t *r);       /* in [0,1) */
double orc_rng_double_pos(orc_rng_t *r);   /* in (0,1] */
/* Raw MRG32k3a helpers exported for the known-answer tests. */
void orc_mrg_set_state(orc_rng_t *r, const uint64_t s[6]);
void orc_mrg_get_state(const orc_rng_t *r, uint64_t s[6]);
void orc_mrg_jump(orc_rng_t *r, uint64_t step);
void orc_mrg_matpow(int comp, int log2n, uint64_t out[9]);
uint32_t orc_mt_next(orc_rng_t *r);

/* ---------------------------------------------------------------- math -- */
int orc_hyp2f1(double a, double b, double c, double z, double tol,
    int maxiter, double *res);
void orc_cspline_ypp(const double *x, const double *y, size_t n, double *ypp);
double orc_cspline_eval(const double *x, const double *y, const double *ypp,
    double xv, size_t i);

/* ---------------------------------------------------------------- FFT --- */
/* 3-D complex-to-real, unnormalised, e^{+ikx} (FFTW backward semantics):
   in: Ng*Ng*(Ng/2+1) complex (interleaved real), out: Ng^3 real. */
int orc_fft_c2r_3d(int Ng, const real *in, real *out, int nthread);
/* 3-D real-to-complex, unnormalised, e^{-ikx} (FFTW forward semantics). */
int orc_fft_r2c_3d(int Ng, const real *in, real *out, int nthread);
/* 1-D complex FFT, in place, sign = +1 (backward) or -1 (forward). */
int orc_fft_1d(int n, real *data, int sign);

/* ---------------------------------------------------------------- API --- */
typedef struct orc_ez orc_ez;

orc_ez *orc_init(double Lbox, int Ngrid, int randgen, uint64_t seed,
    int nthread, int *err);
int orc_set_cosmology(orc_ez *ez, double pk_norm, double fHa, bool eval,
    double z, double z_pk, double Omega_m, double Omega_nu, double w,
    int *err);
int orc_setup_linear_pk(orc_ez *ez, const double *k, int n, const double *Pk,
    const double *Pnw, double mBAO, bool logint, int *err);
int orc_create_dens_field(orc_ez *ez, real *psi[3], bool deepcopy,
    real *delta, bool fixamp, bool iphase, int *err);
int orc_populate_tracer(orc_ez *ez, const real *params, size_t nexp,
    bool att_part, size_t *ntracer, real **x, real **y, real **z,
    real **vx, real **vy, real **vz, int *err);
void orc_destroy(orc_ez *ez);
const char *orc_errmsg(int err);
void orc_free(void *p);
double orc_rsd_z(const orc_ez *ez, real z, real vz, double rsd_fac);

/* Accessors for tests (copy internal state out). */
int orc_get_info(const orc_ez *ez, double *out8);
int orc_copy_field(const orc_ez *ez, int which, void *dst);
size_t orc_pk_nbin(const orc_ez *ez);
int orc_pk_table(const orc_ez *ez, double *k, double *P, double *P2);
double orc_pk_interp(const orc_ez *ez, double k);
int orc_pdf_info(const orc_ez *ez, size_t *out, size_t n);
int orc_eval_pdf(size_t nexp, real pbase, size_t Ncells, size_t *nmax,
    size_t *ncell, size_t *ntot, size_t *num, size_t nnum);
double orc_stage_time(const orc_ez *ez, int stage);
void orc_set_keep_spec(orc_ez *ez, int keep);

/* Literal restatement of the reference PDF walk (pop_tracer.c:413-455) on a
   given rhot array: fills nt (int16) and returns the final ncell/ntot. */
int orc_pdf_walk(const real *rhot, size_t num, size_t ncell_in,
    const size_t *pdf_in, int nmax, int16_t *nt, size_t *pdf_out,
    size_t *ncell_out, size_t *ntot_out);

#endif
