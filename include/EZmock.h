/*
 * EZmock.h -- public C ABI of the MI355X EZmock engine (libEZmock_f32.so,
 * libEZmock_f64.so).
 *
 * Drop-in replacement for the reference's libEZmock/EZmock/EZmock.h:38-216:
 * the same type `EZMOCK`, the same eight functions with the same argument
 * meaning, ownership and error convention (a sticky `int *err`; every call
 * returns immediately with *err if it is non-zero on entry).  The hot path
 * (EZmock_create_dens_field, EZmock_populate_tracer) runs on one AMD GPU
 * (gfx950) through hand-written HIP kernels; the set-up calls are host code.
 *
 * `real` is float when SINGLE_PREC is defined (libEZmock_f32.so, the
 * reference default, options.mk:4) and double otherwise (libEZmock_f64.so).
 */
#ifndef EZMOCK_AMD_EZMOCK_H
#define EZMOCK_AMD_EZMOCK_H

#include <stdint.h>
#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef SINGLE_PREC
typedef float real;
#else
typedef double real;
#endif

/* Opaque instance: six pointers, 48 bytes (reference EZmock.h:49-56). */
typedef struct {
  void *conf;           /* box and grid                                  */
  void *rng;            /* random stream selection                       */
  void *cosmo;          /* growth parameters                             */
  void *pk;             /* linear power spectrum table                   */
  void *par;            /* last effective-bias parameters                */
  void *mesh;           /* device-resident fields                        */
} EZMOCK;

/* Reference EZmock.h:76-77.  randgen: 0 = MRG32k3a, 1 = MT19937.
   nthread >= 0 (0 = all cores) drives the host-side stages only. */
EZMOCK *EZmock_init(const double Lbox, const int Ngrid, const int randgen,
    const uint64_t seed, const int nthread, int *err);

/* Reference EZmock.h:98-100.  eval = false: set growth2 = pk_norm and
   vfac = fHa directly; eval = true: evaluate them in flat wCDM. */
int EZmock_set_cosmology(EZMOCK *ez, const double pk_norm, const double fHa,
    const bool eval, const double z, const double z_pk, const double Omega_m,
    const double Omega_nu, const double w, int *err);

/* Reference EZmock.h:119-121. */
int EZmock_setup_linear_pk(EZMOCK *ez, const double *k, const int n,
    const double *Pk, const double *Pnw, const double mBAO, const bool logint,
    int *err);

/* Reference EZmock.h:144-145.  psi != NULL: displacements given (copied to
   the GPU; with deepcopy = false they are re-read from `psi` at each
   populate call, as the reference keeps only the pointers); delta != NULL:
   white-noise field given; otherwise a random initial field from the seed. */
int EZmock_create_dens_field(EZMOCK *ez, real *psi[3], bool deepcopy,
    real *delta, const bool fixamp, const bool iphase, int *err);

/* Reference EZmock.h:170-172.  params = {rho_c, rho_exp, pdf_base, sigma_v}.
   The six output arrays are malloc'd host memory owned by the caller
   (release with free()). */
int EZmock_populate_tracer(EZMOCK *ez, const real *params, const size_t nexp,
    const bool att_part, size_t *ntracer, real **x, real **y, real **z,
    real **vx, real **vy, real **vz, int *err);

/* Reference EZmock.h:196-198.  rsd_fac > 0 writes (x, y, z + rsd_fac*vz)
   wrapped into the box; otherwise (x, y, z, vx, vy, vz). */
int EZmock_write_ascii(EZMOCK *ez, const real *x, const real *y, const real *z,
    const real *vx, const real *vy, const real *vz, const size_t ntracer,
    const double rsd_fac, const bool header, const char *fname, int *err);

/* Reference EZmock.h:206. */
void EZmock_destroy(EZMOCK *ez);

/* Reference EZmock.h:216; codes and texts of errmsg.h:33-88 / errmsg.c:25-70
   (OpenMP numbering: FFT_INIT = 39, FILE = 40, UNKNOWN = 41). */
const char *EZmock_errmsg(const int err);

#ifdef __cplusplus
}
#endif

#endif
