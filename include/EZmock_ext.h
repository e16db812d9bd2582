/*
 * EZmock_ext.h -- extensions of the MI355X EZmock engine that have no
 * counterpart in the reference API (kept out of EZmock.h so the drop-in ABI
 * stays exactly the reference's).  Used by the benchmark, the parity tests and
 * callers that keep the catalogue on the GPU.
 */
#ifndef EZMOCK_AMD_EZMOCK_EXT_H
#define EZMOCK_AMD_EZMOCK_EXT_H

#include "EZmock.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Build flavour: sizeof(real) in bytes (4 or 8). */
int EZmock_ext_real_size(void);

/* Number of visible HIP devices (0 without a GPU); never faults. */
int EZmock_ext_device_count(void);

/* Same as EZmock_populate_tracer but the catalogue stays in device memory
   (read it with EZmock_ext_copy_tracers or EZmock_ext_device_tracers). */
int EZmock_ext_populate_device(EZMOCK *ez, const real *params,
    const size_t nexp, const bool att_part, size_t *ntracer, int *err);

/* One mock: EZmock_create_dens_field from the initial conditions
   (fixamp, iphase) followed by EZmock_ext_populate_device without attached
   particles, with the same results.  The first call runs those two calls
   and captures their launches as two HIP graphs (field with its CIC; the
   populate with the selection on the device), which later calls replay
   for any seed (EZmock_ext_set_seed) and bias parameters, with one host
   synchronisation per mock and no allocation.  The graphs are recaptured
   after anything they depend on changes (P(k) table, cosmology, options,
   a buffer that had to grow) and are not used with MT19937 streams, the
   per-kernel timers, CIC by tile scan or exact ranges, or option 7 or 8
   off; the plain calls then run instead.  A CIC list range or a capacity
   that overflows during a replay makes the call redo that stage by the
   plain calls.  The reference's counterpart is the pair of calls
   (dens_field.c:337-432, pop_tracer.c:1140-1219), as the Python loop of a
   calibration sweep makes them. */
int EZmock_ext_mock(EZMOCK *ez, const real *params, const size_t nexp, const bool fixamp,
    const bool iphase, size_t *ntracer, int *err);

/* Copy the device-resident catalogue of the last populate call into six
   caller-owned host arrays of at least `n` elements (n <= ntracer). */
int EZmock_ext_copy_tracers(EZMOCK *ez, real *x, real *y, real *z, real *vx,
    real *vy, real *vz, const size_t n, int *err);

/* Device pointers of the catalogue (valid until the next populate/destroy). */
int EZmock_ext_device_tracers(EZMOCK *ez, void *ptrs[6], size_t *ntracer,
    int *err);

/* Redshift-space z of the device-resident catalogue, computed on the GPU in
   double as save_res.c:147-158 does: z + rsd_fac*vz wrapped once into
   [0, Lbox).  Written to a caller-owned host array of `ntracer` doubles. */
int EZmock_ext_rsd_z(EZMOCK *ez, const double rsd_fac, double *z_rsd,
    int *err);

/* Copy an internal field to host memory:
     0,1,2  psi_x, psi_y, psi_z          Ng^3 real
     3      rho (CIC density)            Ng^3 real
     4      rhot (tracer density)        Ng^3 real
     5      tracers per cell             Ng^3 int16
     6      generated spectrum rhok      Ng*Ng*(Ng/2+1) complex (2 real each),
            reference layout (kept only if option 1 is set)
     7      log2 histogram               64 uint64                        */
int EZmock_ext_copy_field(EZMOCK *ez, const int which, void *dst, int *err);

/* How the instance's calls ran, counted since EZmock_init: out[0] =
   populates with the selection on the device, [1] = host-driven populates,
   [2] = device-selection populates redone host-driven (capacity), [3] =
   mocks replayed from captured graphs, [4] = graph captures, [5] =
   candidate cells (threshold bin and above) of the last device selection. */
int EZmock_ext_path_counts(EZMOCK *ez, uint64_t *out6);

/* PDF of the last populate call: out[0]=nmax, [1]=ncell, [2]=ntot,
   [3]=ncell before ties, [4]=ntot before ties, [5..5+nmax) cells per bin. */
int EZmock_ext_pdf_info(EZMOCK *ez, size_t *out, const size_t n);

/* Options: 1 = keep a copy of the generated spectrum (tests);
            2 = record per-kernel device times with HIP events;
            3 = CIC method: 0 = particle binning into fixed list ranges
                per tile (default: 4 entries per cell; after an overflow
                the field is repainted with exact ranges and the ranges of
                later fields become 1.25 x the fullest tile, if that is at
                most 6 per cell, or at most 16 per cell where a quarter of
                the device memory stays free, else exact ranges), 1 =
                particle binning
                with exact ranges (count pass), 2 = tile scan of the
                displaced lattice wherever the displacement range admits it
                (no particle lists), else binning.  All give the same
                fields;
            4 = run the last two FFT passes of the three displacement
                components on three streams (1, default) or on one (0);
                one stream while option 2 is on;
            5 = bytes of device memory the fixed list ranges may take
                (0 = no limit, the default); above it, or when their
                allocation fails, exact ranges are used;
            6 = FFT autotuning (1, default): the first field of a grid size
                >= 256 in the process times the schedule candidates
                (EZmock_ext_fft_schedule) and keeps the fastest; 0 = use the
                grid's current schedule;
            7 = selection on the device (1, default; 2 = with capacities of
                a fraction of the counts, which exercises the redo): with
                MRG32k3a and
                tracers at uniform positions, the valid-cell count, the
                threshold bin, the candidate count and the PDF walk's
                results stay on the device (no host round trip before the
                end of populate, or before the velocities with host
                arrays); candidate keys and catalogue buffers are kept while
                they hold 4 ntot + 64 Ki keys and ntot + ntot/32 + 1024
                tracers (ntot: eval_pdf's count before ties), else grown to
                6 ntot + 64 Ki and ntot + ntot/8 + 4096, and a populate that
                outgrows them is redone by the host-driven path (0), which
                sizes them from the counts.  Both give the same catalogue;
            8 = HIP-graph replay in EZmock_ext_mock (1, default; 0 drops
                the instance's graphs);
            9 = spectrum generated inside the first FFT pass (1, default,
                with MRG32k3a initial conditions once the grid's FFT
                schedule is tuned and its first pass is fused; 0 = the
                spectrum written by its own kernel and read back).  Both
                give the same fields; the spectrum copy of option 1 needs
                the separate kernel. */
int EZmock_ext_set_option(EZMOCK *ez, const int option, const long value);

/* Per-kernel device time totals (ms) and launch counts since the last reset
   (option 2).  Returns the number of kernel classes; names are written as a
   '\n'-separated list into `names` (capacity `cap`). */
int EZmock_ext_kernel_times(EZMOCK *ez, double *ms, long *count, const int n,
    char *names, const size_t cap);
void EZmock_ext_reset_kernel_times(EZMOCK *ez);

/* Host-side state: out[0] = Ng, [1] = Lbox, [2] = growth2, [3] = vfac,
   [4] = number of P(k) table bins (0 if unset), [5] = kmin, [6] = kmax,
   [7] = 1 if the spline is in log space. */
int EZmock_ext_info(EZMOCK *ez, double *out8);

/* The trimmed P(k) table and its spline second derivatives (n entries). */
int EZmock_ext_pk_table(EZMOCK *ez, double *k, double *P, double *P2,
    const int n);

/* Power-spectrum multipoles of the device-resident catalogue of the last
   populate call, for calibration loops (the reference fits its mocks with an
   external interlaced-CIC estimator, example/fit/conf/powspec.conf).  Tracers
   are CIC-assigned to the mock's Ng^3 grid (a second grid shifted by half a
   cell when `interlace`); z is shifted to redshift space by rsd_fac*vz and
   wrapped once into [0, Lbox) (rsd_fac = 0: real space, line of sight z).
   The window of the assignment is divided out.  Bin b holds the modes with
   b*dk <= |k| < (b+1)*dk (nbin <= 1024): mean k, P0 (shot noise V/ntracer
   subtracted), P2 and the number of modes (the +k/-k pair counts twice). */
int EZmock_ext_powspec(EZMOCK *ez, const double rsd_fac, const bool interlace,
    const double dk, const int nbin, double *k, double *P0, double *P2,
    double *nmodes, int *err);

/* Re-seed the instance's random streams (the next create_dens_field and
   populate use `seed`; 0 is rejected as by EZmock_init). */
int EZmock_ext_set_seed(EZMOCK *ez, const uint64_t seed, int *err);

/* ---- request pool: continuous batching of mocks on one GPU ----
   nslot worker threads run queued requests concurrently, each on an engine
   instance with its own HIP stream.  Instances stay resident, keyed by grid
   size and box, and an idle one serves the next request of its shape; a
   request without one starts when its device footprint (the estimate of
   EZmock_ext_footprint) fits mem_budget next to every resident instance,
   after evicting least recently used idle instances of other shapes if that
   makes room.  Every request runs the same calls as a stand-alone mock with
   the pool's cosmology (set_cosmology with eval = true) and linear P(k) (no
   BAO modification, linear spline). */
typedef struct EZMOCK_POOL EZMOCK_POOL;

typedef struct {
  int Ngrid;                    /* grid size                               */
  double Lbox;                  /* box size                                */
  uint64_t seed;                /* MRG32k3a seed (> 0)                     */
  real params[4];               /* rho_c, rho_exp, pdf_base, sigma_v       */
  size_t nexp;                  /* expected number of tracers              */
  double rsd_fac;               /* P(k): redshift-space shift (0 = real)   */
  double pk_dk;                 /* P(k): bin width                         */
  int pk_nbin;                  /* P(k): number of bins (0 = no P(k))      */
  int want_catalog;             /* copy the catalogue to the result        */
} EZMOCK_REQUEST;

typedef struct {
  long id;
  int err;                      /* 0 or the engine's error code            */
  size_t ntracer;
  double t_submit, t_start, t_end;   /* steady-clock seconds               */
  double *k, *P0, *P2, *nmodes; /* pk_nbin each (malloc'd) or NULL        */
  real *cat[6];                 /* x, y, z, vx, vy, vz (malloc'd) or NULL */
} EZMOCK_RESULT;

/* Device bytes accounted to one request of grid Ngrid. */
size_t EZmock_ext_footprint(const int Ngrid, const int with_powspec);

EZMOCK_POOL *EZmock_ext_pool_init(const int nslot, const size_t mem_budget,
    const double *k, const int nk, const double *Pk, const double z,
    const double z_pk, const double Omega_m, const double Omega_nu,
    const double w, int *err);

/* An EZmock_ext_set_option applied to every instance of the pool (before
   the first submit; the pool's own settings come first: option 5, the list
   limit of its memory accounting, and option 8 = 0, plain launches, which
   ran a sweep of 64 mocks at 256^3 at 863 mocks/s against 708 with four
   instances replaying HIP graphs at once). */
int EZmock_ext_pool_option(EZMOCK_POOL *p, const int option, const long value);

/* Queue a request; returns its id (> 0), or -1 with *err set. */
long EZmock_ext_pool_submit(EZMOCK_POOL *pool, const EZMOCK_REQUEST *req,
    int *err);

/* Block until request `id` is done; fills *res (free its arrays with
   EZmock_ext_result_free) and returns res->err. */
int EZmock_ext_pool_wait(EZMOCK_POOL *pool, const long id, EZMOCK_RESULT *res);

void EZmock_ext_result_free(EZMOCK_RESULT *res);

/* Finishes the running requests, drops the queued ones, frees everything. */
void EZmock_ext_pool_destroy(EZMOCK_POOL *pool);

/* Wait for all device work of the instance. */
int EZmock_ext_sync(EZMOCK *ez, int *err);

/* ---- stand-alone device routines (parity tests) ---- */

/* n uniform draws of stream `randgen` seeded with `seed`, starting at draw
   index `start` (reset(seed, start) then get_double n times; `pos` selects
   get_double_pos).  Computed on the GPU with per-thread jump-ahead. */
int EZmock_ext_uniform(const int randgen, const uint64_t seed,
    const uint64_t start, const size_t n, const bool pos, double *out,
    int *err);

/* 3-D complex-to-real FFT (unnormalised, e^{+ikx}) on the GPU.
   in: Ng*Ng*(Ng/2+1) complex (interleaved), out: Ng^3 real. */
int EZmock_ext_fft_c2r(const int Ng, const real *in, real *out, int *err);

/* 3-D real-to-complex FFT (unnormalised, e^{-ikx}) on the GPU. */
int EZmock_ext_fft_r2c(const int Ng, const real *in, real *out, int *err);

/* PDF walk of pop_tracer.c:361-514 on the GPU for an arbitrary tracer-density
   array: nt (int16, num) and the final pdf (nmax entries), ncell, ntot. */
int EZmock_ext_pdf_walk(const real *rhot, const size_t num,
    const size_t ncell, const size_t *pdf, const int nmax, int16_t *nt,
    size_t *pdf_out, size_t *ncell_out, size_t *ntot_out, int *err);

/* FFT schedule of grid size Ng in this process: out[0] = column tile width
   of the strided (j) passes, out[1] = tile width of the first (i) pass
   (0 = the built-in default, else 4, 8 or 16), out[2] = 1 if the first pass
   runs as three single-component passes instead of the fused
   three-component pass, out[3] = 1 if chosen by the autotuner (or set),
   out[4] = tile width of the first pass that generates its own spectrum
   (option 9; 0 = that of out[1]).  All schedules give bit-identical
   fields. */
int EZmock_ext_fft_schedule(const int Ngrid, int *out5);
/* Sets the schedule of grid size Ng (tile widths 0, 4, 8, 16; split 0/1;
   the generating first pass takes tk_i). */
int EZmock_ext_fft_set_schedule(const int Ngrid, const int tk_j, const int tk_i,
    const int split);

/* Plane-wave check of the c2r transforms at any supported grid size, on
   the device: a spectrum of nmodes (<= 4) modes (ijk[3m..3m+2] = i, j, kk
   with 0 < kk < Ng/2, amplitude amp[2m] + i amp[2m+1]) goes through the
   generic pass sequence (path 0: one output, the modes as given) or the
   production first pass of the three displacement components (path 1:
   three outputs, the modes times n_x, n_y, n_z), and each output is
   compared with sum_m 2 w_m Re(A_m exp(2 pi i (i x + j y + kk z) / Ng)).
   maxerr[c] / maxref[c]: largest |output - closed form| and |closed form|
   of output c. */
int EZmock_ext_fft_check(const int Ng, const int path, const int nmodes, const int *ijk,
    const double *amp, double *maxerr, double *maxref, int *err);

/* ---- no-allocation entry point (no counterpart in the reference) ----
   For callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops): the PDF walk of EZmock_ext_pdf_walk (the
   selection of pop_tracer.c:359-514) on device-resident arrays, calling no
   hipMalloc/hipFree.  rhot_dev: `num` tracer densities, nt_dev: `num` int16
   tracers per cell (output), both device memory of the current device;
   pdf[nmax], pdf_out[nmax], *ncell_out and *ntot_out are host memory.
   stream: a hipStream_t (NULL: the null stream); the call synchronises it.
   Every temporary lives in one static device buffer per GPU
   (EZmock_ext_noalloc_scratch_bytes(), held until the process ends), so at
   most EZMOCK_NOALLOC_MAX_CELLS cells (else EZMOCK_ERR_ARG_NGRID) and
   EZMOCK_NOALLOC_MAX_CAND candidate cells, those in the log2 density bins
   down to the threshold bin (else EZMOCK_ERR_ARG_HIEXP).  The buffer keeps
   the last call's contents; calls that overlap on one GPU race on it. */
#define EZMOCK_NOALLOC_MAX_CELLS ((size_t) 1 << 31)
#define EZMOCK_NOALLOC_MAX_CAND ((size_t) 1 << 22)
size_t EZmock_ext_noalloc_scratch_bytes(void);
int EZmock_ext_pdf_walk_noalloc(const real *rhot_dev, const size_t num,
    const size_t ncell, const size_t *pdf, const int nmax, int16_t *nt_dev,
    size_t *pdf_out, size_t *ncell_out, size_t *ntot_out, void *stream, int *err);

#ifdef __cplusplus
}
#endif

#endif
