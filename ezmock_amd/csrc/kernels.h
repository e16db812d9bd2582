// kernels.h -- launch wrappers of the gfx950 kernels (field.hip, tracer.hip).
#pragma once
#include "common.h"
#include "rng.h"

struct GenParams {
  int Ng;
  size_t Nh;                    // Ng*Ng*(Ng/2+1) half-complex cells
  double kfac, logkfac, Pfac;
  int logpk, fixamp, iphase;
  MrgState seed_state;
};

// field.hip
hipError_t launch_pk_table(double *tab, int nsq, const double *kt,
    const double *Pt, const double *P2t, int n, int logpk, double kfac,
    double logkfac, hipStream_t st);
// mtw: MT19937 stream words of draws 0 .. (mult (Nh - 1)) (nullptr: MRG32k3a)
// bases: gen_base_count(gp.Nh) states of scratch (MRG32k3a; unused with
// MT19937 words).
size_t gen_base_count(size_t Nh);
// seedp (optional): the MRG32k3a seed state in device memory, in place of
// gp.seed_state.
hipError_t launch_gen_rhok(cplx *rhok, cplx *rhok_ny, real *ref_copy,
    const GenParams &gp, const double *ptab, const uint32_t *jtab,
    const uint32_t *mtw, MrgState *bases, hipStream_t st, const MrgState *seedp = nullptr);
// The kz = 0 and Nyquist planes only (for fft_first_pass3_gen), MRG32k3a.
hipError_t launch_gen_planes(cplx *rhok, cplx *rhok_ny, const GenParams &gp,
    const double *ptab, const uint32_t *jtab, hipStream_t st, const MrgState *seedp = nullptr);
hipError_t launch_gen_rhok_wn(cplx *rhok, cplx *rhok_ny, const cplx *dk,
    const cplx *dk_ny, const GenParams &gp, const double *ptab,
    real *ref_copy, hipStream_t st);
hipError_t launch_unpack_r2c(cplx *buf, cplx *ny2, int Ng, hipStream_t st);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t st);
// Plane-wave check of a c2r output (see k_plane_wave_check).
#define EZ_FFT_MAX_MODES 4
struct FftModes {
  int n;
  int ijk[EZ_FFT_MAX_MODES][3];
  double re[EZ_FFT_MAX_MODES], im[EZ_FFT_MAX_MODES], w[EZ_FFT_MAX_MODES];
};
hipError_t launch_plane_wave_check(const real *out, int Ng, const FftModes &md,
    unsigned long long *res, hipStream_t st);
hipError_t launch_scale(real *p, real s, size_t n, hipStream_t st);

// cic.hip
struct CicGeom {
  int Ng;
  real igs;
  int TX, TY, TZ;               // tile shape in cells (powers of two)
  int lx, ly, lz;               // their base-2 logarithms
  int ntx, nty, ntz;            // tiles per dimension
  size_t N;                     // particles (< 2^32)
  FastDiv fNg, fntz;            // division by Ng and by ntz
  // bounds blocks of the Lagrangian lattice (tile scan): BX x BY x BZ
  // lattice points, powers of two
  int BX, BY, BZ, lbx, lby, lbz;
  int nbx, nby, nbz;
};
CicGeom cic_geometry(int Ng, real igs);
// Slot of cell (x, y, z) in the velocity gather's field {rho_c, u0, u1, u2}:
// 2 x 2 x 2 bricks of consecutive slots (one 128-byte line in FP32), bricks
// in cell-major order of (x/2, y/2, z/2).  A tracer's eight interpolation
// corners span 1.5^3 = 3.4 lines on average instead of the 4.5 lines of
// the cell-major layout (two z-neighbours per line pair, 1/8 of them split).
// Every supported grid size is even.
__host__ __device__ __forceinline__ size_t vel_brick_index(int x, int y, int z, int Ng) {
  const size_t h = (size_t) (Ng >> 1);
  return ((((size_t) (x >> 1) * h + (size_t) (y >> 1)) * h + (size_t) (z >> 1)) << 3)
         | (size_t) (((x & 1) << 2) | ((y & 1) << 1) | (z & 1));
}
size_t cic_nkeys(const CicGeom &g);
size_t cic_nblocks(const CicGeom &g);
// Words of the scan key array: displacement bounds (3 minima, 3 maxima as
// order-preserving keys), a non-finite flag and the scan status.
#define CIC_GKEY_WORDS 8
#define CIC_GKEY_STATUS 7
// Per bounds block, the minimum and maximum of psi_a * igs over its lattice
// points (bnd: 2 cic_nblocks(g) real4), and their extrema over the grid.
hipError_t launch_cic_bounds(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, real4 *bnd, rkey *gkey, hipStream_t st);
// Tile scan: the density (and optional momentum) tiles from the lattice
// points whose bounds admit them.  When the displacement range makes the
// scan inapplicable, nothing is written and gkey[CIC_GKEY_STATUS] is set.
hipError_t launch_cic_scan(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const real4 *bnd, rkey *gkey, real *rho,
    const real *rho_norm, real4 *vel, hipStream_t st);
hipError_t launch_cic_count(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, uint32_t *counts, uint64_t *off, uint64_t *cursor,
    uint64_t *scan_tmp, hipStream_t st);
// cap = 0: exact list ranges from the count pass (cursor = off).  cap > 0:
// tile t owns list[t cap, (t + 1) cap) after launch_cic_fixed_init; entries
// past a tile's range are dropped and *ovf set (the painting must then be
// redone with exact ranges).
hipError_t launch_cic_scatter(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, uint64_t *cursor, uint32_t *list, uint32_t cap, uint32_t *ovf,
    hipStream_t st);
hipError_t launch_cic_fixed_init(uint64_t *cursor, const CicGeom &g, uint32_t cap,
    uint32_t *ovf, hipStream_t st);
// vel (optional): per cell {rho_c, u0, u1, u2}, u_a = m_a / rho_c where
// rho_c > REAL_TOL and m_a otherwise (m_a the painted momentum); rho_c is
// the density painted in the same pass, or rho_norm when rho is null.
// vel is stored in 2 x 2 x 2 bricks (vel_brick_index).
// off: exact ranges [off[t], off[t + 1]) (cap = 0), or the cursor array of
// fixed ranges [t cap, min(cursor[t], (t + 1) cap)) (cap > 0).
hipError_t launch_cic_tiles(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const uint64_t *off, const uint32_t *list, real *rho,
    const real *rho_norm, real4 *vel, uint32_t cap, hipStream_t st);

// tracer.hip
struct BiasParams {
  size_t N;
  real thres, rho_sat;
  double iexp;
  double lambda;
  int imax;
  MrgState seed_state;
};

// Cells per block of the per-cell passes (256 threads x 16 cells).
#define EZ_CELLS_PER_BLOCK 4096

// Control block of a populate with the selection on the device (no host
// round trip between the bias model and the catalogue), in device memory.
// The host writes the parameters before the launch, the kernels the
// results; kernels given a null control block use their own arguments.
struct PopCtl {
  BiasParams bp;                // seed_state advanced to draw 2N
  MrgState pos_seed, vel_seed;  // advanced to draws 5N and 8N
  double sigma_v;
  uint64_t ncell;               // cells eval_pdf asks for
  uint64_t kcap, ccap;          // capacities: candidate keys, catalogue
  uint64_t kfcap;               // refined candidates the sort's grid covers
  int nmax;
  // written by the kernels
  int imax;                     // bin where the cumulative histogram reaches ncell
  real rho_thres;
  uint64_t nvalid, cnt, nsort, ncell_out, ntot;
  uint32_t flags;               // POP_* bits
  // candidates kept for the sort: those at or above a threshold refined
  // inside the threshold's log2 bin (key <= kfine), nfine of them
  rkey kfine;
  uint64_t nfine;
};
#define POP_RHO_LESS 1u         // fewer valid cells than ncell
#define POP_NO_BIN 2u           // the histogram never reaches ncell
#define POP_KEYS_FULL 4u        // more candidates than kcap
#define POP_CAT_FULL 8u         // more tracers than ccap
#define POP_WALK_ERR 16u        // the PDF walk ran past the candidates
#define POP_CAND_ERR 32u        // candidate counts disagree

hipError_t launch_count_valid(const real *rho, real thres, size_t N,
    uint32_t *blockcnt, hipStream_t st, const PopCtl *ctl = nullptr);
hipError_t launch_scan_u32(const uint32_t *in, uint64_t *out, size_t n,
    uint64_t *tmp, hipStream_t st);
size_t scan_tmp_words(size_t n);
// blockhist (optional): per-block 64-bin histograms (nblk x 64 u32).
// mtw (MT19937, else nullptr): stream words of draws 2N + 2 p, pair p.
// ctl: bp (its seed already at draw 2N) from the control block.
hipError_t launch_bias(const real *rho, real *rhot, const uint64_t *blockoff,
    unsigned long long *hist, uint32_t *blockhist, const BiasParams &bp,
    const uint32_t *jtab, const uint32_t *mtw, hipStream_t st, const PopCtl *ctl = nullptr);
// The threshold bin, candidate count and error flags of the control block
// from the bias histogram and the valid-cell count *nvalid.
hipError_t launch_select(const unsigned long long *hist, const uint64_t *nvalid, PopCtl *ctl,
    hipStream_t st);
// Refinement of the candidates (keys[0, ctl->nsort)) inside the threshold
// bin by the top four mantissa bits: the refined threshold and count
// (ctl->kfine, ctl->nfine), and the kept keys into out in key-block order;
// kcap: the key capacity, which sizes the grids; bc (17 u32 per block of
// CPB keys) and boff (one u64 per block) are scratch, written before read
// (the sort's digit counts and offsets serve).
hipError_t launch_refine(const rkey *keys, size_t kcap, PopCtl *ctl, uint32_t *bc,
    uint64_t *boff, rkey *out, hipStream_t st);
hipError_t launch_block_cand(const uint32_t *blockhist, size_t nblk, int imax,
    uint32_t *blockcnt, hipStream_t st, const PopCtl *ctl = nullptr);
hipError_t launch_compact_at(const real *rhot, real thres, size_t N,
    const uint64_t *blockoff, const uint32_t *blockcnt, rkey *keys, size_t cap,
    unsigned long long *flag, hipStream_t st, const PopCtl *ctl = nullptr);
hipError_t launch_compact(const real *rhot, real thres, size_t N, rkey *keys,
    size_t cap, unsigned long long *counter, hipStream_t st);
// 64-bin log2 histogram of nonzero tracer densities, added to hist.
hipError_t launch_hist_rhot(const real *rhot, size_t N, unsigned long long *hist,
    hipStream_t st);
// np (optional): the number of keys in device memory, n then being the
// capacity (<= n keys sorted, grids sized by n).
hipError_t launch_sort(rkey *keys, rkey *tmp, size_t n, uint32_t *ghist,
    uint64_t *gofs, uint64_t *scan_tmp, hipStream_t st, const uint64_t *np = nullptr);
size_t sort_hist_words(size_t n);
// ctl: M and nmax from the control block, results and flags into it
// (ncand: the candidate count of the compaction's scan).
hipError_t launch_walk(const rkey *sorted, size_t M, const uint64_t *num,
    int nmax, rkey *vb, uint64_t *cnt, uint64_t *res, hipStream_t st,
    PopCtl *ctl = nullptr, const uint64_t *ncand = nullptr);
hipError_t launch_assign_nt(const real *rhot, size_t N, const rkey *vb,
    int nmax, int16_t *nt, uint32_t *blocktr, hipStream_t st, const PopCtl *ctl = nullptr);
// mtw (MT19937, else nullptr): stream words of draws 5N + 3 t, tracer t.
// ctl: the seed from the control block (ntot: the catalogue capacity).
hipError_t launch_positions(const int16_t *nt, const uint64_t *blockofs,
    int Ng, double Lbox, real gs, MrgState seed, const uint32_t *jtab,
    const uint32_t *mtw, real *x, real *y, real *z, size_t ntot, hipStream_t st,
    const PopCtl *ctl = nullptr);
// mtw (MT19937, else nullptr): stream words of draws 8N + 6 q, pair q.
// Tracers t < natt (attached to particles) keep the velocity already in
// vx/vy/vz; the others take the CIC gather; all get the sigma_v scatter.
// ctl: seed, sigma_v and the tracer count from the control block (ntot:
// the catalogue capacity, which sizes the grid).
hipError_t launch_velocities(const real *x, const real *y, const real *z,
    const real4 *vel, size_t ntot, size_t natt, int Ng, real igs, double vfac, double sigma_v,
    MrgState seed, const uint32_t *jtab, const uint32_t *mtw, real *vx, real *vy,
    real *vz, hipStream_t st, const PopCtl *ctl = nullptr);
// tracers attached to particles (pop_tracer.c:517-582, 804-879)
struct AttGeom {
  int Ng;
  real igs, gs;
  size_t N;
  FastDiv fNg;
};
hipError_t launch_att_count(const real *p0, const real *p1, const real *p2,
    const AttGeom &g, const int16_t *nt, uint32_t *np, uint32_t *ovf, hipStream_t st);
// blockoff null: per-block contested-cell counts; else write cid / cinfo.
hipError_t launch_att_cells(const int16_t *nt, uint32_t *np, size_t N, uint32_t *blockcnt,
    const uint64_t *blockoff, uint32_t *cinfo, hipStream_t st);
hipError_t launch_att_classify(const real *p0, const real *p1, const real *p2,
    const AttGeom &g, const int16_t *nt, const uint32_t *np, uint8_t *cls,
    uint32_t *blockcnt, hipStream_t st);
hipError_t launch_att_list(const real *p0, const real *p1, const real *p2,
    const AttGeom &g, const uint32_t *np, const uint8_t *cls, const uint64_t *blockoff,
    uint32_t *plist, hipStream_t st);
// blockoff null: per-block selected counts; else write the attached tracers.
hipError_t launch_att_select(const real *p0, const real *p1, const real *p2,
    const AttGeom &g, double vfac, const uint8_t *cls, const uint64_t *contoff,
    const uint8_t *sel, uint32_t *blockcnt, const uint64_t *blockoff, real *const cat[6],
    hipStream_t st);
hipError_t launch_att_rest(const int16_t *nt, const uint32_t *np, size_t N, int16_t *rest,
    uint32_t *blockcnt, hipStream_t st);
hipError_t launch_uniform(MrgState seed, uint64_t start, size_t n,
    const uint32_t *jtab, double *out, hipStream_t st);
hipError_t launch_uniform_mt(const uint32_t *w, size_t n, int pos, double *out,
    hipStream_t st);
hipError_t launch_rsd(const real *z, const real *vz, size_t n, double fac,
    double Lbox, double *out, hipStream_t st);

// powspec.hip
hipError_t launch_pk_paint(const real *x, const real *y, const real *z,
    const real *vz, size_t n, int Ng, double Lbox, double rsd_fac,
    double shift, real *grid, hipStream_t st);
hipError_t launch_pk_bin(const cplx *X1, const cplx *N1, const cplx *X2,
    const cplx *N2, int Ng, double kf, double dk, int nbin, double pnorm,
    double *acc, hipStream_t st);
