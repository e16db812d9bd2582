// fft.h -- gfx950 3-D real FFT passes (see fft.hip for the data layout).
#pragma once
#include "common.h"

// Supported grid sizes: 2^n and 3*2^n from 8 to 2048.
bool fft_supported(int Ng);

// One strided c2c pass along axis 0 (i) or 1 (j) of the packed half-complex
// field buf[Ng][Ng][Ng/2].  sign +1: e^{+ikx}, -1: e^{-ikx}.  mode 1 reads
// the generated spectrum (rhok, rhok_ny) instead of buf and applies the
// component multiply `comp` (0 x, 1 y, 2 z) and kz = 0 / Nyquist packing.
// tk: column tile width (4, 8, 16; 0 = built-in default; -1 = the grid's
// schedule, see fft_schedule).
hipError_t fft_strided_pass(cplx *buf, const cplx *rhok, const cplx *rhok_ny,
    const cplx *tw, int Ng, int axis, int sign, int mode, int comp,
    hipStream_t st, int tk = 0);

// Row pass along kk: sign +1 = c2r (packed half spectrum -> Ng reals per
// row), -1 = r2c.  twN: exp(+2 pi i m / Ng), twh: exp(+2 pi i m / (Ng/2)).
hipError_t fft_row_pass(cplx *buf, const cplx *twN, const cplx *twh, int Ng,
    int sign, hipStream_t st);

// Packed kz = 0 / Nyquist slots (kk = 0) of the three displacement
// components, slot0[comp][j][i] (3 Ng^2 complex), for fft_first_pass3.
hipError_t fft_slot0_pack(const cplx *rhok, const cplx *rhok_ny, cplx *slot0,
    int Ng, hipStream_t st);

// Schedule of a grid size: tile widths of the strided j pass (tk_j) and of
// the first pass (tk_i; 0 = defaults), and the first pass fused (split = 0:
// the spectrum read once, three transforms per workgroup) or split into
// three single-component strided passes (split = 1).
// tk_g: tile width of the first pass that generates its spectrum
// (fft_first_pass3_gen; 0 = that of the first pass).
struct FftSchedule { int tk_j, tk_i, split, tk_g; };
struct GenParams;
FftSchedule fft_schedule(int Ng, bool *tuned);
void fft_set_schedule(int Ng, const FftSchedule &s);

// First c2r pass (axis i) of the three displacement components rhok * n_a,
// a = x, y, z, into out0..out2 (kk = 0 from slot0, or from rhok / rhok_ny
// when split), by `sched` or the grid's schedule (null).
hipError_t fft_first_pass3(cplx *out0, cplx *out1, cplx *out2, const cplx *rhok,
    const cplx *rhok_ny, const cplx *slot0, const cplx *tw, int Ng, hipStream_t st,
    const FftSchedule *sched = nullptr);

// The same first pass generating the spectrum (MRG32k3a initial
// conditions, gp) tile by tile instead of reading rhok; the kz = 0 and
// Nyquist planes come through slot0 (from launch_gen_planes and
// fft_slot0_pack).  hipErrorNotSupported when the grid's schedule splits
// the first pass.  seedp (optional): the seed in device memory.
hipError_t fft_first_pass3_gen(cplx *out0, cplx *out1, cplx *out2, const cplx *slot0,
    const cplx *tw, int Ng, const GenParams &gp, const double *ptab, const uint32_t *jtab,
    const struct MrgState *seedp, hipStream_t st, int tk = -1);

// Times every candidate schedule of grid Ng on the given buffers (out0..2
// overwritten, the spectrum read) and records the fastest for the grid;
// ms_best[0], [1]: best times of the j pass and of the first pass.  With gp
// (MRG32k3a initial conditions), also the tile widths of the first pass that
// generates its spectrum (tk_g).
hipError_t fft_autotune(int Ng, cplx *out0, cplx *out1, cplx *out2, const cplx *rhok,
    const cplx *rhok_ny, const cplx *slot0, const cplx *tw, hipStream_t st,
    FftSchedule *best, double *ms_best,
    const GenParams *gp = nullptr, const double *ptab = nullptr, const uint32_t *jtab = nullptr);
