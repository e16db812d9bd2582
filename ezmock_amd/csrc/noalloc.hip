// noalloc.hip -- the PDF walk on device-resident arrays without any device
// allocation (EZmock_ext_pdf_walk_noalloc).
//
// The selection of tracer cells (pop_tracer.c:359-514: log2 histogram,
// threshold bin, candidate compaction, the walk with its tie rule, tracers
// per cell) as EZmock_ext_pdf_walk runs it, but on the caller's device
// arrays and with every temporary in one static device buffer per GPU,
// `scratch`.  It is for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops).
//
// `scratch` holds NA_SCRATCH_BYTES on every GPU the library runs on, for
// the life of the process, and keeps the last call's contents: every call
// clears the counters it accumulates into, and calls that overlap on one GPU
// race on it.  In host code the array's name is a host-side stand-in, not a
// device address: its address on the current device comes from
// hipGetSymbolAddress, once per device, and only that address reaches the
// kernels.
#include <mutex>

#include "common.h"
#include "kernels.h"
#include "EZmock_ext.h"

namespace {

constexpr size_t NA_MAX_CAND = EZMOCK_NOALLOC_MAX_CAND;
constexpr size_t NA_MAX_CELLS = EZMOCK_NOALLOC_MAX_CELLS;
constexpr size_t NA_NMAX = EZ_MAX_NUM_TRACER_PER_CELL + 1;
constexpr size_t NA_SORT_HW = 256 * ((NA_MAX_CAND + EZ_CELLS_PER_BLOCK - 1) / EZ_CELLS_PER_BLOCK);
constexpr size_t NA_SORT_SCAN = (NA_SORT_HW + EZ_CELLS_PER_BLOCK - 1) / EZ_CELLS_PER_BLOCK + 1 + 8;

constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }

// byte offsets of the parts of the buffer
constexpr size_t OFF_HIST = 0;                                           // 64 u64
constexpr size_t OFF_CNTR = OFF_HIST + al256(EZ_NUM_LOGB_BIN * 8);       // 2 u64
constexpr size_t OFF_NUM = OFF_CNTR + al256(2 * 8);                      // nmax u64
constexpr size_t OFF_PCNT = OFF_NUM + al256(NA_NMAX * 8);                // nmax u64
constexpr size_t OFF_VB = OFF_PCNT + al256(NA_NMAX * 8);                 // nmax rkey
constexpr size_t OFF_RES = OFF_VB + al256(NA_NMAX * sizeof(rkey));       // 4 u64
constexpr size_t OFF_BLK = OFF_RES + al256(4 * 8);                       // u32 per block
constexpr size_t OFF_KEYS = OFF_BLK + al256(NA_MAX_CELLS / EZ_CELLS_PER_BLOCK * 4);
constexpr size_t OFF_TMP = OFF_KEYS + al256(NA_MAX_CAND * sizeof(rkey));
constexpr size_t OFF_GH = OFF_TMP + al256(NA_MAX_CAND * sizeof(rkey));  // u32
constexpr size_t OFF_GO = OFF_GH + al256(NA_SORT_HW * 4);                // u64
constexpr size_t OFF_GS = OFF_GO + al256(NA_SORT_HW * 8);                // u64
constexpr size_t NA_SCRATCH_BYTES = OFF_GS + al256(NA_SORT_SCAN * 8);

}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

namespace {

// device address of `scratch` on each device, found on the first call there
constexpr int NA_MAX_DEV = 64;
std::mutex g_na_mu;
unsigned char *g_na_ptr[NA_MAX_DEV] = {nullptr};

int scratch_ptr(unsigned char **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEV) {
    (void) hipGetLastError();
    fprintf(stderr, "EZmock: no usable HIP device for the no-allocation walk\n");
    return EZMOCK_ERR_UNKNOWN;
  }
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void) hipGetLastError();
      fprintf(stderr, "EZmock: hipGetSymbolAddress of the walk scratch failed (%s)\n",
          hipGetErrorString(e));
      return EZMOCK_ERR_UNKNOWN;
    }
    g_na_ptr[dev] = (unsigned char *) p;
  }
  *out = g_na_ptr[dev];
  return 0;
}

}  // namespace

extern "C" size_t EZmock_ext_noalloc_scratch_bytes(void) { return NA_SCRATCH_BYTES; }

extern "C" int EZmock_ext_pdf_walk_noalloc(const real *rhot_dev, const size_t num,
    const size_t ncell, const size_t *pdf, const int nmax, int16_t *nt_dev,
    size_t *pdf_out, size_t *ncell_out, size_t *ntot_out, void *stream, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err) return *err;
  if (!rhot_dev || !nt_dev || !pdf || !pdf_out || !ncell_out || !ntot_out || num == 0 ||
      nmax < 1 || nmax > EZ_MAX_NUM_TRACER_PER_CELL)
    return (*err = EZMOCK_ERR_ARG_NULL);
  if (num > NA_MAX_CELLS) {
    fprintf(stderr, "EZmock: %zu cells exceed the no-allocation walk's maximum of %zu\n",
        num, NA_MAX_CELLS);
    return (*err = EZMOCK_ERR_ARG_NGRID);
  }
  unsigned char *sp = nullptr;
  int e = scratch_ptr(&sp);
  if (e) return (*err = e);
  hipStream_t st = (hipStream_t) stream;
  unsigned long long *hist = (unsigned long long *) (sp + OFF_HIST);
  unsigned long long *counter = (unsigned long long *) (sp + OFF_CNTR);
  uint64_t *d_num = (uint64_t *) (sp + OFF_NUM);
  uint64_t *d_cnt = (uint64_t *) (sp + OFF_PCNT);
  rkey *d_vb = (rkey *) (sp + OFF_VB);
  uint64_t *d_res = (uint64_t *) (sp + OFF_RES);
  uint32_t *blk = (uint32_t *) (sp + OFF_BLK);
  rkey *keys = (rkey *) (sp + OFF_KEYS);
  rkey *tmp = (rkey *) (sp + OFF_TMP);
  uint32_t *gh = (uint32_t *) (sp + OFF_GH);
  uint64_t *go = (uint64_t *) (sp + OFF_GO);
  uint64_t *gs = (uint64_t *) (sp + OFF_GS);

  // histogram on the device (counters cleared by a fill kernel)
  unsigned long long hist_h[EZ_NUM_LOGB_BIN];
  if (launch_zero(sp, OFF_NUM, st) != hipSuccess ||
      launch_hist_rhot(rhot_dev, num, hist, st) != hipSuccess ||
      hipMemcpyAsync(hist_h, hist, sizeof(hist_h), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return (*err = EZMOCK_ERR_UNKNOWN);
  // threshold bin, pop_tracer.c:361-369
  int imax;
  size_t cnt = 0;
  for (imax = 0; imax < EZ_NUM_LOGB_BIN; imax++) {
    cnt += hist_h[imax];
    if (cnt >= ncell) break;
  }
  if (imax == EZ_NUM_LOGB_BIN) return (*err = EZMOCK_ERR_UNKNOWN);
  if (cnt > NA_MAX_CAND) {
    fprintf(stderr, "EZmock: %zu candidate cells exceed the no-allocation walk's maximum "
        "of %zu\n", cnt, NA_MAX_CAND);
    return (*err = EZMOCK_ERR_ARG_HIEXP);
  }
  const real thres = (imax == EZ_NUM_LOGB_BIN - 1) ? (real) EZ_REAL_MIN : (real) pow(2, 5 - imax);
  uint64_t numu[NA_NMAX];
  for (int i = 0; i < nmax; i++) numu[i] = pdf[i];
  uint64_t res[3] = {0, 0, 0}, ncand = 0;
  uint64_t pcnt[NA_NMAX];
  if (hipMemcpyAsync(d_num, numu, nmax * sizeof(uint64_t), hipMemcpyHostToDevice, st) != hipSuccess ||
      launch_compact(rhot_dev, thres, num, keys, NA_MAX_CAND, counter, st) != hipSuccess ||
      launch_sort(keys, tmp, cnt, gh, go, gs, st) != hipSuccess ||
      launch_walk(keys, cnt, d_num, nmax, d_vb, d_cnt, d_res, st) != hipSuccess ||
      launch_assign_nt(rhot_dev, num, d_vb, nmax, nt_dev, blk, st) != hipSuccess ||
      hipMemcpyAsync(res, d_res, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(&ncand, counter, sizeof(uint64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(pcnt, d_cnt, nmax * sizeof(uint64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return (*err = EZMOCK_ERR_UNKNOWN);
  if (ncand != cnt || res[2]) return (*err = EZMOCK_ERR_UNKNOWN);
  for (int i = 0; i < nmax; i++) pdf_out[i] = pcnt[i];
  *ncell_out = res[0];
  *ntot_out = res[1];
  return (*err = EZMOCK_SUCCESS);
}
