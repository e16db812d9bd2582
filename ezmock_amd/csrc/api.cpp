// api.cpp -- the libEZmock C ABI on the GPU engine.
//
// Host orchestration of the eight reference entry points
// (libEZmock/EZmock/EZmock.h:76-216) and of the extensions in EZmock_ext.h.
// Argument validation, error codes and their order follow the reference
// functions cited at each entry point; the hot path is launched on one HIP
// stream per instance and its fields stay resident in device memory between
// calls (the calibration loop re-populates one field many times).
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>
#include <unistd.h>
#include <atomic>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.h"
#include "mt.h"
#include "rng.h"
#include "fft.h"
#include "kernels.h"
#include "host_math.h"
#include "EZmock_ext.h"

// ------------------------------------------------------------ instance
namespace {

struct Conf { int nthread, Ng; double Lbox; };
struct RngConf { int type; uint64_t seed; };
struct Cosmo { double growth2, vfac; };
struct PkTab {
  int n; bool log; double kmin, kmax;
  std::vector<double> k, P, P2;
  long version;
};
struct Par { double bao, rho_c, rho_exp, pdf_base, sigma_v; };

struct Timer { std::string name; hipEvent_t a, b; };

struct Mesh {
  bool have_field = false, psi_ref = false, fixamp = false, iphase = false;
  real *psi_host[3] = {nullptr, nullptr, nullptr};
  bool dev_ready = false;
  hipStream_t st = nullptr;
  // two more streams (default priority) for the per-component FFT passes,
  // forked from and joined into st with events
  hipStream_t sx[2] = {nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  bool fft_streams = true;
  bool fft_tune = true;         // time the FFT schedule candidates on first use
  const uint32_t *jtab = nullptr;
  MrgState *gen_base = nullptr;  // k_gen_rhok workgroup states (MRG32k3a)
  // fields (device)
  real *psi[3] = {nullptr, nullptr, nullptr};
  real *rho = nullptr;          // also the spectrum buffer before the CIC
  cplx *ny = nullptr;           // kz = Nyquist plane (+ r2c scratch), 2 Ng^2
  cplx *slot0 = nullptr;        // packed kz = 0 / Nyquist slots, 3 Ng^2
  real *rhot = nullptr;
  int16_t *nt = nullptr;
  real4 *vel4 = nullptr;        // {rho, u0, u1, u2} per cell, for the velocities
  bool mom_valid = false;
  // CIC binning scratch
  uint32_t *cic_counts = nullptr, *cic_list = nullptr;
  uint64_t *cic_off = nullptr, *cic_cur = nullptr, *cic_scan = nullptr;
  size_t cic_list_cap = 0;
  // CIC tile scan: bounds blocks, extrema keys, the scan status (pinned)
  real4 *cic_bnd = nullptr;
  rkey *cic_gkey = nullptr, *cic_status_h = nullptr;
  // binning with fixed list ranges per tile: overflow flag (device, pinned
  // host copy); cleared for the instance after the first overflow
  uint32_t *cic_ovf = nullptr, *cic_ovf_h = nullptr;
  bool cic_fixed_ok = true;
  uint32_t cic_fcap = 0;        // entries per tile range (0: the default)
  size_t cic_list_limit = 0;    // bytes allowed for fixed ranges (0: no limit)
  // 0: binning, fixed ranges (exact after an overflow); 1: binning, exact
  // ranges; 2: tile scan where valid
  int cic_path = 0;
  // painting completed only after the next synchronisation: kind 1 = tile
  // scan (redo by binning if declined), 2 = fixed ranges (redo with exact
  // ranges on overflow)
  struct { int kind; int Ng; real igs; real *rho; bool mom; } cic_pend = {0, 0, 0, nullptr, false};
  double *ptab = nullptr;
  long ptab_version = -1;
  cplx *twN = nullptr, *twh = nullptr;
  real *spec = nullptr;
  // tracer scratch
  uint32_t *blockcnt = nullptr, *blockhist = nullptr;
  uint64_t *blockoff = nullptr, *scan_tmp = nullptr;
  unsigned long long *hist = nullptr, *counter = nullptr;
  rkey *keys = nullptr, *keys_tmp = nullptr;
  size_t keys_cap = 0;
  uint32_t *sort_gh = nullptr;
  uint64_t *sort_go = nullptr, *sort_scan = nullptr;
  uint64_t *d_num = nullptr, *d_cnt = nullptr, *d_res = nullptr;
  rkey *d_vb = nullptr;
  double *d_rsd = nullptr;
  size_t rsd_cap = 0;
  // power-spectrum estimator: two assignment grids, a Nyquist plane, sums
  real *pk_grid[2] = {nullptr, nullptr};
  // MT19937 stream words of the current stage (two per draw)
  uint32_t *mtw = nullptr, *mtw2 = nullptr;   // MT19937 words (second: velocity draws)
  size_t mtw_cap = 0, mtw2_cap = 0;
  // MT19937 segment windows per stream use (0 spectrum, 1 bias, 2
  // positions, 3 velocities, 4 particle selection), keyed by seed and first draw: the jump trees
  // depend on nothing else, so they are kept across fields and built ahead
  // of use on smt where the draw counts are bounded in advance
  struct MtWin {
    uint32_t *win = nullptr, *scr = nullptr;     // windows; the tree's scratch
    size_t cap = 0, nseg = 0;
    uint64_t seed = 0, first = ~0ull;
    hipEvent_t ready = nullptr, used = nullptr;   // built; last read
  } mtwin[5];
  // (created at the first MT19937 window build: HIP spreads streams over a
  // few hardware queues in the order they are created, and one more stream
  // per instance would put every instance's main stream on one queue)
  hipStream_t smt = nullptr;
  // tracers attached to particles: particles per cell, particle classes,
  // tracers left per cell, block counts / offsets, contested cells and
  // particles, their selection and draws
  uint32_t *att_np = nullptr, *att_bcnt = nullptr, *att_cinfo = nullptr,
      *att_plist = nullptr, *att_ovf = nullptr;
  uint8_t *att_cls = nullptr, *att_sel = nullptr;
  int16_t *att_rest = nullptr;
  uint64_t *att_boff[3] = {nullptr, nullptr, nullptr}, *att_stmp = nullptr;
  double *att_draws = nullptr;
  size_t att_ccap = 0, att_pcap = 0;
  cplx *pk_ny = nullptr;
  double *pk_acc = nullptr;
  real *cat[6] = {nullptr};
  size_t cat_cap = 0, ntracer = 0, natt = 0;
  // pinned host staging of the catalogue for EZmock_populate_tracer
  real *pin = nullptr;
  size_t pin_cap = 0;           // tracers per array
  hipEvent_t ev_pos = nullptr, ev_xyz = nullptr;
  // last PDF
  size_t nmax = 0, ncell = 0, ntot = 0, ncell0 = 0, ntot0 = 0;
  std::vector<size_t> pdf;
  unsigned long long hist_h[EZ_NUM_LOGB_BIN] = {0};
  // populate with the selection on the device: control block, pinned
  // staging (parameters in, results out), spectrum seed of a replayed mock
  PopCtl *ctl_d = nullptr;
  MrgState *seed_d = nullptr;
  struct PopIO {
    PopCtl in;
    MrgState seed;
    uint64_t num[EZ_MAX_NUM_TRACER_PER_CELL + 1];
    PopCtl out;
    uint64_t cnt[EZ_MAX_NUM_TRACER_PER_CELL + 1];
    unsigned long long cflag;
  } *pio = nullptr;
  bool gen_fused = true;                // spectrum generated inside the first FFT pass
  int devsel = 1;                       // 2: capacities of half the expected counts
  // path counters: device-selection populates, host-driven ones, capacity
  // retries, graph replays, graph captures
  uint64_t npath[5] = {0, 0, 0, 0, 0};
  uint64_t last_cnt = 0;                // candidates of the last device selection
  // captured mock (EZmock_ext_mock): field and populate graphs, and what
  // they were captured with
  hipGraphExec_t g_field = nullptr, g_pop = nullptr;
  bool use_graph = true;
  struct GraphKey {
    long pk_version = -1; double vfac = 0;
    bool fixamp = false, iphase = false, keep_spec = false, fft_streams = false;
    rkey *keys = nullptr; size_t keys_cap = 0; real *cat0 = nullptr; size_t cat_cap = 0;
    uint32_t *cic_list = nullptr; size_t cic_list_cap = 0; uint32_t fcap = 0;
    int cic_path = 0; bool cic_fixed_ok = false, gen_fused = false;
    int sched[3] = {0, 0, 0};
    bool operator==(const GraphKey &o) const {
      return pk_version == o.pk_version && vfac == o.vfac && fixamp == o.fixamp &&
          iphase == o.iphase && keep_spec == o.keep_spec && fft_streams == o.fft_streams &&
          keys == o.keys && keys_cap == o.keys_cap && cat0 == o.cat0 && cat_cap == o.cat_cap &&
          cic_list == o.cic_list && cic_list_cap == o.cic_list_cap && fcap == o.fcap &&
          cic_path == o.cic_path && cic_fixed_ok == o.cic_fixed_ok && gen_fused == o.gen_fused &&
          sched[0] == o.sched[0] && sched[1] == o.sched[1] && sched[2] == o.sched[2];
    }
  } gkey;
  // options
  bool keep_spec = false, profile = false;
  std::vector<Timer> pending;
  std::map<std::string, std::pair<double, long>> times;
};

std::atomic<long> g_pk_version{0};   // P(k) tables of all instances

std::shared_mutex g_api_mu;
thread_local int t_api_depth = 0;

Conf *CONF(EZMOCK *ez) { return (Conf *) ez->conf; }
RngConf *RNG(EZMOCK *ez) { return (RngConf *) ez->rng; }
Cosmo *COSMO(EZMOCK *ez) { return (Cosmo *) ez->cosmo; }
PkTab *PK(EZMOCK *ez) { return (PkTab *) ez->pk; }
Par *PAR(EZMOCK *ez) { return (Par *) ez->par; }
Mesh *MESH(EZMOCK *ez) { return (Mesh *) ez->mesh; }

// ------------------------------------------------------- device helpers
std::mutex g_jtab_mu;
std::map<int, uint32_t *> g_jtab;

int device_ready(Mesh *m) {
  if (m->dev_ready) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    fprintf(stderr, "EZmock: no HIP device available; the density field and "
        "tracer stages run only on the GPU\n");
    return EZMOCK_ERR_UNKNOWN;
  }
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  HIP_OK(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
  for (int i = 0; i < 2; i++) {
    HIP_OK(hipStreamCreateWithFlags(&m->sx[i], hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&m->ev_join[i], hipEventDisableTiming));
  }
  HIP_OK(hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming));
  HIP_OK(hipEventCreateWithFlags(&m->ev_pos, hipEventDisableTiming));
  HIP_OK(hipEventCreateWithFlags(&m->ev_xyz, hipEventDisableTiming));
  {
    std::lock_guard<std::mutex> lk(g_jtab_mu);
    auto it = g_jtab.find(dev);
    if (it == g_jtab.end()) {
      uint32_t *d = nullptr;
      HIP_OK(hipMalloc(&d, MRG_JUMP_WORDS * sizeof(uint32_t)));
      HIP_OK(hipMemcpy(d, mrg_host_jump_table(), MRG_JUMP_WORDS * sizeof(uint32_t),
          hipMemcpyHostToDevice));
      g_jtab[dev] = d;
      m->jtab = d;
    }
    else m->jtab = it->second;
  }
  m->dev_ready = true;
  return 0;
}

template <class T> int dalloc(T **p, size_t count) {
  if (*p) return 0;
  if (hipMalloc((void **) p, count * sizeof(T) + 64) != hipSuccess) {
    *p = nullptr;
    (void) hipGetLastError();
    fprintf(stderr, "EZmock: device allocation of %zu bytes failed\n",
        count * sizeof(T));
    return EZMOCK_ERR_MEMORY;
  }
  return 0;
}

template <class T> void dfree(T **p) {
  if (*p) (void) hipFree((void *) *p);
  *p = nullptr;
}

void free_tracer_scratch(Mesh *m) {
  dfree(&m->keys); dfree(&m->keys_tmp); m->keys_cap = 0;
  dfree(&m->sort_gh); dfree(&m->sort_go); dfree(&m->sort_scan);
}

void drop_graphs(Mesh *m) {
  if (m->g_field) (void) hipGraphExecDestroy(m->g_field);
  if (m->g_pop) (void) hipGraphExecDestroy(m->g_pop);
  m->g_field = m->g_pop = nullptr;
}

void free_mesh_device(Mesh *m) {
  if (m->smt) (void) hipStreamSynchronize(m->smt);   // window builds ahead of use
  drop_graphs(m);
  dfree(&m->ctl_d); dfree(&m->seed_d);
  if (m->pio) { (void) hipHostFree(m->pio); m->pio = nullptr; }
  for (int i = 0; i < 3; i++) dfree(&m->psi[i]);
  dfree(&m->vel4); dfree(&m->gen_base);
  dfree(&m->rho); dfree(&m->ny); dfree(&m->slot0); dfree(&m->rhot); dfree(&m->nt);
  dfree(&m->ptab); dfree(&m->twN); dfree(&m->twh); dfree(&m->spec);
  dfree(&m->blockcnt); dfree(&m->blockoff); dfree(&m->scan_tmp);
  dfree(&m->hist); dfree(&m->counter); dfree(&m->blockhist);
  dfree(&m->cic_counts); dfree(&m->cic_list); m->cic_list_cap = 0; dfree(&m->cic_off);
  dfree(&m->cic_cur); dfree(&m->cic_scan);
  dfree(&m->cic_bnd); dfree(&m->cic_gkey); dfree(&m->cic_ovf);
  if (m->cic_status_h) { (void) hipHostFree(m->cic_status_h); m->cic_status_h = nullptr; }
  if (m->cic_ovf_h) { (void) hipHostFree(m->cic_ovf_h); m->cic_ovf_h = nullptr; }
  free_tracer_scratch(m);
  dfree(&m->d_num); dfree(&m->d_cnt); dfree(&m->d_res); dfree(&m->d_vb);
  dfree(&m->d_rsd);
  dfree(&m->pk_grid[0]); dfree(&m->pk_grid[1]); dfree(&m->pk_ny); dfree(&m->pk_acc);
  dfree(&m->mtw); m->mtw_cap = 0; dfree(&m->mtw2); m->mtw2_cap = 0;
  dfree(&m->att_np); dfree(&m->att_bcnt); dfree(&m->att_cinfo); dfree(&m->att_plist);
  dfree(&m->att_ovf); dfree(&m->att_cls); dfree(&m->att_sel); dfree(&m->att_rest);
  for (int i = 0; i < 3; i++) dfree(&m->att_boff[i]);
  dfree(&m->att_stmp); dfree(&m->att_draws);
  m->att_ccap = m->att_pcap = 0;
  for (int i = 0; i < 6; i++) dfree(&m->cat[i]);
  m->cat_cap = 0;
  for (auto &t : m->pending) { (void) hipEventDestroy(t.a); (void) hipEventDestroy(t.b); }
  m->pending.clear();
  if (m->st) (void) hipStreamDestroy(m->st);
  m->st = nullptr;
  for (int i = 0; i < 2; i++) {
    if (m->sx[i]) (void) hipStreamDestroy(m->sx[i]);
    if (m->ev_join[i]) (void) hipEventDestroy(m->ev_join[i]);
    m->sx[i] = nullptr;
    m->ev_join[i] = nullptr;
  }
  if (m->ev_fork) (void) hipEventDestroy(m->ev_fork);
  m->ev_fork = nullptr;
  for (auto &w : m->mtwin) {
    dfree(&w.win);
    dfree(&w.scr);
    w.cap = w.nseg = 0;
    w.first = ~0ull;
    if (w.ready) (void) hipEventDestroy(w.ready);
    if (w.used) (void) hipEventDestroy(w.used);
    w.ready = w.used = nullptr;
  }
  if (m->smt) (void) hipStreamDestroy(m->smt);
  m->smt = nullptr;
  if (m->ev_pos) (void) hipEventDestroy(m->ev_pos);
  if (m->ev_xyz) (void) hipEventDestroy(m->ev_xyz);
  m->ev_pos = m->ev_xyz = nullptr;
  if (m->pin) (void) hipHostFree(m->pin);
  m->pin = nullptr;
  m->pin_cap = 0;
}

// Kernel-class timing with HIP events on the instance stream (option 2).
struct Scope {
  Mesh *m; Timer t; bool on; hipStream_t s;
  Scope(Mesh *mm, const char *name, hipStream_t ss = nullptr)
      : m(mm), on(mm->profile), s(ss ? ss : mm->st) {
    if (!on) return;
    t.name = name;
    (void) hipEventCreate(&t.a);
    (void) hipEventCreate(&t.b);
    (void) hipEventRecord(t.a, s);
  }
  ~Scope() {
    if (!on) return;
    (void) hipEventRecord(t.b, s);
    m->pending.push_back(t);
  }
};

int sync_and_collect(Mesh *m) {
  HIP_OK(hipStreamSynchronize(m->st));
  for (auto &t : m->pending) {
    float ms = 0;
    (void) hipEventElapsedTime(&ms, t.a, t.b);
    auto &e = m->times[t.name];
    e.first += ms;
    e.second += 1;
    (void) hipEventDestroy(t.a);
    (void) hipEventDestroy(t.b);
  }
  m->pending.clear();
  return 0;
}

#define KCHK(expr)                                                          \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      fprintf(stderr, "EZmock: kernel launch failed (%s) at %s:%d\n",       \
              hipGetErrorString(e_), __FILE__, __LINE__);                   \
      return EZMOCK_ERR_UNKNOWN;                                            \
    }                                                                       \
  } while (0)

int ensure_twiddles(Mesh *m, int Ng) {
  if (m->twN) return 0;
  const int Ngh = Ng / 2;
  int e;
  if ((e = dalloc(&m->twN, Ng)) || (e = dalloc(&m->twh, Ngh))) return e;
  std::vector<cplx> a(Ng), b(Ngh);
  for (int i = 0; i < Ng; i++) {
    const double t = 2 * EZ_PI * (double) i / Ng;
    a[i] = cplx{(real) cos(t), (real) sin(t)};
  }
  for (int i = 0; i < Ngh; i++) {
    const double t = 2 * EZ_PI * (double) i / Ngh;
    b[i] = cplx{(real) cos(t), (real) sin(t)};
  }
  HIP_OK(hipMemcpy(m->twN, a.data(), sizeof(cplx) * Ng, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(m->twh, b.data(), sizeof(cplx) * Ngh, hipMemcpyHostToDevice));
  return 0;
}

// CIC painting of the displaced particles (cic.hip): density into rho
// (unless null) and, when with_mom, the three momentum fields.
void cic_pending(Mesh *m, int kind, int Ng, real igs, real *rho, bool with_mom) {
  m->cic_pend.kind = kind;
  m->cic_pend.Ng = Ng;
  m->cic_pend.igs = igs;
  m->cic_pend.rho = rho;
  m->cic_pend.mom = with_mom;
}

// List entries per tile with fixed ranges: 4 per cell of the tile, about
// three times the mean (1.35 per cell); dense tiles rarely exceed twice it
// at cell sizes of a few Mpc/h.  After an overflow the ranges are resized
// from the exact counts: up to CIC_FIXED_MAX_PER_CELL always, and up to
// CIC_FIXED_MEM_PER_CELL where the device keeps a quarter of its memory
// free after the larger lists (small cells cluster more: 1024^3 in a
// 2500 Mpc/h box needs about 8).
#define CIC_FIXED_PER_CELL 4
#define CIC_FIXED_MAX_PER_CELL 6
#define CIC_FIXED_MEM_PER_CELL 16

// CIC by particle binning: scatter into fixed ranges per tile and tiles,
// checked for overflow after the next synchronisation (no count pass, no
// host synchronisation); or count, scan, scatter, tiles with exact ranges.
int paint_cic_binned(Mesh *m, int Ng, real igs, real *rho, bool with_mom) {
  const CicGeom g = cic_geometry(Ng, igs);
  const size_t nk = cic_nkeys(g);
  int e;
  if ((e = dalloc(&m->cic_counts, nk)) || (e = dalloc(&m->cic_off, nk + 1)) ||
      (e = dalloc(&m->cic_cur, nk)) || (e = dalloc(&m->cic_scan, scan_tmp_words(nk) + 8)))
    return e;
  const uint32_t fcap = m->cic_fcap ? m->cic_fcap
                                     : (uint32_t) (CIC_FIXED_PER_CELL * g.TX * g.TY * g.TZ);
  if (m->cic_path == 0 && m->cic_fixed_ok && nk < (1ull << 27)) {
    const size_t need = nk * fcap;
    if (m->cic_list_cap < need) {
      dfree(&m->cic_list);
      m->cic_list_cap = 0;
      // fixed ranges need more memory than exact ones (about 1.35 N
      // entries); where they do not fit (allocation failure or the
      // instance's list limit), this and later fields use exact ranges
      if ((m->cic_list_limit && need * sizeof(uint32_t) > m->cic_list_limit) ||
          dalloc(&m->cic_list, need)) {
        m->cic_fixed_ok = false;
        m->cic_fcap = 0;
        return paint_cic_binned(m, Ng, igs, rho, with_mom);
      }
      m->cic_list_cap = need;
    }
    if ((e = dalloc(&m->cic_ovf, 1))) return e;
    if (!m->cic_ovf_h &&
        hipHostMalloc((void **) &m->cic_ovf_h, sizeof(uint32_t)) != hipSuccess) {
      m->cic_ovf_h = nullptr;
      return EZMOCK_ERR_MEMORY;
    }
    {
      Scope s(m, "cic_bin_scatter");
      KCHK(launch_cic_fixed_init(m->cic_cur, g, fcap, m->cic_ovf, m->st));
      KCHK(launch_cic_scatter(m->psi[0], m->psi[1], m->psi[2], g, m->cic_cur,
          m->cic_list, fcap, m->cic_ovf, m->st));
    }
    {
      Scope s(m, with_mom ? (rho ? "cic_tiles_rho_mom" : "cic_tiles_mom") : "cic_tiles_rho");
      KCHK(launch_cic_tiles(m->psi[0], m->psi[1], m->psi[2], g, m->cic_cur,
          m->cic_list, rho, m->rho, with_mom ? m->vel4 : nullptr, fcap, m->st));
    }
    HIP_OK(hipMemcpyAsync(m->cic_ovf_h, m->cic_ovf, sizeof(uint32_t),
        hipMemcpyDeviceToHost, m->st));
    cic_pending(m, 2, Ng, igs, rho, with_mom);
    return 0;
  }
  {
    Scope s(m, "cic_bin_count");
    KCHK(launch_cic_count(m->psi[0], m->psi[1], m->psi[2], g, m->cic_counts,
        m->cic_off, m->cic_cur, m->cic_scan, m->st));
  }
  uint64_t total = 0;
  HIP_OK(hipMemcpyAsync(&total, m->cic_off + nk, sizeof(uint64_t),
      hipMemcpyDeviceToHost, m->st));
  if ((e = sync_and_collect(m))) return e;
  if (total < g.N || total > 8 * g.N) return EZMOCK_ERR_UNKNOWN;
  if (m->cic_list_cap < total) {
    dfree(&m->cic_list);
    const size_t cap = total + total / 8;
    if ((e = dalloc(&m->cic_list, cap))) { m->cic_list_cap = 0; return e; }
    m->cic_list_cap = cap;
  }
  {
    Scope s(m, "cic_bin_scatter");
    KCHK(launch_cic_scatter(m->psi[0], m->psi[1], m->psi[2], g, m->cic_cur,
        m->cic_list, 0u, nullptr, m->st));
  }
  {
    Scope s(m, with_mom ? (rho ? "cic_tiles_rho_mom" : "cic_tiles_mom") : "cic_tiles_rho");
    KCHK(launch_cic_tiles(m->psi[0], m->psi[1], m->psi[2], g, m->cic_off,
        m->cic_list, rho, m->rho, with_mom ? m->vel4 : nullptr, 0u, m->st));
  }
  return 0;
}

// Density (into rho, unless null) and, with with_mom, the normalised
// momentum fields of the displacement fields psi (dens_field.c:233-310,
// pop_tracer.c:593-682).  By particle binning, or (option) by tile scan
// where the displacement range allows it.  The painting is complete only
// after cic_finish(), called after the next synchronisation of the stream,
// which repeats it when the scan declined or a fixed list range overflowed.
int paint_cic(Mesh *m, int Ng, real igs, real *rho, bool with_mom) {
  const CicGeom g = cic_geometry(Ng, igs);
  if (g.N >= 0xffffffffull) {
    fprintf(stderr, "EZmock: grids of 2^32 cells or more are not supported\n");
    return EZMOCK_ERR_ARG_NGRID;
  }
  int e;
  if (with_mom && (e = dalloc(&m->vel4, g.N))) return e;
  m->cic_pend.kind = 0;
  if (m->cic_path != 2) return paint_cic_binned(m, Ng, igs, rho, with_mom);
  if ((e = dalloc(&m->cic_bnd, 2 * cic_nblocks(g))) ||
      (e = dalloc(&m->cic_gkey, CIC_GKEY_WORDS)))
    return e;
  if (!m->cic_status_h &&
      hipHostMalloc((void **) &m->cic_status_h, sizeof(rkey)) != hipSuccess) {
    m->cic_status_h = nullptr;
    return EZMOCK_ERR_MEMORY;
  }
  {
    Scope s(m, "cic_bounds");
    KCHK(launch_cic_bounds(m->psi[0], m->psi[1], m->psi[2], g, m->cic_bnd, m->cic_gkey,
        m->st));
  }
  {
    Scope s(m, with_mom ? (rho ? "cic_scan_rho_mom" : "cic_scan_mom") : "cic_scan_rho");
    KCHK(launch_cic_scan(m->psi[0], m->psi[1], m->psi[2], g, m->cic_bnd, m->cic_gkey,
        rho, m->rho, with_mom ? m->vel4 : nullptr, m->st));
  }
  HIP_OK(hipMemcpyAsync(m->cic_status_h, m->cic_gkey + CIC_GKEY_STATUS, sizeof(rkey),
      hipMemcpyDeviceToHost, m->st));
  cic_pending(m, 1, Ng, igs, rho, with_mom);
  return 0;
}

// After a stream synchronisation that follows paint_cic: paint again when
// the tile scan declined (by binning) or a fixed list range overflowed
// (with exact ranges, and so on for this instance).
int cic_finish(Mesh *m) {
  while (m->cic_pend.kind) {
    const int kind = m->cic_pend.kind;
    m->cic_pend.kind = 0;
    if (kind == 1 && !*m->cic_status_h) return 0;
    if (kind == 2) {
      if (!*m->cic_ovf_h) return 0;
      m->cic_fixed_ok = false;
    }
    int e;
    if ((e = paint_cic_binned(m, m->cic_pend.Ng, m->cic_pend.igs, m->cic_pend.rho,
                              m->cic_pend.mom)) ||
        (e = sync_and_collect(m)))
      return e;
    if (kind == 2) {
      // ranges for the next fields of this instance: 1.25 x the fullest
      // tile of this one, within the memory bound
      const CicGeom g = cic_geometry(m->cic_pend.Ng, m->cic_pend.igs);
      const size_t nk = cic_nkeys(g);
      std::vector<uint32_t> cnt(nk);
      HIP_OK(hipMemcpy(cnt.data(), m->cic_counts, nk * sizeof(uint32_t),
          hipMemcpyDeviceToHost));
      uint32_t mx = 0;
      for (uint32_t c : cnt) mx = c > mx ? c : mx;
      const uint32_t cap = ((mx + mx / 4) + 63) / 64 * 64;
      const size_t tc = (size_t) g.TX * g.TY * g.TZ;
      bool ok = cap <= CIC_FIXED_MAX_PER_CELL * tc;
      if (!ok && cap <= CIC_FIXED_MEM_PER_CELL * tc) {
        // the exact list is freed before the fixed one is allocated
        size_t fr = 0, tot = 0;
        const size_t need = nk * (size_t) cap * sizeof(uint32_t);
        const size_t have = m->cic_list_cap * sizeof(uint32_t);
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && need <= have + fr && 
            have + fr - need >= tot / 4)
          ok = true;
        else (void) hipGetLastError();
      }
      if (ok && m->cic_list_limit && nk * (size_t) cap * sizeof(uint32_t) > m->cic_list_limit)
        ok = false;
      // the larger list is allocated now, in the field that overflowed, so
      // that the next fields run without an allocation (hipMalloc of tens
      // of GB takes about a second)
      if (ok && m->cic_list_cap < nk * (size_t) cap) {
        dfree(&m->cic_list);
        m->cic_list_cap = 0;
        if (dalloc(&m->cic_list, nk * (size_t) cap)) ok = false;
        else m->cic_list_cap = nk * (size_t) cap;
      }
      if (ok) {
        m->cic_fcap = cap;
        m->cic_fixed_ok = true;
      }
    }
  }
  return 0;
}

// The (j, rows) passes of the c2r of a displacement component whose first
// pass (axis i, fft_first_pass3 from the generated spectrum) is done.
int c2r_finish(Mesh *m, int Ng, real *out, hipStream_t st) {
  cplx *buf = (cplx *) out;
  {
    Scope s(m, "fft_pass_j", st);
    KCHK(fft_strided_pass(buf, nullptr, nullptr, m->twN, Ng, 1, +1, 0, 0, st, -1));
  }
  {
    Scope s(m, "fft_rows_c2r", st);
    KCHK(fft_row_pass(buf, m->twN, m->twh, Ng, +1, st));
  }
  return 0;
}

// r2c of a real field held in buf (in place): packed result, A/B unpacked.
int r2c_field(Mesh *m, int Ng, real *field) {
  cplx *buf = (cplx *) field;
  {
    Scope s(m, "fft_rows_r2c");
    KCHK(fft_row_pass(buf, m->twN, m->twh, Ng, -1, m->st));
  }
  {
    Scope s(m, "fft_pass_j");
    KCHK(fft_strided_pass(buf, nullptr, nullptr, m->twN, Ng, 1, -1, 0, 0, m->st));
  }
  {
    Scope s(m, "fft_pass_i");
    KCHK(fft_strided_pass(buf, nullptr, nullptr, m->twN, Ng, 0, -1, 0, 0, m->st));
  }
  {
    Scope s(m, "r2c_unpack");
    KCHK(launch_unpack_r2c(buf, m->ny, Ng, m->st));
  }
  return 0;
}

// The stream that builds MT19937 windows ahead of use (the instance stream
// while the per-kernel timers are on), created at its first use.
hipStream_t mt_ahead_stream(Mesh *m) {
  if (m->profile) return m->st;
  if (!m->smt && hipStreamCreateWithFlags(&m->smt, hipStreamNonBlocking) != hipSuccess) {
    m->smt = nullptr;
    return m->st;
  }
  return m->smt;
}

// MT19937 segment windows of draws [first, first + n) for stream use k,
// kept from an earlier build when seed and first match and it covers n
// draws; otherwise built on stream st, whose completion `ready` marks.
int mt_windows(EZMOCK *ez, Mesh *m, int k, uint64_t first, uint64_t n, hipStream_t st) {
  Mesh::MtWin &w = m->mtwin[k];
  const uint64_t seed = RNG(ez)->seed;
  const size_t nseg = mt_segments(2 * n);
  if (w.seed == seed && w.first == first && w.nseg >= nseg) return 0;
  if (!w.ready) {
    HIP_OK(hipEventCreateWithFlags(&w.ready, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&w.used, hipEventDisableTiming));
    HIP_OK(hipEventRecord(w.used, st));
  }
  HIP_OK(hipStreamWaitEvent(st, w.used, 0));   // the last fill from them is done
  if (w.cap < nseg) {
    HIP_OK(hipEventSynchronize(w.used));
    dfree(&w.win);
    dfree(&w.scr);
    w.cap = 0;
    int e = dalloc(&w.win, mt_window_words(nseg));
    if (!e) e = dalloc(&w.scr, mt_window_words(nseg));
    if (e) return e;
    w.cap = nseg;
  }
  w.first = ~0ull;                      // invalid until the build is enqueued
  {
    Scope s(m, "mt_stream", st);
    if (mt_stream_windows(seed, 2 * first, nseg, w.win, w.scr, st) != hipSuccess)
      return EZMOCK_ERR_RNG_JUMP;
  }
  HIP_OK(hipEventRecord(w.ready, st));
  w.seed = seed;
  w.first = first;
  w.nseg = nseg;
  return 0;
}

// MT19937 words of draws first .. first + n - 1 (two words per draw) into
// word buffer `slot` (0 or 1) on stream st (default: the instance stream),
// from the windows of stream use k.
int mt_words(EZMOCK *ez, Mesh *m, int k, uint64_t first, uint64_t n, const uint32_t **out,
    int slot = 0, hipStream_t st = nullptr) {
  *out = nullptr;
  if (RNG(ez)->type != 1 || n == 0) return 0;
  uint32_t *&buf = slot ? m->mtw2 : m->mtw;
  size_t &cap = slot ? m->mtw2_cap : m->mtw_cap;
  if (!st) st = m->st;
  if (cap < 2 * n) {
    dfree(&buf);
    cap = 0;
    int e = dalloc(&buf, 2 * n);
    if (e) return e;
    cap = 2 * n;
  }
  int e = mt_windows(ez, m, k, first, n, st);
  if (e) return e;
  HIP_OK(hipStreamWaitEvent(st, m->mtwin[k].ready, 0));
  {
    Scope s(m, "mt_stream", st);
    if (mt_stream_fill(m->mtwin[k].win, 2 * n, buf, st) != hipSuccess)
      return EZMOCK_ERR_RNG_JUMP;
  }
  HIP_OK(hipEventRecord(m->mtwin[k].used, st));
  *out = buf;
  return 0;
}

// pop_tracer.c:94-133
int eval_pdf(size_t nexp, real pbase, size_t ncells, std::vector<size_t> &num,
    size_t &ncell, size_t &ntot) {
  size_t n;
  double A = 0, sum = 0, power = pbase;
  for (n = 1; n <= nexp; n++) {
    sum += power * n;
    A = nexp / sum;
    power *= pbase;
    if (A * power < 0.5) break;
  }
  if (n > EZ_MAX_NUM_TRACER_PER_CELL) return EZMOCK_ERR_RHO_PBASE;
  num.assign(n, 0);
  size_t nc = 0, nobj = 0;
  for (size_t i = 0; i < n; i++) {
    num[i] = (size_t) round(A * pow(pbase, (double) (n - i)));
    nc += num[i];
    nobj += num[i] * (n - i);
  }
  if (nc > ncells || nobj == 0) return EZMOCK_ERR_PAR_PBASE;
  ncell = nc;
  ntot = nobj;
  return 0;
}

int ensure_cat(Mesh *m, size_t n) {
  if (m->cat_cap >= n && m->cat[0]) return 0;
  for (int i = 0; i < 6; i++) dfree(&m->cat[i]);
  m->cat_cap = 0;
  const size_t cap = n + n / 8 + 1024;
  for (int i = 0; i < 6; i++) {
    int e = dalloc(&m->cat[i], cap);
    if (e) return e;
  }
  m->cat_cap = cap;
  return 0;
}

// Exclusive scan of per-block u32 counts into out; returns the total.
int scan_total(Mesh *m, const uint32_t *in, uint64_t *out, size_t nblk, uint64_t *tmp,
    uint64_t *total) {
  KCHK(launch_scan_u32(in, out, nblk, tmp, m->st));
  const size_t nch = (nblk + EZ_CELLS_PER_BLOCK - 1) / EZ_CELLS_PER_BLOCK;
  HIP_OK(hipMemcpyAsync(total, tmp + nch, sizeof(uint64_t), hipMemcpyDeviceToHost, m->st));
  return sync_and_collect(m);
}

// Tracers attached to particles (att_part, pop_tracer.c:517-582 and
// 804-879), written to the front of the catalogue; *iv = their number, and
// *rest / *restoff the tracers per cell left for uniform positions and
// their block offsets.  The particle counts, classes and the writing of the
// catalogue are parallel passes; the selection itself is the reference's
// serial, stream-ordered Bentley selection, run on the host over the
// particles of contested cells only (cells with more particles than
// tracers; the others take all their particles without a draw).  A cell
// keeps max(nt - np, 0) tracers for uniform positions whatever the draws.
int attach_particles(EZMOCK *ez, Mesh *m, size_t ntot, size_t *iv,
    const int16_t **rest, const uint64_t **restoff) {
  Conf *conf = CONF(ez);
  const int Ng = conf->Ng;
  const size_t N = (size_t) Ng * Ng * Ng;
  const size_t nblk = (N + EZ_CELLS_PER_BLOCK - 1) / EZ_CELLS_PER_BLOCK;
  AttGeom ag;
  ag.Ng = Ng;
  ag.igs = conf->Ng / conf->Lbox;
  ag.gs = conf->Lbox / conf->Ng;
  ag.N = N;
  ag.fNg = fastdiv_make((uint32_t) Ng);
  int e;
  if ((e = dalloc(&m->att_np, N)) || (e = dalloc(&m->att_cls, N)) ||
      (e = dalloc(&m->att_rest, N)) || (e = dalloc(&m->att_bcnt, nblk)) ||
      (e = dalloc(&m->att_boff[0], nblk)) || (e = dalloc(&m->att_boff[1], nblk)) ||
      (e = dalloc(&m->att_boff[2], nblk)) ||
      (e = dalloc(&m->att_stmp, scan_tmp_words(nblk) + 8)) || (e = dalloc(&m->att_ovf, 1)))
    return e;
  // particles per tracer cell, with the reference's uint16 limit
  uint32_t ovf = 0;
  {
    Scope s(m, "att_count");
    KCHK(launch_zero(m->att_np, N * sizeof(uint32_t), m->st));
    KCHK(launch_zero(m->att_ovf, sizeof(uint32_t), m->st));
    KCHK(launch_att_count(m->psi[0], m->psi[1], m->psi[2], ag, m->nt, m->att_np, m->att_ovf,
        m->st));
  }
  HIP_OK(hipMemcpyAsync(&ovf, m->att_ovf, sizeof(uint32_t), hipMemcpyDeviceToHost, m->st));
  if ((e = sync_and_collect(m))) return e;
  if (ovf) return EZMOCK_ERR_RHO_NPART;
  // contested cells and particles
  uint64_t ncc = 0, ncont = 0;
  {
    Scope s(m, "att_cells");
    KCHK(launch_att_cells(m->nt, m->att_np, N, m->att_bcnt, nullptr, nullptr, m->st));
  }
  if ((e = scan_total(m, m->att_bcnt, m->att_boff[0], nblk, m->att_stmp, &ncc))) return e;
  if (m->att_ccap < ncc + 1) {
    dfree(&m->att_cinfo);
    m->att_ccap = 0;
    if ((e = dalloc(&m->att_cinfo, ncc + 1))) return e;
    m->att_ccap = ncc + 1;
  }
  {
    Scope s(m, "att_cells");
    KCHK(launch_att_cells(m->nt, m->att_np, N, m->att_bcnt, m->att_boff[0], m->att_cinfo,
        m->st));
  }
  {
    Scope s(m, "att_classify");
    KCHK(launch_att_classify(m->psi[0], m->psi[1], m->psi[2], ag, m->nt, m->att_np,
        m->att_cls, m->att_bcnt, m->st));
  }
  if ((e = scan_total(m, m->att_bcnt, m->att_boff[1], nblk, m->att_stmp, &ncont))) return e;
  std::vector<uint8_t> selh(ncont);
  if (ncont) {
    if (m->att_pcap < ncont) {
      dfree(&m->att_plist); dfree(&m->att_sel); dfree(&m->att_draws);
      m->att_pcap = 0;
      if ((e = dalloc(&m->att_plist, ncont)) || (e = dalloc(&m->att_sel, ncont)) ||
          (e = dalloc(&m->att_draws, ncont)))
        return e;
      m->att_pcap = ncont;
    }
    {
      Scope s(m, "att_list");
      KCHK(launch_att_list(m->psi[0], m->psi[1], m->psi[2], ag, m->att_np, m->att_cls,
          m->att_boff[1], m->att_plist, m->st));
    }
    // at most one draw per contested particle, from stream position 4N
    // (pop_tracer.c:781, 846)
    if (RNG(ez)->type == 1) {
      const uint32_t *w = nullptr;
      if ((e = mt_words(ez, m, 4, 4 * (uint64_t) N, ncont, &w))) return e;
      KCHK(launch_uniform_mt(w, ncont, 0, m->att_draws, m->st));
    }
    else {
      MrgState st0;
      mrg_seed_state(RNG(ez)->seed, st0);
      KCHK(launch_uniform(st0, 4 * (uint64_t) N, ncont, m->jtab, m->att_draws, m->st));
    }
    std::vector<uint32_t> pl(ncont), ci(ncc);
    std::vector<double> dr(ncont);
    HIP_OK(hipMemcpyAsync(pl.data(), m->att_plist, ncont * sizeof(uint32_t),
        hipMemcpyDeviceToHost, m->st));
    HIP_OK(hipMemcpyAsync(ci.data(), m->att_cinfo, ncc * sizeof(uint32_t),
        hipMemcpyDeviceToHost, m->st));
    HIP_OK(hipMemcpyAsync(dr.data(), m->att_draws, ncont * sizeof(double),
        hipMemcpyDeviceToHost, m->st));
    if ((e = sync_and_collect(m))) return e;
    // pop_tracer.c:841-854, in Lagrangian order
    std::vector<uint16_t> tt(ncc), nn(ncc);
    for (size_t c = 0; c < ncc; c++) {
      tt[c] = (uint16_t) (ci[c] & 0xffffu);
      nn[c] = (uint16_t) (ci[c] >> 16);
    }
    size_t d = 0;
    for (size_t k = 0; k < ncont; k++) {
      const uint32_t c = pl[k];
      uint8_t take = 0;
      if (tt[c] != 0) {
        bool skip = false;
        if (tt[c] < nn[c]) {
          const double ran = dr[d++];
          if (ran * nn[c] >= tt[c]) {
            nn[c]--;
            skip = true;
          }
        }
        if (!skip) {
          tt[c]--;
          nn[c]--;
          take = 1;
        }
      }
      selh[k] = take;
    }
    HIP_OK(hipMemcpyAsync(m->att_sel, selh.data(), ncont, hipMemcpyHostToDevice, m->st));
  }
  // the attached tracers at the front of the catalogue, in Lagrangian order
  uint64_t nsel = 0, nrest = 0;
  real *const cat[6] = {m->cat[0], m->cat[1], m->cat[2], m->cat[3], m->cat[4], m->cat[5]};
  {
    Scope s(m, "att_select");
    KCHK(launch_att_select(m->psi[0], m->psi[1], m->psi[2], ag, COSMO(ez)->vfac, m->att_cls,
        m->att_boff[1], m->att_sel, m->att_bcnt, nullptr, cat, m->st));
  }
  if ((e = scan_total(m, m->att_bcnt, m->att_boff[0], nblk, m->att_stmp, &nsel))) return e;
  if (nsel > ntot) return EZMOCK_ERR_UNKNOWN;
  {
    Scope s(m, "att_select");
    KCHK(launch_att_select(m->psi[0], m->psi[1], m->psi[2], ag, COSMO(ez)->vfac, m->att_cls,
        m->att_boff[1], m->att_sel, nullptr, m->att_boff[0], cat, m->st));
    KCHK(launch_att_rest(m->nt, m->att_np, N, m->att_rest, m->att_bcnt, m->st));
  }
  if ((e = scan_total(m, m->att_bcnt, m->att_boff[2], nblk, m->att_stmp, &nrest))) return e;
  if (nsel + nrest != ntot) return EZMOCK_ERR_UNKNOWN;   // pop_tracer.c:989-993
  *iv = nsel;
  *rest = m->att_rest;
  *restoff = m->att_boff[2];
  return 0;
}

// Catalogue arrays for the caller (EZmock_populate_tracer): malloc'd host
// arrays filled from pinned staging buffers.
int ensure_keys(Mesh *m, size_t cap) {
  if (m->keys_cap >= cap) return 0;
  free_tracer_scratch(m);
  const size_t hw = sort_hist_words(cap);
  int e;
  if ((e = dalloc(&m->keys, cap)) || (e = dalloc(&m->keys_tmp, cap)) ||
      (e = dalloc(&m->sort_gh, hw)) || (e = dalloc(&m->sort_go, hw)) ||
      (e = dalloc(&m->sort_scan, scan_tmp_words(hw) + 8)))
    return e;
  m->keys_cap = cap;
  return 0;
}

// internal: a device-selection populate outgrew its capacities (the
// host-driven path then sizes them from the counts)
constexpr int EZ_RETRY = -1;

// Capacities of a device-selection populate.  Candidates (the cells of the
// threshold's log2 bin and above: 1.3 ntot at the median of a sweep of
// bias parameters, up to 3.4 ntot, and 2.8 ntot for the 512^3 bench mock)
// and tracers (ntot and the ties of the PDF walk) are bounded in terms of
// eval_pdf's ntot, which follows the expected number of tracers and barely
// the bias parameters, so that one allocation serves a sweep: the current
// buffers are kept while they hold 4 ntot + 64 Ki keys (at most N) and
// ntot + ntot/32 + 1024 tracers, else they grow to 6 ntot + 64 Ki and
// ntot + ntot/8 + 4096.  With alloc false: whether the current ones are kept.
bool dsel_capacity(Mesh *m, size_t N, size_t ntot, bool alloc, int *e) {
  size_t kacc = 4 * ntot + 65536, kcap = 6 * ntot + 65536;
  if (kacc > N) kacc = N;
  if (kcap > N) kcap = N;
  const size_t cacc = ntot + ntot / 32 + 1024;
  size_t ccap = ntot + ntot / 8 + 4096;
  if (m->devsel == 2) {                 // capacities too small: the redo path
    kacc = kcap = ntot / 4 + 1;
    ccap = ntot / 2 + 1;
  }
  const bool keep = m->keys_cap >= kacc && m->cat_cap >= cacc && m->cat[0];
  if (keep) return true;
  if (!alloc) return false;
  if (m->keys_cap < kacc && (*e = ensure_keys(m, kcap))) return false;
  if (m->cat_cap < cacc || !m->cat[0]) {
    for (int i = 0; i < 6; i++) dfree(&m->cat[i]);
    m->cat_cap = 0;
    for (int i = 0; i < 6; i++)
      if ((*e = dalloc(&m->cat[i], ccap))) return false;
    m->cat_cap = ccap;
  }
  return true;
}

// Refined candidates the sort covers: a third of the key capacity (2 ntot
// + 21 Ki keys at the capacity's growth; 0.95 ntot measured for the 512^3
// bench mock); more are redone by the host-driven path.
size_t dsel_kfcap(const Mesh *m) { return m->keys_cap / 3 ? m->keys_cap / 3 : 1; }

// The control block of a device-selection populate into the pinned
// staging, and its upload with the PDF bin sizes (stream order).  Every
// call of the engine ends with a synchronisation, so the staging is free
// again when the next call writes it.
int dsel_upload(Mesh *m, const BiasParams &bp, const std::vector<size_t> &num, size_t ncell,
    real sigma_v) {
  int e;
  if ((e = dalloc(&m->ctl_d, 1))) return e;
  if (!m->pio && hipHostMalloc((void **) &m->pio, sizeof(Mesh::PopIO)) != hipSuccess) {
    m->pio = nullptr;
    return EZMOCK_ERR_MEMORY;
  }
  PopCtl &c = m->pio->in;
  memset(&c, 0, sizeof c);
  const uint32_t *jt = mrg_host_jump_table();
  c.bp = bp;
  mrg_jump(c.bp.seed_state, 2 * (uint64_t) bp.N, jt);
  c.pos_seed = bp.seed_state;
  mrg_jump(c.pos_seed, 5 * (uint64_t) bp.N, jt);
  c.vel_seed = bp.seed_state;
  mrg_jump(c.vel_seed, 8 * (uint64_t) bp.N, jt);
  c.sigma_v = (double) sigma_v;
  c.ncell = ncell;
  c.kcap = m->keys_cap;
  c.kfcap = dsel_kfcap(m);
  c.ccap = m->cat_cap;
  c.nmax = (int) num.size();
  for (size_t p = 0; p < num.size(); p++) m->pio->num[p] = num[p];
  HIP_OK(hipMemcpyAsync(m->ctl_d, &c, sizeof c, hipMemcpyHostToDevice, m->st));
  HIP_OK(hipMemcpyAsync(m->d_num, m->pio->num, sizeof(uint64_t) * num.size(),
      hipMemcpyHostToDevice, m->st));
  return 0;
}

// The launches of a device-selection populate, pop_tracer.c:149-1007 (and
// with vel :593-753, :1043-1113): valid cells, bias model, threshold bin,
// candidates, sort, PDF walk, tracers per cell, positions (velocities),
// then the control block, the PDF bin counts and the compaction flag into
// the pinned staging.  Every count is read from the control block, so the
// sequence can be captured once and replayed for other seeds and
// parameters.
int dsel_launch(EZMOCK *ez, Mesh *m, bool vel) {
  Conf *conf = CONF(ez);
  const int Ng = conf->Ng;
  const size_t N = (size_t) Ng * Ng * Ng;
  const size_t nblk = (N + EZ_CELLS_PER_BLOCK - 1) / EZ_CELLS_PER_BLOCK;
  const size_t nch = (nblk + EZ_CELLS_PER_BLOCK - 1) / EZ_CELLS_PER_BLOCK;
  PopCtl *ctl = m->ctl_d;
  BiasParams bp{};                      // the grid size only: the rest from ctl
  bp.N = N;
  const MrgState none{};
  {
    Scope s(m, "bias_count");
    KCHK(launch_count_valid(m->rho, 0, N, m->blockcnt, m->st, ctl));
  }
  {
    Scope s(m, "scan");
    KCHK(launch_scan_u32(m->blockcnt, m->blockoff, nblk, m->scan_tmp, m->st));
  }
  KCHK(launch_zero(m->hist, EZ_NUM_LOGB_BIN * sizeof(unsigned long long), m->st));
  {
    Scope s(m, "bias");
    KCHK(launch_bias(m->rho, m->rhot, m->blockoff, m->hist, m->blockhist, bp, m->jtab,
        nullptr, m->st, ctl));
  }
  {
    Scope s(m, "select");
    KCHK(launch_select(m->hist, m->scan_tmp + nch, ctl, m->st));
  }
  KCHK(launch_zero(m->counter, 2 * sizeof(unsigned long long), m->st));
  {
    Scope s(m, "compact");
    KCHK(launch_block_cand(m->blockhist, nblk, 0, m->blockcnt, m->st, ctl));
    KCHK(launch_scan_u32(m->blockcnt, m->blockoff, nblk, m->scan_tmp, m->st));
    KCHK(launch_compact_at(m->rhot, 0, N, m->blockoff, m->blockcnt, m->keys, m->keys_cap,
        m->counter, m->st, ctl));
  }
  // the candidates at or above a threshold refined inside the threshold
  // bin, into keys_tmp (about a third of them at 512^3), then sorted there
  // (the refinement's block counts and offsets in the sort's buffers)
  {
    Scope s(m, "refine");
    KCHK(launch_refine(m->keys, m->keys_cap, ctl, m->sort_gh, m->sort_go, m->keys_tmp, m->st));
  }
  {
    Scope s(m, "sort");
    KCHK(launch_sort(m->keys_tmp, m->keys, dsel_kfcap(m), m->sort_gh, m->sort_go,
        m->sort_scan, m->st, &ctl->nfine));
  }
  {
    Scope s(m, "walk");
    KCHK(launch_walk(m->keys_tmp, m->keys_cap, m->d_num, 0, m->d_vb, m->d_cnt, m->d_res, m->st,
        ctl, m->scan_tmp + nch));
  }
  {
    Scope s(m, "assign_nt");
    KCHK(launch_assign_nt(m->rhot, N, m->d_vb, 0, m->nt, m->blockcnt, m->st, ctl));
  }
  {
    Scope s(m, "scan");
    KCHK(launch_scan_u32(m->blockcnt, m->blockoff, nblk, m->scan_tmp, m->st));
  }
  const real gs = conf->Lbox / conf->Ng;
  const real igs = conf->Ng / conf->Lbox;
  {
    Scope s(m, "positions");
    KCHK(launch_positions(m->nt, m->blockoff, Ng, conf->Lbox, gs, none, m->jtab, nullptr,
        m->cat[0], m->cat[1], m->cat[2], m->cat_cap, m->st, ctl));
  }
  if (vel) {
    Scope s(m, "velocities");
    KCHK(launch_velocities(m->cat[0], m->cat[1], m->cat[2], m->vel4, m->cat_cap, 0, Ng, igs,
        COSMO(ez)->vfac, 0, none, m->jtab, nullptr, m->cat[3], m->cat[4], m->cat[5], m->st,
        ctl));
  }
  HIP_OK(hipMemcpyAsync(&m->pio->out, ctl, sizeof(PopCtl), hipMemcpyDeviceToHost, m->st));
  HIP_OK(hipMemcpyAsync(m->pio->cnt, m->d_cnt, sizeof(m->pio->cnt), hipMemcpyDeviceToHost,
      m->st));
  HIP_OK(hipMemcpyAsync(&m->pio->cflag, m->counter, sizeof(unsigned long long),
      hipMemcpyDeviceToHost, m->st));
  return 0;
}

// After the synchronisation that follows dsel_launch: the errors of the
// host-driven path, in its order (EZ_RETRY where a capacity was short), and
// the PDF of the instance.
int dsel_collect(Mesh *m) {
  const PopCtl &r = m->pio->out;
  m->last_cnt = r.cnt;
  if (r.flags & POP_RHO_LESS) return EZMOCK_ERR_RHO_LESS;
  if (r.flags & (POP_KEYS_FULL | POP_CAT_FULL)) return EZ_RETRY;
  if ((r.flags & (POP_NO_BIN | POP_WALK_ERR | POP_CAND_ERR)) || m->pio->cflag)
    return EZMOCK_ERR_UNKNOWN;
  m->nmax = (size_t) r.nmax;
  m->ncell = r.ncell_out;
  m->ntot = r.ntot;
  m->pdf.assign(m->pio->cnt, m->pio->cnt + r.nmax);
  return 0;
}

struct HostCat {
  real *h[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  void release() {
    for (int i = 0; i < 6; i++) { free(h[i]); h[i] = nullptr; }
  }
};

// dst[a][0, n) = src[a][0, n) for arrays a in [a0, a1), in pieces of 1 M
// elements spread over up to nthread (at most 8) host threads.
void copy_arrays(real *const *dst, real *const *src, int a0, int a1, size_t n, int nthread) {
  const size_t CH = (size_t) 1 << 20;
  const size_t per = (n + CH - 1) / CH;              // pieces per array
  const size_t npiece = per * (size_t) (a1 - a0);
  int nt = nthread < 1 ? 1 : (nthread > 8 ? 8 : nthread);
  if ((size_t) nt > npiece) nt = npiece ? (int) npiece : 1;
  auto work = [&](int t) {
    for (size_t q = (size_t) t; q < npiece; q += (size_t) nt) {
      const int a = a0 + (int) (q / per);
      const size_t off = (q % per) * CH;
      const size_t len = n - off < CH ? n - off : CH;
      memcpy(dst[a] + off, src[a] + off, len * sizeof(real));
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
}

// The device part of EZmock_populate_tracer; catalogue left in m->cat.
int populate_device(EZMOCK *ez, const real *params, size_t nexp,
    bool att_part, size_t *ntracer, HostCat *hc = nullptr, bool allow_dsel = true) {
  Conf *conf = CONF(ez);
  Mesh *m = MESH(ez);
  Cosmo *cosmo = COSMO(ez);
  const int Ng = conf->Ng;
  const size_t N = (size_t) Ng * Ng * Ng;
  const real rho_c = params[0], rho_exp = params[1], pbase = params[2],
      sigma_v = params[3];
  if (!isfinite(rho_c) || rho_c < 0) return EZMOCK_ERR_PAR_RCUT;
  if (!isfinite(rho_exp) || rho_exp <= 0) return EZMOCK_ERR_PAR_REXP;
  if (!isfinite(pbase) || pbase <= 0 || pbase >= 1) return EZMOCK_ERR_PAR_PBASE;
  if (!isfinite(sigma_v) || sigma_v < 0) return EZMOCK_ERR_PAR_SIGV;

  std::vector<size_t> num;
  size_t ncell, ntot;
  int e = eval_pdf(nexp, pbase, N, num, ncell, ntot);
  if (e) return e;
  const int nmax = (int) num.size();
  m->ncell0 = ncell;
  m->ntot0 = ntot;

  if ((e = device_ready(m))) return e;
  if (RNG(ez)->type == 1) {
    // the position and velocity draws' windows on their own stream while
    // the bias model runs: at most 3 and 6 per tracer pair, with a margin
    // over the expected count (a larger final count rebuilds them)
    const uint64_t nb = ntot + ntot / 32 + 4096;
    hipStream_t sp = mt_ahead_stream(m);
    if ((e = mt_windows(ez, m, 2, 5 * (uint64_t) N, 3 * nb, sp)) ||
        (e = mt_windows(ez, m, 3, 8 * (uint64_t) N, 6 * ((nb + 1) / 2), sp)))
      return e;
  }
  if (m->psi_ref) {             // the reference reads the caller's arrays
    for (int i = 0; i < 3; i++)
      HIP_OK(hipMemcpyAsync(m->psi[i], m->psi_host[i], N * sizeof(real),
          hipMemcpyHostToDevice, m->st));
    m->mom_valid = false;
  }
  const size_t nblk = (N + EZ_CELLS_PER_BLOCK - 1) / EZ_CELLS_PER_BLOCK;
  if ((e = dalloc(&m->rhot, N)) || (e = dalloc(&m->nt, N)) ||
      (e = dalloc(&m->blockcnt, nblk)) || (e = dalloc(&m->blockoff, nblk)) ||
      (e = dalloc(&m->scan_tmp, scan_tmp_words(nblk) + 8)) ||
      (e = dalloc(&m->hist, EZ_NUM_LOGB_BIN)) || (e = dalloc(&m->counter, 2)) ||
      (e = dalloc(&m->blockhist, nblk * EZ_NUM_LOGB_BIN)) ||
      (e = dalloc(&m->d_num, EZ_MAX_NUM_TRACER_PER_CELL + 1)) ||
      (e = dalloc(&m->d_cnt, EZ_MAX_NUM_TRACER_PER_CELL + 1)) ||
      (e = dalloc(&m->d_vb, EZ_MAX_NUM_TRACER_PER_CELL + 1)) ||
      (e = dalloc(&m->d_res, 4)))
    return e;

  // ---- bias model, pop_tracer.c:149-348
  const real thres = (rho_c == 0) ? (real) EZ_REAL_MIN : rho_c;
  BiasParams bp;
  bp.N = N;
  bp.thres = thres;
  bp.rho_sat = (real) EZ_RHO_SAT;
  bp.iexp = -1 / rho_exp;                 // evaluated in `real`
  const real lambda = (real) EZ_GAUSS_WIDTH;
  bp.lambda = lambda;
  const real rho_max = 1 + lambda * 5;
  bp.imax = ilogb(rho_max);
  mrg_seed_state(RNG(ez)->seed, bp.seed_state);
  const real gs = conf->Lbox / conf->Ng;
  const real igs = conf->Ng / conf->Lbox;
  size_t iv = 0;
  // MT19937: the velocity draws' words on a second stream while the
  // position draws' words and the positions are made (the jump trees are
  // latency-bound: the two overlap)
  const bool vel_side = RNG(ez)->type == 1 && m->fft_streams && !m->profile;
  uint64_t nvel = 0;
  const uint32_t *mtw_vel = nullptr;
  // selection on the device (MRG32k3a, tracers at uniform positions, fields
  // owned by the engine): no host round trip before the end, or before the
  // velocities when the catalogue goes to host arrays
  const bool dsel = allow_dsel && m->devsel && RNG(ez)->type == 0 && !att_part &&
      !m->psi_ref && m->mom_valid;
  if (dsel) m->npath[0]++;
  else m->npath[1]++;
  if (dsel) {
    if (!dsel_capacity(m, N, ntot, true, &e)) return e;
    if ((e = dsel_upload(m, bp, num, ncell, sigma_v)) || (e = dsel_launch(ez, m, !hc)))
      return e;
    if (hc) {
      if ((e = sync_and_collect(m))) return e;
      if ((e = dsel_collect(m)) == EZ_RETRY) {
        m->npath[2]++;
        return populate_device(ez, params, nexp, att_part, ntracer, hc, false);
      }
      if (e) return e;
      ncell = m->ncell;
      ntot = m->ntot;
    }
  }
  else {
  {
    Scope s(m, "bias_count");
    KCHK(launch_count_valid(m->rho, thres, N, m->blockcnt, m->st));
  }
  {
    Scope s(m, "scan");
    KCHK(launch_scan_u32(m->blockcnt, m->blockoff, nblk, m->scan_tmp, m->st));
  }
  uint64_t nvalid = 0;
  const size_t nch = (nblk + EZ_CELLS_PER_BLOCK - 1) / EZ_CELLS_PER_BLOCK;
  HIP_OK(hipMemcpyAsync(&nvalid, m->scan_tmp + nch, sizeof(uint64_t),
      hipMemcpyDeviceToHost, m->st));
  if ((e = sync_and_collect(m))) return e;
  if (nvalid < ncell) return EZMOCK_ERR_RHO_LESS;
  KCHK(launch_zero(m->hist, EZ_NUM_LOGB_BIN * sizeof(unsigned long long), m->st));
  const uint32_t *mtw_bias = nullptr;
  if ((e = mt_words(ez, m, 1, 2 * (uint64_t) N, 2 * ((nvalid + 1) / 2), &mtw_bias)))
    return e;
  {
    Scope s(m, "bias");
    KCHK(launch_bias(m->rho, m->rhot, m->blockoff, m->hist, m->blockhist, bp,
        m->jtab, mtw_bias, m->st));
  }
  HIP_OK(hipMemcpyAsync(m->hist_h, m->hist, sizeof(m->hist_h),
      hipMemcpyDeviceToHost, m->st));
  if ((e = sync_and_collect(m))) return e;

  // ---- PDF mapping, pop_tracer.c:361-514
  int imax;
  size_t cnt = 0;
  for (imax = 0; imax < EZ_NUM_LOGB_BIN; imax++) {
    cnt += m->hist_h[imax];
    if (cnt >= ncell) break;
  }
  if (imax == EZ_NUM_LOGB_BIN) return EZMOCK_ERR_UNKNOWN;
  const real rho_thres = (imax == EZ_NUM_LOGB_BIN - 1) ? (real) EZ_REAL_MIN
      : (real) pow(2, bp.imax - imax);
  if (m->keys_cap < cnt && (e = ensure_keys(m, cnt + cnt / 4 + 4096))) return e;
  KCHK(launch_zero(m->counter, 2 * sizeof(unsigned long long), m->st));
  {
    Scope s(m, "compact");
    KCHK(launch_block_cand(m->blockhist, nblk, imax, m->blockcnt, m->st));
    KCHK(launch_scan_u32(m->blockcnt, m->blockoff, nblk, m->scan_tmp, m->st));
    KCHK(launch_compact_at(m->rhot, rho_thres, N, m->blockoff, m->blockcnt,
        m->keys, m->keys_cap, m->counter, m->st));
  }
  {
    Scope s(m, "sort");
    KCHK(launch_sort(m->keys, m->keys_tmp, cnt, m->sort_gh, m->sort_go,
        m->sort_scan, m->st));
  }
  std::vector<uint64_t> numu(num.begin(), num.end());
  HIP_OK(hipMemcpyAsync(m->d_num, numu.data(), sizeof(uint64_t) * nmax,
      hipMemcpyHostToDevice, m->st));
  {
    Scope s(m, "walk");
    KCHK(launch_walk(m->keys, cnt, m->d_num, nmax, m->d_vb, m->d_cnt, m->d_res,
        m->st));
  }
  uint64_t res[4] = {0, 0, 0, 0}, ncand = 0;
  std::vector<uint64_t> pcnt(nmax);
  HIP_OK(hipMemcpyAsync(res, m->d_res, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, m->st));
  uint64_t cflag = 0;
  HIP_OK(hipMemcpyAsync(&ncand, m->scan_tmp + nch, sizeof(uint64_t), hipMemcpyDeviceToHost, m->st));
  HIP_OK(hipMemcpyAsync(&cflag, m->counter, sizeof(uint64_t), hipMemcpyDeviceToHost, m->st));
  HIP_OK(hipMemcpyAsync(pcnt.data(), m->d_cnt, sizeof(uint64_t) * nmax,
      hipMemcpyDeviceToHost, m->st));
  if ((e = sync_and_collect(m))) return e;
  if (ncand != cnt || cflag || res[2]) return EZMOCK_ERR_UNKNOWN;
  ncell = res[0];
  ntot = res[1];
  m->nmax = nmax;
  m->ncell = ncell;
  m->ntot = ntot;
  m->pdf.assign(pcnt.begin(), pcnt.end());

  // ---- tracers per cell and positions, pop_tracer.c:881-1007
  {
    Scope s(m, "assign_nt");
    KCHK(launch_assign_nt(m->rhot, N, m->d_vb, nmax, m->nt, m->blockcnt, m->st));
  }
  {
    Scope s(m, "scan");
    KCHK(launch_scan_u32(m->blockcnt, m->blockoff, nblk, m->scan_tmp, m->st));
  }
  if ((e = ensure_cat(m, ntot))) return e;
  // tracers attached to particles first (att_part), then the others at
  // uniform positions in their cells, pop_tracer.c:881-1007
  const int16_t *ntpos = m->nt;
  const uint64_t *offpos = m->blockoff;
  if (att_part && (e = attach_particles(ez, m, ntot, &iv, &ntpos, &offpos))) return e;
  nvel = 6 * (((uint64_t) ntot + 1) / 2);
  if (vel_side) {
    HIP_OK(hipEventRecord(m->ev_fork, m->st));
    HIP_OK(hipStreamWaitEvent(m->sx[1], m->ev_fork, 0));
    if ((e = mt_words(ez, m, 3, 8 * (uint64_t) N, nvel, &mtw_vel, 1, m->sx[1]))) return e;
    HIP_OK(hipEventRecord(m->ev_join[1], m->sx[1]));
  }
  const uint32_t *mtw_pos = nullptr;
  if ((e = mt_words(ez, m, 2, 5 * (uint64_t) N, 3 * (uint64_t) (ntot - iv), &mtw_pos)))
    return e;
  if (ntot > iv) {
    Scope s(m, "positions");
    KCHK(launch_positions(ntpos, offpos, Ng, conf->Lbox, gs, bp.seed_state,
        m->jtab, mtw_pos, m->cat[0] + iv, m->cat[1] + iv, m->cat[2] + iv, ntot - iv, m->st));
  }
  }
  m->natt = iv;
  // host catalogue: the positions go to pinned staging on a second stream
  // while the velocities are computed
  real *stage[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (hc) {
    if (m->pin_cap < ntot) {
      if (m->pin) (void) hipHostFree(m->pin);
      m->pin = nullptr;
      m->pin_cap = 0;
      const size_t cap = ntot + ntot / 8 + 1024;
      if (hipHostMalloc((void **) &m->pin, 6 * cap * sizeof(real)) != hipSuccess) {
        m->pin = nullptr;
        return EZMOCK_ERR_MEMORY;
      }
      m->pin_cap = cap;
    }
    for (int i = 0; i < 6; i++) {
      stage[i] = m->pin + (size_t) i * m->pin_cap;
      if (!(hc->h[i] = (real *) malloc((ntot ? ntot : 1) * sizeof(real)))) {
        hc->release();
        return EZMOCK_ERR_MEMORY;
      }
    }
    HIP_OK(hipEventRecord(m->ev_pos, m->st));
    HIP_OK(hipStreamWaitEvent(m->sx[0], m->ev_pos, 0));
    for (int i = 0; i < 3; i++)
      HIP_OK(hipMemcpyAsync(stage[i], m->cat[i], ntot * sizeof(real), hipMemcpyDeviceToHost,
          m->sx[0]));
    HIP_OK(hipEventRecord(m->ev_xyz, m->sx[0]));
  }
  // ---- velocities, pop_tracer.c:593-753 and :1043-1113
  if (!m->mom_valid) {
    // fields given by reference: the momentum fields follow the caller's
    // current arrays, the density stays the one of create_dens_field
    if ((e = paint_cic(m, Ng, igs, nullptr, true)) || (e = sync_and_collect(m)) ||
        (e = cic_finish(m)))
      return e;
    m->mom_valid = !m->psi_ref;
  }
  if (vel_side) HIP_OK(hipStreamWaitEvent(m->st, m->ev_join[1], 0));
  else if ((e = mt_words(ez, m, 3, 8 * (uint64_t) N, nvel, &mtw_vel))) return e;
  if (!dsel || hc) {                    // (launched with the selection otherwise)
    Scope s(m, "velocities");
    KCHK(launch_velocities(m->cat[0], m->cat[1], m->cat[2], m->vel4, ntot, iv, Ng, igs,
        cosmo->vfac, (double) sigma_v,
        bp.seed_state, m->jtab, mtw_vel, m->cat[3], m->cat[4], m->cat[5], m->st,
        dsel ? m->ctl_d : nullptr));
  }
  if (hc) {
    for (int i = 3; i < 6; i++)
      HIP_OK(hipMemcpyAsync(stage[i], m->cat[i], ntot * sizeof(real), hipMemcpyDeviceToHost,
          m->st));
    // positions into the caller's arrays while the velocities finish
    if (hipEventSynchronize(m->ev_xyz) != hipSuccess) { hc->release(); return EZMOCK_ERR_UNKNOWN; }
    copy_arrays(hc->h, stage, 0, 3, ntot, conf->nthread);
  }
  if ((e = sync_and_collect(m))) {
    if (hc) hc->release();
    return e;
  }
  if (hc) copy_arrays(hc->h, stage, 3, 6, ntot, conf->nthread);
  if (dsel && !hc) {
    if ((e = dsel_collect(m)) == EZ_RETRY) {
      m->npath[2]++;
      return populate_device(ez, params, nexp, att_part, ntracer, hc, false);
    }
    if (e) return e;
    ntot = m->ntot;
  }
  m->ntracer = ntot;
  *ntracer = ntot;
  Par *par = PAR(ez);
  par->rho_c = rho_c;
  par->rho_exp = rho_exp;
  par->pdf_base = pbase;
  par->sigma_v = sigma_v;
  return 0;
}

// Validation shared by both populate entry points, pop_tracer.c:1144-1160.
int populate_checks(EZMOCK *ez, const real *params, size_t nexp,
    size_t *ntracer) {
  if (!ez) return EZMOCK_ERR_ARG_EZ;
  if (!params) return EZMOCK_ERR_ARG_PARAMS;
  if (nexp < EZ_MIN_NUM_TRACER) return EZMOCK_ERR_ARG_LOWEXP;
  if (!ntracer) return EZMOCK_ERR_ARG_NDATA;
  Conf *conf = CONF(ez);
  if (nexp >= (size_t) conf->Ng * conf->Ng * conf->Ng) return EZMOCK_ERR_ARG_HIEXP;
  Mesh *m = MESH(ez);
  if (!m->have_field) return EZMOCK_ERR_ARG_MESH;
  if (COSMO(ez)->vfac == HUGE_VAL) return EZMOCK_ERR_ARG_COSMO;
  return 0;
}

// The displacement fields of the initial conditions (delta null) or of a
// white-noise field (delta, host): spectrum, packed kz = 0 / Nyquist slots,
// first FFT pass of the three components, then their last two passes (on
// three streams unless disabled or timed), dens_field.c:337-432 and
// perturb.c:78-399.  seedp (optional): the spectrum's MRG32k3a seed in
// device memory, for a captured sequence.  Launches only, apart from the
// first field of a grid size, which times the FFT schedules.
int field_launch(EZMOCK *ez, Mesh *m, real *delta, bool fixamp, bool iphase,
    const MrgState *seedp) {
  Conf *conf = CONF(ez);
  PkTab *pk = PK(ez);
  const int Ng = conf->Ng;
  const size_t N = (size_t) Ng * Ng * Ng;
  const int Ngh = Ng / 2;
  const size_t plane = (size_t) Ng * Ng;
  int e;
  bool gen_fused = false, mrg_ic = false;
  GenParams gp;
  gp.Ng = Ng;
  gp.Nh = plane * (Ngh + 1);
  gp.kfac = EZ_PI * 2 / conf->Lbox;
  gp.logkfac = log(gp.kfac);
  gp.Pfac = pow(conf->Lbox, -1.5);
  gp.logpk = pk->log;
  gp.fixamp = fixamp;
  gp.iphase = iphase;
  mrg_seed_state(RNG(ez)->seed, gp.seed_state);
  real *spec = nullptr;
  if (m->keep_spec) {
    if ((e = dalloc(&m->spec, 2 * gp.Nh))) return e;
    spec = m->spec;
  }
  cplx *rhok = (cplx *) m->rho;
  if (delta) {
    HIP_OK(hipMemcpyAsync(m->psi[0], delta, N * sizeof(real),
        hipMemcpyHostToDevice, m->st));
    if ((e = r2c_field(m, Ng, m->psi[0]))) return e;
    Scope s(m, "gen_spectrum_wn");
    KCHK(launch_gen_rhok_wn(rhok, m->ny + plane, (const cplx *) m->psi[0],
        m->ny, gp, m->ptab, spec, m->st));
  }
  else {
    const uint32_t *mtw = nullptr;
    if ((e = mt_words(ez, m, 0, 0, (fixamp ? 1 : 2) * (gp.Nh - 1), &mtw)))
      return e;
    // the bias draws' windows (at most N + 1 draws from 2N) on their own
    // stream while the field is made
    if (RNG(ez)->type == 1 && (e = mt_windows(ez, m, 1, 2 * (uint64_t) N, (uint64_t) N + 1,
        mt_ahead_stream(m))))
      return e;
    // MRG32k3a with a tuned, fused first pass: the spectrum is generated by
    // the first FFT pass itself, tile by tile, and only its kz = 0 and
    // Nyquist planes here (they enter the pass through the packed slots)
    bool tuned = false;
    const FftSchedule fs = fft_schedule(Ng, &tuned);
    mrg_ic = !mtw;
    gen_fused = !mtw && !spec && m->gen_fused && fs.split == 0 &&
        (tuned || !m->fft_tune || Ng < 256);
    if (gen_fused) {
      Scope s(m, "gen_planes");
      KCHK(launch_gen_planes(rhok, m->ny + plane, gp, m->ptab, m->jtab, m->st, seedp));
    }
    else {
      if (!mtw && (e = dalloc(&m->gen_base, gen_base_count(gp.Nh)))) return e;
      Scope s(m, "gen_spectrum");
      KCHK(launch_gen_rhok(rhok, m->ny + plane, spec, gp, m->ptab, m->jtab, mtw, m->gen_base,
          m->st, seedp));
    }
  }
  // the generated Nyquist plane lives in ny[plane, 2 plane)
  cplx *ny_save = m->ny;
  m->ny = m->ny + plane;
  if ((e = dalloc(&m->slot0, 3 * plane))) { m->ny = ny_save; return e; }
  {
    Scope s(m, "fft_slot0");
    if (fft_slot0_pack(rhok, m->ny, m->slot0, Ng, m->st) != hipSuccess) {
      m->ny = ny_save;
      return EZMOCK_ERR_UNKNOWN;
    }
  }
  // the first field of a grid size (>= 256) times the FFT schedule
  // candidates on these buffers and keeps the fastest for the process
  bool tuned = false;
  (void) fft_schedule(Ng, &tuned);
  if (m->fft_tune && Ng >= 256 && !tuned &&
      fft_autotune(Ng, (cplx *) m->psi[0], (cplx *) m->psi[1], (cplx *) m->psi[2], rhok,
          m->ny, m->slot0, m->twN, m->st, nullptr, nullptr, mrg_ic ? &gp : nullptr, m->ptab,
          m->jtab) != hipSuccess) {
    m->ny = ny_save;
    return EZMOCK_ERR_UNKNOWN;
  }
  if (gen_fused) {
    Scope s(m, "fft_pass_i_gen");
    if (fft_first_pass3_gen((cplx *) m->psi[0], (cplx *) m->psi[1], (cplx *) m->psi[2],
        m->slot0, m->twN, Ng, gp, m->ptab, m->jtab, seedp, m->st) != hipSuccess) {
      m->ny = ny_save;
      return EZMOCK_ERR_UNKNOWN;
    }
  }
  else {
    Scope s(m, "fft_pass_i");
    if (fft_first_pass3((cplx *) m->psi[0], (cplx *) m->psi[1], (cplx *) m->psi[2],
        rhok, m->ny, m->slot0, m->twN, Ng, m->st) != hipSuccess) {
      m->ny = ny_save;
      return EZMOCK_ERR_UNKNOWN;
    }
  }
  // the three components' remaining passes are independent: components
  // 1 and 2 on their own streams, forked after the first pass and joined
  // before the CIC
  // (one stream while per-kernel timers are on, so that their intervals
  // do not overlap)
  const bool fork = m->fft_streams && !m->profile;
  hipStream_t cs[3] = {m->st, m->st, m->st};
  if (fork) {
    HIP_OK(hipEventRecord(m->ev_fork, m->st));
    for (int i = 0; i < 2; i++) {
      HIP_OK(hipStreamWaitEvent(m->sx[i], m->ev_fork, 0));
      cs[1 + i] = m->sx[i];
    }
  }
  for (int c = 0; c < 3; c++) {
    if ((e = c2r_finish(m, Ng, m->psi[c], cs[c]))) { m->ny = ny_save; return e; }
  }
  m->ny = ny_save;
  if (fork) {
    for (int i = 0; i < 2; i++) {
      HIP_OK(hipEventRecord(m->ev_join[i], m->sx[i]));
      HIP_OK(hipStreamWaitEvent(m->st, m->ev_join[i], 0));
    }
  }
  if (delta) {
    const real norm = 1 / sqrt((double) N);
    Scope s(m, "scale");
    for (int c = 0; c < 3; c++) KCHK(launch_scale(m->psi[c], norm, N, m->st));
  }
  return 0;
}

// What a captured mock depends on besides its seed and parameters (which
// it reads from device memory): buffers that can be reallocated, the P(k)
// table, the velocity factor, the spectrum options and the FFT schedule.
Mesh::GraphKey graph_key(EZMOCK *ez, Mesh *m, bool fixamp, bool iphase) {
  Mesh::GraphKey k;
  k.pk_version = PK(ez) ? PK(ez)->version : -1;
  k.vfac = COSMO(ez)->vfac;
  k.fixamp = fixamp;
  k.iphase = iphase;
  k.keep_spec = m->keep_spec;
  k.fft_streams = m->fft_streams;
  k.keys = m->keys;
  k.keys_cap = m->keys_cap;
  k.cat0 = m->cat[0];
  k.cat_cap = m->cat_cap;
  k.cic_list = m->cic_list;
  k.cic_list_cap = m->cic_list_cap;
  k.fcap = m->cic_fcap;
  k.cic_path = m->cic_path;
  k.cic_fixed_ok = m->cic_fixed_ok;
  k.gen_fused = m->gen_fused;
  bool tuned = false;
  const FftSchedule fs = fft_schedule(CONF(ez)->Ng, &tuned);
  k.sched[0] = fs.tk_j;
  k.sched[1] = fs.tk_i;
  k.sched[2] = fs.split;
  return k;
}

// A mock's launches (the field from the initial conditions with its CIC,
// then the populate with the selection on the device) captured as two HIP
// graphs, right after a mock by the plain calls has sized every buffer.
// Nothing is captured where a replay could not reproduce the plain calls
// without a host round trip: MT19937 streams, timers on, fields by
// reference, CIC by tile scan or exact ranges, FFT schedule not yet tuned.
bool capture_ok(EZMOCK *ez, Mesh *m) {
  const int Ng = CONF(ez)->Ng;
  bool tuned = false;
  (void) fft_schedule(Ng, &tuned);
  return m->use_graph && !m->profile && m->devsel == 1 && RNG(ez)->type == 0 && !m->psi_ref &&
      m->mom_valid && m->cic_path == 0 && m->cic_fixed_ok && m->cic_list &&
      !(m->fft_tune && Ng >= 256 && !tuned) && m->pio && m->ctl_d && m->cic_ovf_h;
}

void capture_mock(EZMOCK *ez, Mesh *m, bool fixamp, bool iphase) {
  drop_graphs(m);
  Conf *conf = CONF(ez);
  const int Ng = conf->Ng;
  if (!capture_ok(ez, m)) return;
  if (dalloc(&m->seed_d, 1)) return;
  hipGraphExec_t x[2] = {nullptr, nullptr};
  bool ok = true;
  for (int i = 0; i < 2 && ok; i++) {
    if (hipStreamBeginCapture(m->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      ok = false;
      break;
    }
    int e = 0;
    if (i == 0) {
      e = field_launch(ez, m, nullptr, fixamp, iphase, m->seed_d);
      if (!e) e = paint_cic(m, Ng, (real) (conf->Ng / conf->Lbox), m->rho, true);
    }
    else e = dsel_launch(ez, m, true);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(m->st, &g);
    if (e || ce != hipSuccess || !g ||
        hipGraphInstantiate(&x[i], g, nullptr, nullptr, 0) != hipSuccess) {
      x[i] = nullptr;
      ok = false;
    }
    if (g) (void) hipGraphDestroy(g);
  }
  m->cic_pend.kind = 0;                 // the capture's painting is not pending
  (void) hipGetLastError();
  if (!ok) {
    for (auto &g : x) if (g) (void) hipGraphExecDestroy(g);
    fprintf(stderr, "EZmock: the mock could not be captured as a HIP graph; "
        "plain launches are used\n");
    m->use_graph = false;
    return;
  }
  m->g_field = x[0];
  m->g_pop = x[1];
  m->gkey = graph_key(ez, m, fixamp, iphase);
  m->npath[4]++;
}

}  // namespace

ApiGuard::ApiGuard() : held(t_api_depth++ == 0) {
  if (held) g_api_mu.lock_shared();
}

ApiGuard::~ApiGuard() {
  --t_api_depth;
  if (held) g_api_mu.unlock_shared();
}

namespace {

// EZmock_ext_mock under the API guard: a replay of the instance's graphs,
// or the plain calls (with *recapture set when they succeeded, for a
// capture after the guard is released).
int mock_call(EZMOCK *ez, const real *params, const size_t nexp, const bool fixamp,
    const bool iphase, size_t *ntracer, int *err, bool *recapture) {
  ApiGuard guard;
  Mesh *m = MESH(ez);
  Conf *conf = CONF(ez);
  const int Ng = conf->Ng;
  const size_t N = (size_t) Ng * Ng * Ng;
  // the plain calls, then a capture for the next mock, unless the graphs
  // of an earlier mock still describe this instance
  auto plain = [&](bool *cap) {
    EZmock_create_dens_field(ez, nullptr, false, nullptr, fixamp, iphase, err);
    EZmock_ext_populate_device(ez, params, nexp, false, ntracer, err);
    *cap = *err == EZMOCK_SUCCESS;
    return *err;
  };
  if (!m->g_field || !m->use_graph || m->profile || m->devsel != 1 || RNG(ez)->type != 0 ||
      !PK(ez) || !(m->gkey == graph_key(ez, m, fixamp, iphase)) ||
      m->ptab_version != PK(ez)->version)
    return plain(recapture);
  // the checks of the two calls, in their order
  const double kmin = 2 * EZ_PI / conf->Lbox;
  const double kmax = EZ_PI * conf->Ng / conf->Lbox * sqrt(3);
  PkTab *pk = PK(ez);
  if (pk->kmin > kmin - EZ_DOUBLE_TOL || pk->kmax < kmax + EZ_DOUBLE_TOL)
    return (*err = EZMOCK_ERR_PK_RANGE);
  int e = populate_checks(ez, params, nexp, ntracer);
  if (e) return (*err = e);
  const real rho_c = params[0], rho_exp = params[1], pbase = params[2], sigma_v = params[3];
  if (!isfinite(rho_c) || rho_c < 0) return (*err = EZMOCK_ERR_PAR_RCUT);
  if (!isfinite(rho_exp) || rho_exp <= 0) return (*err = EZMOCK_ERR_PAR_REXP);
  if (!isfinite(pbase) || pbase <= 0 || pbase >= 1) return (*err = EZMOCK_ERR_PAR_PBASE);
  if (!isfinite(sigma_v) || sigma_v < 0) return (*err = EZMOCK_ERR_PAR_SIGV);
  std::vector<size_t> num;
  size_t ncell, ntot;
  if ((e = eval_pdf(nexp, pbase, N, num, ncell, ntot))) return (*err = e);
  if (!dsel_capacity(m, N, ntot, false, &e)) return plain(recapture);
  m->ncell0 = ncell;
  m->ntot0 = ntot;
  BiasParams bp;
  bp.N = N;
  bp.thres = (rho_c == 0) ? (real) EZ_REAL_MIN : rho_c;
  bp.rho_sat = (real) EZ_RHO_SAT;
  bp.iexp = -1 / rho_exp;
  const real lambda = (real) EZ_GAUSS_WIDTH;
  bp.lambda = lambda;
  const real rho_max = 1 + lambda * 5;
  bp.imax = ilogb(rho_max);
  mrg_seed_state(RNG(ez)->seed, bp.seed_state);
  m->pio->seed = bp.seed_state;
  m->have_field = false;
  m->psi_ref = false;
  for (int i = 0; i < 3; i++) m->psi_host[i] = nullptr;
  HIP_OK(hipMemcpyAsync(m->seed_d, &m->pio->seed, sizeof(MrgState), hipMemcpyHostToDevice,
      m->st));
  if ((e = dsel_upload(m, bp, num, ncell, sigma_v))) return (*err = e);
  m->npath[3]++;
  HIP_OK(hipGraphLaunch(m->g_field, m->st));
  HIP_OK(hipGraphLaunch(m->g_pop, m->st));
  if ((e = sync_and_collect(m))) return (*err = e);
  if (*m->cic_ovf_h) {
    // a CIC list range overflowed: the plain calls repaint with exact
    // ranges and resize the ranges of later fields
    drop_graphs(m);
    return plain(recapture);
  }
  m->fixamp = fixamp;
  m->iphase = iphase;
  m->mom_valid = true;
  m->have_field = true;
  if ((e = dsel_collect(m)) == EZ_RETRY) {
    m->npath[2]++;
    drop_graphs(m);
    e = populate_device(ez, params, nexp, false, ntracer);
    *recapture = !e;
    return (*err = e);
  }
  if (e) return (*err = e);
  m->natt = 0;
  m->ntracer = m->ntot;
  *ntracer = m->ntot;
  Par *par = PAR(ez);
  par->rho_c = rho_c;
  par->rho_exp = rho_exp;
  par->pdf_base = pbase;
  par->sigma_v = sigma_v;
  return EZMOCK_SUCCESS;
}


// capture_mock with every other thread's entry into the C ABI held off;
// skipped when called from inside another entry point of this thread
void capture_exclusive(EZMOCK *ez, Mesh *m, bool fixamp, bool iphase) {
  if (t_api_depth != 0 || !capture_ok(ez, m)) return;
  std::unique_lock<std::shared_mutex> lk(g_api_mu);
  t_api_depth++;
  capture_mock(ez, m, fixamp, iphase);
  t_api_depth--;
}

}  // namespace

// ============================================================== C ABI ===
extern "C" {

// EZmock.c:41-124
EZMOCK *EZmock_init(const double Lbox, const int Ngrid, const int randgen,
    const uint64_t seed, const int nthread, int *err) {
  if (!err || *err != EZMOCK_SUCCESS) return NULL;
  if (Lbox <= 0) { *err = EZMOCK_ERR_ARG_LBOX; return NULL; }
  if (Ngrid <= 0 || Ngrid > EZ_MAX_GRID) { *err = EZMOCK_ERR_ARG_NGRID; return NULL; }
  if (randgen != 0 && randgen != 1) { *err = EZMOCK_ERR_ARG_RNG; return NULL; }
  if (seed == 0) { *err = EZMOCK_ERR_ARG_SEED; return NULL; }
  if (nthread < 0) { *err = EZMOCK_ERR_ARG_THREAD; return NULL; }
  EZMOCK *ez = (EZMOCK *) calloc(1, sizeof(EZMOCK));
  if (!ez) { *err = EZMOCK_ERR_MEMORY; return NULL; }
  Conf *conf = new (std::nothrow) Conf;
  RngConf *rng = new (std::nothrow) RngConf;
  Cosmo *cosmo = new (std::nothrow) Cosmo;
  Par *par = new (std::nothrow) Par;
  Mesh *mesh = new (std::nothrow) Mesh;
  ez->conf = conf; ez->rng = rng; ez->cosmo = cosmo; ez->par = par;
  ez->mesh = mesh; ez->pk = NULL;
  if (!conf || !rng || !cosmo || !par || !mesh) {
    *err = EZMOCK_ERR_MEMORY;
    EZmock_destroy(ez);
    return NULL;
  }
  conf->nthread = nthread ? nthread : (int) sysconf(_SC_NPROCESSORS_ONLN);
  conf->Ng = Ngrid;
  conf->Lbox = Lbox;
  rng->type = randgen;
  rng->seed = seed;
  cosmo->growth2 = cosmo->vfac = HUGE_VAL;
  par->bao = par->rho_c = par->rho_exp = HUGE_VAL;
  par->pdf_base = par->sigma_v = HUGE_VAL;
  return ez;
}

// EZmock.c:133-142
void EZmock_destroy(EZMOCK *ez) {
  ApiGuard guard;
  if (!ez) return;
  Mesh *m = MESH(ez);
  if (m) { free_mesh_device(m); delete m; }
  delete CONF(ez);
  delete RNG(ez);
  delete COSMO(ez);
  delete PK(ez);
  delete PAR(ez);
  free(ez);
}

// errmsg.c:25-70
const char *EZmock_errmsg(const int err) {
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
    "failed to write to the output file", "unknown error"};
  if (err < 0 || err > EZMOCK_ERR_UNKNOWN) return "unregistered error code";
  return msg[err];
}

// cosmology.c:53-95 (growth in flat wCDM) and :121-157 (interface)
int EZmock_set_cosmology(EZMOCK *ez, const double pk_norm, const double fHa,
    const bool eval, const double z, const double z_pk, const double Omega_m,
    const double Omega_nu, const double w, int *err) {
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez) return EZMOCK_ERR_ARG_EZ;      // *err untouched, as upstream
  Cosmo *cosmo = COSMO(ez);
  if (!eval) {
    if (pk_norm <= 0 || fHa <= 0) return (*err = EZMOCK_ERR_PAR_COSMO);
    cosmo->growth2 = pk_norm;
    cosmo->vfac = fHa;
    return EZMOCK_SUCCESS;
  }
  if (z < 0 || z_pk < 0 || Omega_m <= 0 || Omega_m > 1 || Omega_nu < 0 ||
      Omega_nu >= 1 || w >= -1.0 / 3.0) return (*err = EZMOCK_ERR_PAR_COSMO);
  const double a = 1 / (1 + z), a0 = 1 / (1 + z_pk);
  if (a <= 0 || a > 1 || a0 <= 0 || a0 > 1) return (*err = EZMOCK_ERR_PAR_COSMO);
  const double iw = 1 / w;
  const double c1 = 0.5 * (1 - iw);
  const double c2 = -0x1.5555555555555p-2 * iw;
  const double c3 = 1 - 0x1.aaaaaaaaaaaabp-1 * iw;
  const double x = (1 - 1 / Omega_m) * pow(a, -3 * w);
  double h1, h2, D, f;
  if (ez_hyp2f1(c1, c2, c3, x, 1e-12, 10000, &h1) ||
      ez_hyp2f1(c1 + 1, c2 + 1, c3 + 1, x, 1e-12, 10000, &h2))
    return (*err = EZMOCK_ERR_COS_GROWTH);
  f = 1 + 3 * (w - 1) * x / (6 * w - 5) * h2 / h1;
  if (a == a0) D = 1;
  else {
    const double x0 = (1 - 1 / Omega_m) * pow(a0, -3 * w);
    if (ez_hyp2f1(c1, c2, c3, x0, 1e-12, 10000, &h2))
      return (*err = EZMOCK_ERR_COS_GROWTH);
    D = (a * h1) / (a0 * h2);
  }
  const double Omn = Omega_m + Omega_nu;
  const double H = 100 * sqrt((1 - Omn) * pow(a, -3 * (1 + w)) + Omn * pow(a, -3));
  cosmo->growth2 = D * D;
  cosmo->vfac = f * H * a;
  return EZMOCK_SUCCESS;
}

// linear_pk.c:75-183
int EZmock_setup_linear_pk(EZMOCK *ez, const double *k, const int n,
    const double *Pk, const double *Pnw, const double mBAO, const bool logint,
    int *err) {
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez) return (*err = EZMOCK_ERR_ARG_EZ);
  if (!k || !Pk || (mBAO != 0 && !Pnw)) return (*err = EZMOCK_ERR_ARG_NULL);
  if (n <= 3 || n > EZ_MAX_PK_NBIN) return (*err = EZMOCK_ERR_ARG_NBIN);
  Cosmo *cosmo = COSMO(ez);
  if (cosmo->growth2 == HUGE_VAL) return (*err = EZMOCK_ERR_ARG_COSMO);
  for (int i = 0; i < n; i++) {
    if (!isfinite(k[i]) || !isfinite(Pk[i])) return (*err = EZMOCK_ERR_PK_FINITE);
    if (k[i] <= 0 || Pk[i] <= 0) return (*err = EZMOCK_ERR_PK_NONPOS);
    if (i && k[i] <= k[i - 1]) return (*err = EZMOCK_ERR_PK_NONASC);
  }
  if (mBAO != 0) {
    for (int i = 0; i < n; i++) {
      if (!isfinite(Pnw[i])) return (*err = EZMOCK_ERR_PK_FINITE);
      if (Pnw[i] <= 0) return (*err = EZMOCK_ERR_PK_NONPOS);
    }
  }
  Conf *conf = CONF(ez);
  const double kmin = 2 * EZ_PI / conf->Lbox;
  const double kmax = EZ_PI * conf->Ng / conf->Lbox * sqrt(3);
  if (k[0] > kmin - EZ_DOUBLE_TOL || k[n - 1] < kmax + EZ_DOUBLE_TOL)
    return (*err = EZMOCK_ERR_PK_RANGE);
  int lo, hi;
  for (lo = 0; lo < n; lo++) if (k[lo] > kmin) break;
  for (hi = lo; hi < n; hi++) if (k[hi] > kmax) break;
  lo = (lo < 3) ? 0 : lo - 3;
  hi = (hi + 3 >= n) ? n : hi + 3;
  const int nb = hi - lo;
  if (nb < 3) return (*err = EZMOCK_ERR_PK_NBIN);
  PkTab *pk = new (std::nothrow) PkTab;
  if (!pk) return (*err = EZMOCK_ERR_MEMORY);
  pk->k.resize(nb);
  pk->P.resize(nb);
  pk->P2.resize(2 * (size_t) nb);
  for (int i = 0; i < nb; i++) {
    const int j = i + lo;
    if (mBAO == 0) pk->P[i] = Pk[j] * cosmo->growth2;
    else {
      pk->P[i] = (k[j] <= EZ_MAX_K_BAO) ?
          (Pk[j] - Pnw[j]) * exp(k[j] * k[j] * mBAO) + Pnw[j] : Pk[j];
      if (!isfinite(pk->P[i]) || pk->P[i] <= 0) {
        delete pk;
        return (*err = EZMOCK_ERR_PK_MODBAO);
      }
      pk->P[i] *= cosmo->growth2;
    }
  }
  for (int i = 0; i < nb; i++) {
    if (logint) { pk->k[i] = log(k[i + lo]); pk->P[i] = log(pk->P[i]); }
    else pk->k[i] = k[i + lo];
  }
  pk->n = nb;
  pk->log = logint;
  pk->kmin = k[lo];
  pk->kmax = k[hi - 1];
  ez_cspline_ypp(pk->k.data(), pk->P.data(), nb, pk->P2.data());
  pk->version = ++g_pk_version;
  delete PK(ez);
  ez->pk = pk;
  PAR(ez)->bao = mBAO;
  return EZMOCK_SUCCESS;
}

// dens_field.c:337-432
int EZmock_create_dens_field(EZMOCK *ez, real *psi[3], bool deepcopy,
    real *delta, const bool fixamp, const bool iphase, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez) return (*err = EZMOCK_ERR_ARG_EZ);
  Conf *conf = CONF(ez);
  Mesh *m = MESH(ez);
  const int Ng = conf->Ng;
  const size_t N = (size_t) Ng * Ng * Ng;
  int e;
  const bool given = psi && psi[0] && psi[1] && psi[2];
  if (!given) {
    PkTab *pk = PK(ez);
    if (!pk) return (*err = EZMOCK_ERR_ARG_PK);
    const double kmin = 2 * EZ_PI / conf->Lbox;
    const double kmax = EZ_PI * conf->Ng / conf->Lbox * sqrt(3);
    if (pk->kmin > kmin - EZ_DOUBLE_TOL || pk->kmax < kmax + EZ_DOUBLE_TOL)
      return (*err = EZMOCK_ERR_PK_RANGE);
    if (!fft_supported(Ng)) {
      fprintf(stderr, "EZmock: grid size %d is not supported by the GPU FFT "
          "(supported: 2^n and 3*2^n, 8..2048)\n", Ng);
      return (*err = EZMOCK_ERR_ARG_NGRID);
    }
  }
  if ((e = device_ready(m))) return (*err = e);
  m->have_field = false;
  m->mom_valid = false;
  m->psi_ref = false;
  for (int i = 0; i < 3; i++) m->psi_host[i] = nullptr;
  for (int i = 0; i < 3; i++)
    if ((e = dalloc(&m->psi[i], N))) return (*err = e);
  if ((e = dalloc(&m->rho, N))) return (*err = e);
  if (given) {
    for (int i = 0; i < 3; i++)
      HIP_OK(hipMemcpyAsync(m->psi[i], psi[i], N * sizeof(real),
          hipMemcpyHostToDevice, m->st));
    if (!deepcopy) {
      m->psi_ref = true;
      for (int i = 0; i < 3; i++) m->psi_host[i] = psi[i];
    }
  }
  else {
    PkTab *pk = PK(ez);
    const int Ngh = Ng / 2;
    const size_t plane = (size_t) Ng * Ng;
    const int nsq = 3 * Ngh * Ngh + 1;
    double *dk = nullptr, *dP = nullptr, *dP2 = nullptr;
    if ((e = dalloc(&m->ny, 2 * plane)) || (e = dalloc(&m->ptab, (size_t) nsq)) ||
        (e = ensure_twiddles(m, Ng)))
      return (*err = e);
    if (m->ptab_version != pk->version) {
      if ((e = dalloc(&dk, pk->n)) || (e = dalloc(&dP, pk->n)) ||
          (e = dalloc(&dP2, pk->n))) {
        dfree(&dk); dfree(&dP);
        return (*err = e);
      }
      HIP_OK(hipMemcpyAsync(dk, pk->k.data(), sizeof(double) * pk->n, hipMemcpyHostToDevice, m->st));
      HIP_OK(hipMemcpyAsync(dP, pk->P.data(), sizeof(double) * pk->n, hipMemcpyHostToDevice, m->st));
      HIP_OK(hipMemcpyAsync(dP2, pk->P2.data(), sizeof(double) * pk->n, hipMemcpyHostToDevice, m->st));
      const double kfac = EZ_PI * 2 / conf->Lbox;
      {
        Scope s(m, "pk_table");
        KCHK(launch_pk_table(m->ptab, nsq, dk, dP, dP2, pk->n, pk->log, kfac,
            log(kfac), m->st));
      }
      if ((e = sync_and_collect(m))) return (*err = e);
      dfree(&dk); dfree(&dP); dfree(&dP2);
      m->ptab_version = pk->version;
    }
    if ((e = field_launch(ez, m, delta, fixamp, iphase, nullptr))) return (*err = e);
  }
  // density and (for fields owned by the engine) the momentum fields used
  // by every later populate call
  if ((e = paint_cic(m, Ng, (real) (conf->Ng / conf->Lbox), m->rho, !m->psi_ref)))
    return (*err = e);
  m->mom_valid = !m->psi_ref;
  if ((e = sync_and_collect(m)) || (e = cic_finish(m))) return (*err = e);
  m->fixamp = fixamp;
  m->iphase = iphase;
  m->have_field = true;
  return EZMOCK_SUCCESS;
}

// pop_tracer.c:1140-1219
int EZmock_populate_tracer(EZMOCK *ez, const real *params, const size_t nexp,
    const bool att_part, size_t *ntracer, real **x, real **y, real **z,
    real **vx, real **vy, real **vz, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  int e = populate_checks(ez, params, nexp, ntracer);
  if (e) return (*err = e);
  size_t n = 0;
  HostCat hc;
  if ((e = populate_device(ez, params, nexp, att_part, &n, &hc))) return (*err = e);
  real **out[6] = {x, y, z, vx, vy, vz};
  for (int i = 0; i < 6; i++) *out[i] = hc.h[i];
  *ntracer = n;
  return EZMOCK_SUCCESS;
}

// save_res.c:95-178
int EZmock_write_ascii(EZMOCK *ez, const real *x, const real *y, const real *z,
    const real *vx, const real *vy, const real *vz, const size_t ntracer,
    const double rsd_fac, const bool header, const char *fname, int *err) {
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez) return (*err = EZMOCK_ERR_ARG_EZ);
  // the reference also rejects a first element equal to zero (`!(*x)`)
  if (!x || !(*x) || !y || !(*y) || !z || !(*z) || !vx || !(*vx) || !vy ||
      !(*vy) || !vz || !(*vz))
    return (*err = EZMOCK_ERR_ARG_NULL);
  if (!ntracer) return (*err = EZMOCK_ERR_ARG_NDATA);
  Conf *conf = CONF(ez);
  if (rsd_fac > conf->Lbox) return (*err = EZMOCK_ERR_ARG_RSD);
  if (!fname || !(*fname)) return (*err = EZMOCK_ERR_ARG_NULL);
  FILE *fp = fopen(fname, "w");
  if (!fp) return (*err = EZMOCK_ERR_FILE);
  const size_t CH = 1048576;
  std::vector<char> buf(CH);
  size_t used = 0;
  auto line = [&](const char *fmt, ...) -> int {
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(buf.data() + used, CH - used, fmt, ap);
    va_end(ap);
    if (n < 0 || (size_t) n >= CH) return 1;
    if (used + n >= CH) {
      if (fwrite(buf.data(), used, 1, fp) != 1) return 1;
      used = 0;
      va_start(ap, fmt);
      n = vsnprintf(buf.data(), CH, fmt, ap);
      va_end(ap);
      if ((size_t) n >= CH) return 1;
    }
    used += n;
    return 0;
  };
#ifdef SINGLE_PREC
#define RF "%.6g"
#else
#define RF "%.10g"
#endif
  int bad = 0;
  if (header) {
    Mesh *m = MESH(ez);
    Cosmo *c = COSMO(ez);
    Par *p = PAR(ez);
    const char *names[2] = {"MRG32K3A", "MT19937"};
    bad |= line("# BOX_SIZE=%.10lg , NUM_GRID=%d , FIX_AMPLITUDE=%c , "
        "INVERT_PHASE=%c\n", conf->Lbox, conf->Ng, m->fixamp ? 'T' : 'F',
        m->iphase ? 'T' : 'F');
    bad |= line("# GROWTH_PK=%.10lg , VELOCITY_FAC=%.10lg\n", c->growth2, c->vfac);
    bad |= line("# BAO_ENHANCE=%.10lg , RHO_CRITICAL=%.10lg , RHO_EXP=%.10lg , "
        "PDF_BASE=%.10lg , SIGMA_VELOCITY=%.10lg\n# RAND_GENERATOR=%s , "
        "RAND_SEED=%ld\n", p->bao, p->rho_c, p->rho_exp, p->pdf_base,
        p->sigma_v, names[RNG(ez)->type], (long) RNG(ez)->seed);
  }
  if (rsd_fac > 0) {
    bad |= line("# x(1) y(2) z_rsd(3)\n");
    for (size_t i = 0; i < ntracer && !bad; i++) {
      double zr = z[i] + rsd_fac * vz[i];
      if (zr >= conf->Lbox) zr -= conf->Lbox;
      if (zr < 0) zr += conf->Lbox;
      bad |= line(RF " " RF " " RF "\n", x[i], y[i], zr);
    }
  }
  else {
    bad |= line("# x(1) y(2) z(3) vx(4) vy(5) vz(6)\n");
    for (size_t i = 0; i < ntracer && !bad; i++)
      bad |= line(RF " " RF " " RF " " RF " " RF " " RF "\n", x[i], y[i], z[i],
          vx[i], vy[i], vz[i]);
  }
#undef RF
  if (!bad && fwrite(buf.data(), used, 1, fp) != 1) bad = 1;
  fclose(fp);
  return (*err = bad ? EZMOCK_ERR_FILE : EZMOCK_SUCCESS);
}

// ------------------------------------------------------------ extensions
int EZmock_ext_real_size(void) { return (int) sizeof(real); }

int EZmock_ext_device_count(void) {
  ApiGuard guard;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int EZmock_ext_populate_device(EZMOCK *ez, const real *params,
    const size_t nexp, const bool att_part, size_t *ntracer, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  int e = populate_checks(ez, params, nexp, ntracer);
  if (e) return (*err = e);
  if ((e = populate_device(ez, params, nexp, att_part, ntracer))) return (*err = e);
  return EZMOCK_SUCCESS;
}

int EZmock_ext_mock(EZMOCK *ez, const real *params, const size_t nexp, const bool fixamp,
    const bool iphase, size_t *ntracer, int *err) {
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez) return (*err = EZMOCK_ERR_ARG_EZ);
  Mesh *m = MESH(ez);
  bool recapture = false;
  const int r = mock_call(ez, params, nexp, fixamp, iphase, ntracer, err, &recapture);
  if (recapture) capture_exclusive(ez, m, fixamp, iphase);
  return r;
}

int EZmock_ext_copy_tracers(EZMOCK *ez, real *x, real *y, real *z, real *vx,
    real *vy, real *vz, const size_t n, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez) return (*err = EZMOCK_ERR_ARG_EZ);
  Mesh *m = MESH(ez);
  if (!m->cat[0] || n > m->ntracer) return (*err = EZMOCK_ERR_ARG_NDATA);
  real *h[6] = {x, y, z, vx, vy, vz};
  for (int i = 0; i < 6; i++) {
    if (!h[i]) continue;
    if (hipMemcpyAsync(h[i], m->cat[i], n * sizeof(real), hipMemcpyDeviceToHost,
        m->st) != hipSuccess) return (*err = EZMOCK_ERR_UNKNOWN);
  }
  if (hipStreamSynchronize(m->st) != hipSuccess) return (*err = EZMOCK_ERR_UNKNOWN);
  return EZMOCK_SUCCESS;
}

int EZmock_ext_device_tracers(EZMOCK *ez, void *ptrs[6], size_t *ntracer,
    int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez) return (*err = EZMOCK_ERR_ARG_EZ);
  Mesh *m = MESH(ez);
  for (int i = 0; i < 6; i++) ptrs[i] = m->cat[i];
  *ntracer = m->ntracer;
  return EZMOCK_SUCCESS;
}

int EZmock_ext_rsd_z(EZMOCK *ez, const double rsd_fac, double *z_rsd, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez) return (*err = EZMOCK_ERR_ARG_EZ);
  Mesh *m = MESH(ez);
  if (!m->cat[0] || !z_rsd) return (*err = EZMOCK_ERR_ARG_NULL);
  if (rsd_fac > CONF(ez)->Lbox) return (*err = EZMOCK_ERR_ARG_RSD);
  const size_t n = m->ntracer;
  if (m->rsd_cap < n) {
    dfree(&m->d_rsd);
    int e = dalloc(&m->d_rsd, n);
    if (e) return (*err = e);
    m->rsd_cap = n;
  }
  {
    Scope s(m, "rsd");
    if (launch_rsd(m->cat[2], m->cat[5], n, rsd_fac, CONF(ez)->Lbox, m->d_rsd,
        m->st) != hipSuccess) return (*err = EZMOCK_ERR_UNKNOWN);
  }
  if (hipMemcpyAsync(z_rsd, m->d_rsd, n * sizeof(double), hipMemcpyDeviceToHost,
      m->st) != hipSuccess) return (*err = EZMOCK_ERR_UNKNOWN);
  if (sync_and_collect(m)) return (*err = EZMOCK_ERR_UNKNOWN);
  return EZMOCK_SUCCESS;
}

int EZmock_ext_powspec(EZMOCK *ez, const double rsd_fac, const bool interlace,
    const double dk, const int nbin, double *k, double *P0, double *P2,
    double *nmodes, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez) return (*err = EZMOCK_ERR_ARG_EZ);
  Mesh *m = MESH(ez);
  const Conf *conf = CONF(ez);
  if (!m->cat[0] || !k || !P0 || !P2 || !nmodes) return (*err = EZMOCK_ERR_ARG_NULL);
  if (!(dk > 0) || nbin <= 0 || nbin > 1024) return (*err = EZMOCK_ERR_ARG_NBIN);
  if (rsd_fac < 0 || rsd_fac > conf->Lbox) return (*err = EZMOCK_ERR_ARG_RSD);
  const int Ng = conf->Ng;
  const size_t N = (size_t) Ng * Ng * Ng, plane = (size_t) Ng * Ng;
  const size_t ntr = m->ntracer;
  if (ntr == 0) return (*err = EZMOCK_ERR_ARG_NULL);
  const int ng = interlace ? 2 : 1;
  int e;
  for (int g = 0; g < ng; g++)
    if ((e = dalloc(&m->pk_grid[g], N))) return (*err = e);
  if ((e = dalloc(&m->pk_ny, plane)) || (e = dalloc(&m->pk_acc, 4 * (size_t) 1024)) ||
      (e = dalloc(&m->ny, 2 * plane)) || (e = ensure_twiddles(m, Ng)))
    return (*err = e);
  for (int g = 0; g < ng; g++) {
    {
      Scope s(m, "pk_paint");
      if (launch_zero(m->pk_grid[g], N * sizeof(real), m->st) != hipSuccess ||
          launch_pk_paint(m->cat[0], m->cat[1], m->cat[2], m->cat[5], ntr, Ng,
              conf->Lbox, rsd_fac, g ? 0.5 : 0.0, m->pk_grid[g], m->st) != hipSuccess)
        return (*err = EZMOCK_ERR_UNKNOWN);
    }
    if ((e = r2c_field(m, Ng, m->pk_grid[g]))) return (*err = e);
    // the Nyquist plane of grid 0 must survive the r2c of grid 1
    if (g == 0 && interlace &&
        hipMemcpyAsync(m->pk_ny, m->ny, plane * sizeof(cplx), hipMemcpyDeviceToDevice,
            m->st) != hipSuccess)
      return (*err = EZMOCK_ERR_UNKNOWN);
  }
  const double L = conf->Lbox;
  const double kf = 2 * EZ_PI / L;
  // P(k) = V |X(k)|^2 / ntr^2 for the unnormalised transform X of the counts
  const double pnorm = L * L * L / ((double) ntr * (double) ntr);
  {
    Scope s(m, "pk_bin");
    const cplx *X1 = (const cplx *) m->pk_grid[0];
    const cplx *N1 = interlace ? m->pk_ny : m->ny;
    const cplx *X2 = interlace ? (const cplx *) m->pk_grid[1] : nullptr;
    const cplx *N2 = interlace ? m->ny : nullptr;
    if (launch_zero(m->pk_acc, 4 * (size_t) nbin * sizeof(double), m->st) != hipSuccess ||
        launch_pk_bin(X1, N1, X2, N2, Ng, kf, dk, nbin, pnorm, m->pk_acc, m->st) != hipSuccess)
      return (*err = EZMOCK_ERR_UNKNOWN);
  }
  std::vector<double> acc(4 * (size_t) nbin);
  if (hipMemcpyAsync(acc.data(), m->pk_acc, acc.size() * sizeof(double),
      hipMemcpyDeviceToHost, m->st) != hipSuccess) return (*err = EZMOCK_ERR_UNKNOWN);
  if ((e = sync_and_collect(m))) return (*err = e);
  const double shot = L * L * L / (double) ntr;
  for (int b = 0; b < nbin; b++) {
    const double w = acc[4 * b];
    nmodes[b] = w;
    if (w > 0) {
      k[b] = acc[4 * b + 1] / w;
      P0[b] = acc[4 * b + 2] / w - shot;
      P2[b] = 5 * acc[4 * b + 3] / w;
    }
    else k[b] = P0[b] = P2[b] = 0;
  }
  return EZMOCK_SUCCESS;
}

int EZmock_ext_copy_field(EZMOCK *ez, const int which, void *dst, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez || !dst) return (*err = EZMOCK_ERR_ARG_NULL);
  Mesh *m = MESH(ez);
  const int Ng = CONF(ez)->Ng;
  const size_t N = (size_t) Ng * Ng * Ng;
  const void *src = nullptr;
  size_t bytes = 0;
  switch (which) {
    case 0: case 1: case 2: src = m->psi[which]; bytes = N * sizeof(real); break;
    case 3: src = m->rho; bytes = N * sizeof(real); break;
    case 4: src = m->rhot; bytes = N * sizeof(real); break;
    case 5: src = m->nt; bytes = N * sizeof(int16_t); break;
    case 6: src = m->spec; bytes = 2 * (size_t) Ng * Ng * (Ng / 2 + 1) * sizeof(real); break;
    case 7: src = m->hist; bytes = EZ_NUM_LOGB_BIN * sizeof(unsigned long long); break;
    default: return (*err = EZMOCK_ERR_ARG_NULL);
  }
  if (!src) return (*err = EZMOCK_ERR_ARG_MESH);
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
      hipStreamSynchronize(m->st) != hipSuccess)
    return (*err = EZMOCK_ERR_UNKNOWN);
  return EZMOCK_SUCCESS;
}

int EZmock_ext_path_counts(EZMOCK *ez, uint64_t *out) {
  if (!ez || !out) return EZMOCK_ERR_ARG_NULL;
  for (int i = 0; i < 5; i++) out[i] = MESH(ez)->npath[i];
  out[5] = MESH(ez)->last_cnt;
  return EZMOCK_SUCCESS;
}

int EZmock_ext_pdf_info(EZMOCK *ez, size_t *out, const size_t n) {
  if (!ez || !out || n < 5) return EZMOCK_ERR_ARG_NULL;
  Mesh *m = MESH(ez);
  out[0] = m->nmax; out[1] = m->ncell; out[2] = m->ntot;
  out[3] = m->ncell0; out[4] = m->ntot0;
  for (size_t i = 0; i < m->pdf.size() && 5 + i < n; i++) out[5 + i] = m->pdf[i];
  return EZMOCK_SUCCESS;
}

int EZmock_ext_set_option(EZMOCK *ez, const int option, const long value) {
  ApiGuard guard;
  if (!ez) return EZMOCK_ERR_ARG_EZ;
  Mesh *m = MESH(ez);
  if (option == 1) m->keep_spec = value != 0;
  else if (option == 2) m->profile = value != 0;
  else if (option == 3 && value >= 0 && value <= 2) {
    m->cic_path = (int) value;
    m->cic_fixed_ok = true;
  }
  else if (option == 4) m->fft_streams = value != 0;
  else if (option == 6) m->fft_tune = value != 0;
  else if (option == 7 && value >= 0 && value <= 2) m->devsel = (int) value;
  else if (option == 9) m->gen_fused = value != 0;
  else if (option == 8) {
    m->use_graph = value != 0;
    if (!m->use_graph) drop_graphs(m);
  }
  else if (option == 5 && value >= 0) {
    m->cic_list_limit = (size_t) value;
    m->cic_fixed_ok = true;
    m->cic_fcap = 0;
  }
  else return EZMOCK_ERR_ARG_PARAMS;
  return EZMOCK_SUCCESS;
}

int EZmock_ext_kernel_times(EZMOCK *ez, double *ms, long *count, const int n,
    char *names, const size_t cap) {
  if (!ez) return -1;
  Mesh *m = MESH(ez);
  std::string all;
  int i = 0;
  for (auto &kv : m->times) {
    if (i < n) {
      if (ms) ms[i] = kv.second.first;
      if (count) count[i] = kv.second.second;
    }
    all += kv.first;
    all += "\n";
    i++;
  }
  if (names && cap) {
    strncpy(names, all.c_str(), cap - 1);
    names[cap - 1] = 0;
  }
  return i;
}

void EZmock_ext_reset_kernel_times(EZMOCK *ez) {
  if (ez) MESH(ez)->times.clear();
}

int EZmock_ext_info(EZMOCK *ez, double *o) {
  if (!ez || !o) return EZMOCK_ERR_ARG_NULL;
  PkTab *pk = PK(ez);
  o[0] = CONF(ez)->Ng; o[1] = CONF(ez)->Lbox;
  o[2] = COSMO(ez)->growth2; o[3] = COSMO(ez)->vfac;
  o[4] = pk ? pk->n : 0; o[5] = pk ? pk->kmin : 0; o[6] = pk ? pk->kmax : 0;
  o[7] = pk ? pk->log : 0;
  return EZMOCK_SUCCESS;
}

int EZmock_ext_pk_table(EZMOCK *ez, double *k, double *P, double *P2,
    const int n) {
  if (!ez) return EZMOCK_ERR_ARG_EZ;
  PkTab *pk = PK(ez);
  if (!pk) return EZMOCK_ERR_ARG_PK;
  if (n < pk->n || !k || !P || !P2) return EZMOCK_ERR_ARG_NBIN;
  memcpy(k, pk->k.data(), sizeof(double) * pk->n);
  memcpy(P, pk->P.data(), sizeof(double) * pk->n);
  memcpy(P2, pk->P2.data(), sizeof(double) * pk->n);
  return EZMOCK_SUCCESS;
}

int EZmock_ext_set_seed(EZMOCK *ez, const uint64_t seed, int *err) {
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err != EZMOCK_SUCCESS) return *err;
  if (!ez) return (*err = EZMOCK_ERR_ARG_EZ);
  if (seed == 0) return (*err = EZMOCK_ERR_ARG_SEED);
  RNG(ez)->seed = seed;
  return EZMOCK_SUCCESS;
}

int EZmock_ext_sync(EZMOCK *ez, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (!ez) return (*err = EZMOCK_ERR_ARG_EZ);
  Mesh *m = MESH(ez);
  if (m->st && sync_and_collect(m)) return (*err = EZMOCK_ERR_UNKNOWN);
  return EZMOCK_SUCCESS;
}

// ---- stand-alone routines for the parity tests
int EZmock_ext_uniform(const int randgen, const uint64_t seed,
    const uint64_t start, const size_t n, const bool pos, double *out,
    int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err) return *err;
  if (randgen != 0 && randgen != 1) return (*err = EZMOCK_ERR_ARG_RNG);
  Mesh tmp;
  int e = device_ready(&tmp);
  if (e) return (*err = e);
  double *d = nullptr;
  if ((e = dalloc(&d, n))) { free_mesh_device(&tmp); return (*err = e); }
  if (randgen == 0) {           // MRG32k3a: get_double == get_double_pos
    MrgState s;
    mrg_seed_state(seed, s);
    if (launch_uniform(s, start, n, tmp.jtab, d, tmp.st) != hipSuccess) e = EZMOCK_ERR_UNKNOWN;
  }
  else {                        // MT19937: two output words per draw
    uint32_t *w = nullptr, *win = nullptr;
    const size_t nseg = mt_segments(2 * n), ww = mt_window_words(nseg);
    if ((e = dalloc(&w, 2 * n)) == 0 && (e = dalloc(&win, 2 * ww)) == 0) {
      if (mt_stream_windows(seed, 2 * start, nseg, win, win + ww, tmp.st) != hipSuccess ||
          mt_stream_fill(win, 2 * n, w, tmp.st) != hipSuccess ||
          launch_uniform_mt(w, n, pos ? 1 : 0, d, tmp.st) != hipSuccess)
        e = EZMOCK_ERR_RNG_JUMP;
      if (hipStreamSynchronize(tmp.st) != hipSuccess) e = EZMOCK_ERR_UNKNOWN;
    }
    dfree(&w);
    dfree(&win);
  }
  if (!e && (hipMemcpyAsync(out, d, n * sizeof(double), hipMemcpyDeviceToHost, tmp.st) != hipSuccess ||
      hipStreamSynchronize(tmp.st) != hipSuccess))
    e = EZMOCK_ERR_UNKNOWN;
  dfree(&d);
  free_mesh_device(&tmp);
  return (*err = e);
}

static int fft_test(const int Ng, const real *in, real *out, int sign, int *err) {
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err) return *err;
  if (!fft_supported(Ng)) return (*err = EZMOCK_ERR_ARG_NGRID);
  Mesh tmp;
  int e = device_ready(&tmp);
  if (e) return (*err = e);
  const size_t N = (size_t) Ng * Ng * Ng, plane = (size_t) Ng * Ng;
  const int Ngh = Ng / 2;
  const size_t nh = plane * (Ngh + 1);
  real *buf = nullptr, *half = nullptr;
  if ((e = dalloc(&buf, N)) || (e = dalloc(&half, 2 * nh)) ||
      (e = dalloc(&tmp.ny, 2 * plane)) || (e = dalloc(&tmp.rho, N)) ||
      (e = ensure_twiddles(&tmp, Ng))) {
    dfree(&buf); dfree(&half); free_mesh_device(&tmp);
    return (*err = e);
  }
  if (sign > 0) {
    // reference layout in -> generator layout (rhok + Nyquist plane)
    std::vector<real> main(N), nyq(2 * plane);
    for (size_t ij = 0; ij < plane; ij++) {
      for (int k = 0; k < Ngh; k++) {
        main[2 * (ij * Ngh + k)] = in[2 * (ij * (Ngh + 1) + k)];
        main[2 * (ij * Ngh + k) + 1] = in[2 * (ij * (Ngh + 1) + k) + 1];
      }
      nyq[2 * ij] = in[2 * (ij * (Ngh + 1) + Ngh)];
      nyq[2 * ij + 1] = in[2 * (ij * (Ngh + 1) + Ngh) + 1];
    }
    e = 0;
    if (hipMemcpy(tmp.rho, main.data(), N * sizeof(real), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(tmp.ny, nyq.data(), 2 * plane * sizeof(real), hipMemcpyHostToDevice) != hipSuccess)
      e = EZMOCK_ERR_UNKNOWN;
    // comp 1 multiplies by n_y: use a dedicated "identity" component instead
    if (!e) {
      cplx *b = (cplx *) buf;
      if (fft_strided_pass(b, (const cplx *) tmp.rho, tmp.ny, tmp.twN, Ng, 0, +1, 1, 3, tmp.st) != hipSuccess ||
          fft_strided_pass(b, nullptr, nullptr, tmp.twN, Ng, 1, +1, 0, 0, tmp.st) != hipSuccess ||
          fft_row_pass(b, tmp.twN, tmp.twh, Ng, +1, tmp.st) != hipSuccess ||
          hipMemcpyAsync(out, buf, N * sizeof(real), hipMemcpyDeviceToHost, tmp.st) != hipSuccess ||
          hipStreamSynchronize(tmp.st) != hipSuccess)
        e = EZMOCK_ERR_UNKNOWN;
    }
  }
  else {
    e = 0;
    if (hipMemcpy(buf, in, N * sizeof(real), hipMemcpyHostToDevice) != hipSuccess ||
        r2c_field(&tmp, Ng, buf) ||
        hipStreamSynchronize(tmp.st) != hipSuccess)
      e = EZMOCK_ERR_UNKNOWN;
    if (!e) {
      std::vector<real> main(N), nyq(2 * plane);
      if (hipMemcpy(main.data(), buf, N * sizeof(real), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(nyq.data(), tmp.ny, 2 * plane * sizeof(real), hipMemcpyDeviceToHost) != hipSuccess)
        e = EZMOCK_ERR_UNKNOWN;
      for (size_t ij = 0; ij < plane && !e; ij++) {
        for (int k = 0; k < Ngh; k++) {
          out[2 * (ij * (Ngh + 1) + k)] = main[2 * (ij * Ngh + k)];
          out[2 * (ij * (Ngh + 1) + k) + 1] = main[2 * (ij * Ngh + k) + 1];
        }
        out[2 * (ij * (Ngh + 1) + Ngh)] = nyq[2 * ij];
        out[2 * (ij * (Ngh + 1) + Ngh) + 1] = nyq[2 * ij + 1];
      }
    }
  }
  dfree(&buf); dfree(&half);
  free_mesh_device(&tmp);
  return (*err = e);
}

int EZmock_ext_fft_schedule(const int Ng, int *out) {
  if (!out) return EZMOCK_ERR_ARG_NULL;
  if (!fft_supported(Ng)) return EZMOCK_ERR_ARG_NGRID;
  bool tuned = false;
  const FftSchedule sc = fft_schedule(Ng, &tuned);
  out[0] = sc.tk_j; out[1] = sc.tk_i; out[2] = sc.split; out[3] = tuned ? 1 : 0;
  out[4] = sc.tk_g;
  return EZMOCK_SUCCESS;
}

int EZmock_ext_fft_set_schedule(const int Ng, const int tk_j, const int tk_i,
    const int split) {
  ApiGuard guard;
  if (!fft_supported(Ng)) return EZMOCK_ERR_ARG_NGRID;
  auto ok = [](int t) { return t == 0 || t == 4 || t == 8 || t == 16; };
  if (!ok(tk_j) || !ok(tk_i) || (split != 0 && split != 1)) return EZMOCK_ERR_ARG_PARAMS;
  fft_set_schedule(Ng, FftSchedule{tk_j, tk_i, split, 0});
  return EZMOCK_SUCCESS;
}

// Plane waves through the c2r passes at any supported size, checked on the
// device against their closed form (no host copies of the fields).
int EZmock_ext_fft_check(const int Ng, const int path, const int nmodes, const int *ijk,
    const double *amp, double *maxerr, double *maxref, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err) return *err;
  if (!fft_supported(Ng)) return (*err = EZMOCK_ERR_ARG_NGRID);
  if ((path != 0 && path != 1) || nmodes < 1 || nmodes > EZ_FFT_MAX_MODES || !ijk || !amp ||
      !maxerr || !maxref)
    return (*err = EZMOCK_ERR_ARG_NULL);
  const int Ngh = Ng / 2;
  FftModes md;
  md.n = nmodes;
  for (int m = 0; m < nmodes; m++) {
    const int i0 = ijk[3 * m], j0 = ijk[3 * m + 1], k0 = ijk[3 * m + 2];
    if (i0 < 0 || i0 >= Ng || j0 < 0 || j0 >= Ng || k0 <= 0 || k0 >= Ngh)
      return (*err = EZMOCK_ERR_ARG_NULL);
    for (int q = 0; q < m; q++)
      if (md.ijk[q][0] == i0 && md.ijk[q][1] == j0 && md.ijk[q][2] == k0)
        return (*err = EZMOCK_ERR_ARG_NULL);
    md.ijk[m][0] = i0; md.ijk[m][1] = j0; md.ijk[m][2] = k0;
    md.re[m] = amp[2 * m];
    md.im[m] = amp[2 * m + 1];
  }
  Mesh tmp;
  int e = device_ready(&tmp);
  if (e) return (*err = e);
  const size_t N = (size_t) Ng * Ng * Ng, plane = (size_t) Ng * Ng;
  const int nout = path ? 3 : 1;
  real *out[3] = {nullptr, nullptr, nullptr};
  unsigned long long *res = nullptr;
  e = 0;
  for (int c = 0; c < nout && !e; c++) e = dalloc(&out[c], N);
  if (!e && ((e = dalloc(&tmp.rho, N)) || (e = dalloc(&tmp.ny, plane)) ||
      (e = dalloc(&res, 2 * nout)) || (e = ensure_twiddles(&tmp, Ng)) ||
      (path && (e = dalloc(&tmp.slot0, 3 * plane)))))
    ;
  if (!e) {
    cplx *rhok = (cplx *) tmp.rho;
    if (launch_zero(tmp.rho, N * sizeof(real), tmp.st) != hipSuccess ||
        launch_zero(tmp.ny, plane * sizeof(cplx), tmp.st) != hipSuccess ||
        launch_zero(res, 2 * nout * sizeof(unsigned long long), tmp.st) != hipSuccess)
      e = EZMOCK_ERR_UNKNOWN;
    for (int m = 0; m < nmodes && !e; m++) {
      const cplx a = cplx{(real) md.re[m], (real) md.im[m]};
      md.re[m] = a.x;                       // the closed form uses the stored values
      md.im[m] = a.y;
      const size_t at = ((size_t) md.ijk[m][0] * Ng + md.ijk[m][1]) * Ngh + md.ijk[m][2];
      if (hipMemcpyAsync(rhok + at, &a, sizeof(cplx), hipMemcpyHostToDevice, tmp.st) != hipSuccess ||
          hipStreamSynchronize(tmp.st) != hipSuccess)
        e = EZMOCK_ERR_UNKNOWN;
    }
    if (!e && path == 0) {
      cplx *b = (cplx *) out[0];
      if (fft_strided_pass(b, rhok, tmp.ny, tmp.twN, Ng, 0, +1, 1, 3, tmp.st, -1) != hipSuccess ||
          fft_strided_pass(b, nullptr, nullptr, tmp.twN, Ng, 1, +1, 0, 0, tmp.st, -1) != hipSuccess ||
          fft_row_pass(b, tmp.twN, tmp.twh, Ng, +1, tmp.st) != hipSuccess)
        e = EZMOCK_ERR_UNKNOWN;
    }
    else if (!e) {
      if (fft_slot0_pack(rhok, tmp.ny, tmp.slot0, Ng, tmp.st) != hipSuccess ||
          fft_first_pass3((cplx *) out[0], (cplx *) out[1], (cplx *) out[2], rhok, tmp.ny, tmp.slot0,
              tmp.twN, Ng, tmp.st) != hipSuccess)
        e = EZMOCK_ERR_UNKNOWN;
      for (int c = 0; c < 3 && !e; c++) e = c2r_finish(&tmp, Ng, out[c], tmp.st);
    }
    for (int c = 0; c < nout && !e; c++) {
      FftModes mc = md;
      for (int m = 0; m < nmodes; m++) {
        const int n = c == 2 ? md.ijk[m][2] : md.ijk[m][c];
        mc.w[m] = path == 0 ? 1.0 : (double) ((c < 2 && n > Ngh) ? n - Ng : n);
      }
      if (launch_plane_wave_check(out[c], Ng, mc, res + 2 * c, tmp.st) != hipSuccess)
        e = EZMOCK_ERR_UNKNOWN;
    }
    unsigned long long rh[6] = {0};
    if (!e && (hipMemcpyAsync(rh, res, 2 * nout * sizeof(unsigned long long),
        hipMemcpyDeviceToHost, tmp.st) != hipSuccess ||
        hipStreamSynchronize(tmp.st) != hipSuccess))
      e = EZMOCK_ERR_UNKNOWN;
    for (int c = 0; c < nout && !e; c++) {
      memcpy(&maxerr[c], &rh[2 * c], sizeof(double));
      memcpy(&maxref[c], &rh[2 * c + 1], sizeof(double));
    }
  }
  for (int c = 0; c < 3; c++) dfree(&out[c]);
  dfree(&res);
  free_mesh_device(&tmp);
  return (*err = e);
}

int EZmock_ext_fft_c2r(const int Ng, const real *in, real *out, int *err) {
  ApiGuard guard;
  return fft_test(Ng, in, out, +1, err);
}

int EZmock_ext_fft_r2c(const int Ng, const real *in, real *out, int *err) {
  ApiGuard guard;
  return fft_test(Ng, in, out, -1, err);
}

int EZmock_ext_pdf_walk(const real *rhot, const size_t num, const size_t ncell,
    const size_t *pdf, const int nmax, int16_t *nt, size_t *pdf_out,
    size_t *ncell_out, size_t *ntot_out, int *err) {
  ApiGuard guard;
  if (!err) return EZMOCK_ERR_ARG_ECODE;
  if (*err) return *err;
  if (!rhot || !pdf || !nt || nmax < 1 || nmax > EZ_MAX_NUM_TRACER_PER_CELL)
    return (*err = EZMOCK_ERR_ARG_NULL);
  Mesh m;
  int e = device_ready(&m);
  if (e) return (*err = e);
  // histogram on the host (64 bins, pop_tracer.c:265-268 with ilogb_max 5)
  unsigned long long hist[EZ_NUM_LOGB_BIN] = {0};
  for (size_t i = 0; i < num; i++) {
    if (rhot[i] == 0) continue;
    int d = (int) ((unsigned) 5 - (unsigned) ilogb(rhot[i]));
    if (d < 0) d = 0;
    else if (d >= EZ_NUM_LOGB_BIN) d = EZ_NUM_LOGB_BIN - 1;
    hist[d]++;
  }
  int imax;
  size_t cnt = 0;
  for (imax = 0; imax < EZ_NUM_LOGB_BIN; imax++) {
    cnt += hist[imax];
    if (cnt >= ncell) break;
  }
  if (imax == EZ_NUM_LOGB_BIN) { free_mesh_device(&m); return (*err = EZMOCK_ERR_UNKNOWN); }
  const real thres = (imax == EZ_NUM_LOGB_BIN - 1) ? (real) EZ_REAL_MIN : (real) pow(2, 5 - imax);
  const size_t hw = sort_hist_words(cnt + 1);
  real *d_rhot = nullptr;
  if ((e = dalloc(&d_rhot, num)) || (e = dalloc(&m.keys, cnt + 1)) ||
      (e = dalloc(&m.keys_tmp, cnt + 1)) || (e = dalloc(&m.sort_gh, hw)) ||
      (e = dalloc(&m.sort_go, hw)) || (e = dalloc(&m.sort_scan, scan_tmp_words(hw) + 8)) ||
      (e = dalloc(&m.counter, 2)) || (e = dalloc(&m.d_num, nmax)) ||
      (e = dalloc(&m.d_cnt, nmax)) || (e = dalloc(&m.d_vb, nmax)) ||
      (e = dalloc(&m.d_res, 4)) || (e = dalloc(&m.nt, num)) ||
      (e = dalloc(&m.blockcnt, (num + EZ_CELLS_PER_BLOCK - 1) / EZ_CELLS_PER_BLOCK))) {
    dfree(&d_rhot); free_mesh_device(&m);
    return (*err = e);
  }
  std::vector<uint64_t> numu(pdf, pdf + nmax);
  uint64_t res[3] = {0, 0, 0}, ncand = 0;
  std::vector<uint64_t> pcnt(nmax);
  e = 0;
  if (hipMemcpy(d_rhot, rhot, num * sizeof(real), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m.d_num, numu.data(), nmax * sizeof(uint64_t), hipMemcpyHostToDevice) != hipSuccess ||
      launch_zero(m.counter, 16, m.st) != hipSuccess ||
      launch_compact(d_rhot, thres, num, m.keys, cnt + 1, m.counter, m.st) != hipSuccess ||
      launch_sort(m.keys, m.keys_tmp, cnt, m.sort_gh, m.sort_go, m.sort_scan, m.st) != hipSuccess ||
      launch_walk(m.keys, cnt, m.d_num, nmax, m.d_vb, m.d_cnt, m.d_res, m.st) != hipSuccess ||
      launch_assign_nt(d_rhot, num, m.d_vb, nmax, m.nt, m.blockcnt, m.st) != hipSuccess ||
      hipMemcpyAsync(res, m.d_res, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, m.st) != hipSuccess ||
      hipMemcpyAsync(&ncand, m.counter, sizeof(uint64_t), hipMemcpyDeviceToHost, m.st) != hipSuccess ||
      hipMemcpyAsync(pcnt.data(), m.d_cnt, nmax * sizeof(uint64_t), hipMemcpyDeviceToHost, m.st) != hipSuccess ||
      hipMemcpyAsync(nt, m.nt, num * sizeof(int16_t), hipMemcpyDeviceToHost, m.st) != hipSuccess ||
      hipStreamSynchronize(m.st) != hipSuccess)
    e = EZMOCK_ERR_UNKNOWN;
  if (!e && (ncand != cnt || res[2])) e = EZMOCK_ERR_UNKNOWN;
  if (!e) {
    for (int i = 0; i < nmax; i++) pdf_out[i] = pcnt[i];
    *ncell_out = res[0];
    *ntot_out = res[1];
  }
  dfree(&d_rhot);
  free_mesh_device(&m);
  return (*err = e);
}

}  // extern "C"
