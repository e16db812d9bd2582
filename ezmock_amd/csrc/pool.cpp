// pool.cpp -- continuous batching of mock requests on one GPU.
//
// The reference produces one mock per process call (run_mock.c, or the
// Python loop of example/fit/fit_EZmock.ipynb over seeds and bias
// parameters).  A pool runs many such requests on one device at once:
//   * nslot worker threads; each request runs on an engine instance
//     (EZMOCK, with its own HIP stream), so the kernels and host round trips
//     of several mocks overlap on the GPU;
//   * instances stay resident after their request, keyed by grid size and
//     box: a later request of the same shape reuses an idle one (fields and
//     scratch buffers included), so steady traffic allocates and frees no
//     device memory -- hipFree synchronises the whole device, and churning
//     GB-sized allocations between grid sizes stalled every worker;
//   * a shared FIFO of requests; an idle worker takes the first queued
//     request that has an idle instance of its shape, or whose device
//     footprint fits the memory budget next to every resident instance,
//     after evicting least recently used idle instances of other shapes if
//     that makes room;
//   * every request is computed by the same calls as a stand-alone mock
//     (EZmock_set_cosmology, EZmock_setup_linear_pk, EZmock_ext_mock, i.e.
//     EZmock_create_dens_field + EZmock_ext_populate_device, with plain
//     launches unless the pool's option 8 turns on the replay of captured
//     HIP graphs, and optionally EZmock_ext_powspec), so its result does not
//     depend on what else runs.
// Written against the public C API only.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"            // EZmock.h and the error codes
#include "EZmock_ext.h"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
  long id;
  EZMOCK_REQUEST req;
  double t_submit;
};

struct Done {
  EZMOCK_RESULT res;
};

struct Inst {
  EZMOCK *ez = nullptr;         // null while being created
  int Ng = 0;
  double Lbox = 0;
  size_t bytes = 0;             // footprint accounted to the budget
  bool busy = false;
  double last_used = 0;
};

}  // namespace

struct EZMOCK_POOL {
  // model shared by every request
  std::vector<double> k, P;
  double z = 0, z_pk = 0, Om = 0, Onu = 0, w = -1;
  size_t budget = 0;
  std::vector<std::pair<int, long>> options;    // EZmock_ext_set_option of every instance
  // scheduling
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job> queue;
  std::map<long, Done> done;
  std::vector<Inst *> insts;    // resident instances
  size_t held = 0;              // sum of their bytes
  long next_id = 1;
  bool stop = false;
  std::vector<std::thread> threads;
};

size_t EZmock_ext_footprint(const int Ngrid, const int with_powspec) {
  const size_t N = (size_t) Ngrid * Ngrid * Ngrid;
  const size_t s = sizeof(real);
  // psi x3, rho, vel4 (x4), rhot; nt (int16); CIC lists (fixed ranges of
  // up to 6 u32 per cell); candidate keys and sort scratch (<= N keys worst
  // case is covered by the head room of the estimate); catalogue
  size_t b = N * s * 9 + 2 * N + 24 * N + N * sizeof(real) + (64u << 20);
  if (with_powspec) b += 2 * N * s;
  return b;
}

namespace {

int init_inst(EZMOCK_POOL *p, Inst &in, uint64_t seed) {
  int err = 0;
  in.ez = EZmock_init(in.Lbox, in.Ng, 0, seed, 0, &err);
  if (!in.ez) return err ? err : EZMOCK_ERR_UNKNOWN;
  // keep the CIC lists within the footprint's 24 bytes per cell
  EZmock_ext_set_option(in.ez, 5, (long) (24 * (size_t) in.Ng * in.Ng * in.Ng));
  // plain launches: with four instances replaying HIP graphs at once the
  // pool ran a 64 x 256^3 sweep at 708 mocks/s, with plain launches (and
  // the selection on the device) at 863
  EZmock_ext_set_option(in.ez, 8, 0);
  for (auto &o : p->options) EZmock_ext_set_option(in.ez, o.first, o.second);
  EZmock_set_cosmology(in.ez, 0, 0, true, p->z, p->z_pk, p->Om, p->Onu, p->w, &err);
  EZmock_setup_linear_pk(in.ez, p->k.data(), (int) p->k.size(), p->P.data(),
      nullptr, 0, false, &err);
  return err;
}

int run_one(EZMOCK_POOL *p, Inst &in, const Job &job, EZMOCK_RESULT &r) {
  const EZMOCK_REQUEST &q = job.req;
  int err = 0;
  if (!in.ez && (err = init_inst(p, in, q.seed))) return err;
  EZmock_ext_set_seed(in.ez, q.seed, &err);
  size_t ntr = 0;
  EZmock_ext_mock(in.ez, q.params, q.nexp, false, false, &ntr, &err);
  if (err) return err;
  r.ntracer = ntr;
  if (q.pk_nbin > 0) {
    const int nb = q.pk_nbin;
    r.k = (double *) malloc(sizeof(double) * nb);
    r.P0 = (double *) malloc(sizeof(double) * nb);
    r.P2 = (double *) malloc(sizeof(double) * nb);
    r.nmodes = (double *) malloc(sizeof(double) * nb);
    if (!r.k || !r.P0 || !r.P2 || !r.nmodes) return EZMOCK_ERR_MEMORY;
    EZmock_ext_powspec(in.ez, q.rsd_fac, true, q.pk_dk, nb, r.k, r.P0, r.P2,
        r.nmodes, &err);
    if (err) return err;
  }
  if (q.want_catalog && ntr) {
    for (int i = 0; i < 6; i++) {
      r.cat[i] = (real *) malloc(sizeof(real) * ntr);
      if (!r.cat[i]) return EZMOCK_ERR_MEMORY;
    }
    EZmock_ext_copy_tracers(in.ez, r.cat[0], r.cat[1], r.cat[2], r.cat[3],
        r.cat[4], r.cat[5], ntr, &err);
  }
  else EZmock_ext_sync(in.ez, &err);
  return err;
}

// What a worker will do next: run queued job `job` on resident instance
// `inst` (or on a new one, inst = nullptr, after evicting `evict`).
struct Plan {
  long job = -1;
  Inst *inst = nullptr;
  std::vector<Inst *> evict;
};

// The first queued job that can run now (caller holds the lock).
bool pick(EZMOCK_POOL *p, Plan &pl) {
  for (size_t i = 0; i < p->queue.size(); i++) {
    const EZMOCK_REQUEST &q = p->queue[i].req;
    for (Inst *in : p->insts)
      if (!in->busy && in->Ng == q.Ngrid && in->Lbox == q.Lbox) {
        pl.job = (long) i;
        pl.inst = in;
        pl.evict.clear();
        return true;
      }
    const size_t need = EZmock_ext_footprint(q.Ngrid, q.pk_nbin > 0);
    if (p->held + need <= p->budget) {
      pl.job = (long) i;
      pl.inst = nullptr;
      pl.evict.clear();
      return true;
    }
    // idle instances of other shapes, least recently used first
    std::vector<Inst *> idle;
    for (Inst *in : p->insts)
      if (!in->busy) idle.push_back(in);
    std::sort(idle.begin(), idle.end(),
        [](const Inst *a, const Inst *b) { return a->last_used < b->last_used; });
    size_t freed = 0;
    std::vector<Inst *> ev;
    for (Inst *in : idle) {
      if (p->held - freed + need <= p->budget) break;
      ev.push_back(in);
      freed += in->bytes;
    }
    if (p->held - freed + need <= p->budget) {
      pl.job = (long) i;
      pl.inst = nullptr;
      pl.evict = ev;
      return true;
    }
  }
  return false;
}

void drop(EZMOCK_POOL *p, Inst *in) {        // caller holds the lock
  p->insts.erase(std::find(p->insts.begin(), p->insts.end(), in));
  p->held -= in->bytes;
}

void worker(EZMOCK_POOL *p) {
  std::unique_lock<std::mutex> lk(p->mu);
  for (;;) {
    Plan pl;
    p->cv_work.wait(lk, [&] { return p->stop || pick(p, pl); });
    if (p->stop) break;
    Job job = p->queue[pl.job];
    p->queue.erase(p->queue.begin() + pl.job);
    const size_t need = EZmock_ext_footprint(job.req.Ngrid, job.req.pk_nbin > 0);
    for (Inst *in : pl.evict) drop(p, in);
    Inst *in = pl.inst;
    if (!in) {
      in = new Inst;
      in->Ng = job.req.Ngrid;
      in->Lbox = job.req.Lbox;
      in->bytes = need;
      p->insts.push_back(in);
      p->held += need;
    }
    else if (need > in->bytes) {             // now with the estimator grids
      p->held += need - in->bytes;
      in->bytes = need;
    }
    in->busy = true;
    lk.unlock();
    for (Inst *ev : pl.evict) {
      EZmock_destroy(ev->ez);
      delete ev;
    }
    EZMOCK_RESULT r;
    memset(&r, 0, sizeof(r));
    r.id = job.id;
    r.t_submit = job.t_submit;
    r.t_start = now_s();
    r.err = run_one(p, *in, job, r);
    r.t_end = now_s();
    EZMOCK *bad = nullptr;
    lk.lock();
    in->busy = false;
    in->last_used = r.t_end;
    if (r.err) {                     // an instance with a sticky error is dropped
      bad = in->ez;
      drop(p, in);
      delete in;
    }
    p->done[job.id] = Done{r};
    p->cv_done.notify_all();
    p->cv_work.notify_all();         // idle instances and memory may admit others
    if (bad) {
      lk.unlock();
      EZmock_destroy(bad);
      lk.lock();
    }
  }
}

}  // namespace

EZMOCK_POOL *EZmock_ext_pool_init(const int nslot, const size_t mem_budget,
    const double *k, const int nk, const double *Pk, const double z,
    const double z_pk, const double Omega_m, const double Omega_nu,
    const double w, int *err) {
  if (!err) return nullptr;
  if (*err) return nullptr;
  if (nslot < 1 || nslot > 64) { *err = EZMOCK_ERR_ARG_THREAD; return nullptr; }
  if (!k || !Pk || nk < 4) { *err = EZMOCK_ERR_ARG_PK; return nullptr; }
  EZMOCK_POOL *p = new EZMOCK_POOL();
  p->k.assign(k, k + nk);
  p->P.assign(Pk, Pk + nk);
  p->z = z; p->z_pk = z_pk; p->Om = Omega_m; p->Onu = Omega_nu; p->w = w;
  p->budget = mem_budget;
  for (int i = 0; i < nslot; i++) p->threads.emplace_back(worker, p);
  return p;
}

int EZmock_ext_pool_option(EZMOCK_POOL *p, const int option, const long value) {
  if (!p) return EZMOCK_ERR_ARG_NULL;
  std::lock_guard<std::mutex> lk(p->mu);
  if (!p->insts.empty() || p->next_id != 1) return EZMOCK_ERR_ARG_PARAMS;
  p->options.emplace_back(option, value);
  return EZMOCK_SUCCESS;
}

long EZmock_ext_pool_submit(EZMOCK_POOL *p, const EZMOCK_REQUEST *req, int *err) {
  if (!err) return -1;
  if (*err) return -1;
  if (!p || !req) { *err = EZMOCK_ERR_ARG_NULL; return -1; }
  if (req->seed == 0) { *err = EZMOCK_ERR_ARG_SEED; return -1; }
  if (EZmock_ext_footprint(req->Ngrid, req->pk_nbin > 0) > p->budget) {
    *err = EZMOCK_ERR_MEMORY;
    return -1;
  }
  std::lock_guard<std::mutex> lk(p->mu);
  Job j;
  j.id = p->next_id++;
  j.req = *req;
  j.t_submit = now_s();
  p->queue.push_back(j);
  p->cv_work.notify_all();
  return j.id;
}

int EZmock_ext_pool_wait(EZMOCK_POOL *p, const long id, EZMOCK_RESULT *res) {
  if (!p || !res) return EZMOCK_ERR_ARG_NULL;
  std::unique_lock<std::mutex> lk(p->mu);
  p->cv_done.wait(lk, [&] { return p->done.count(id) > 0; });
  *res = p->done[id].res;
  p->done.erase(id);
  return res->err;
}

void EZmock_ext_result_free(EZMOCK_RESULT *res) {
  if (!res) return;
  free(res->k); free(res->P0); free(res->P2); free(res->nmodes);
  for (int i = 0; i < 6; i++) free(res->cat[i]);
  memset(res, 0, sizeof(*res));
}

void EZmock_ext_pool_destroy(EZMOCK_POOL *p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    p->stop = true;
    p->cv_work.notify_all();
  }
  for (auto &t : p->threads) t.join();
  for (Inst *in : p->insts) {
    if (in->ez) EZmock_destroy(in->ez);
    delete in;
  }
  for (auto &d : p->done) EZmock_ext_result_free(&d.second.res);
  delete p;
}
