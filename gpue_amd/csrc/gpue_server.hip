// gpue_server.hip -- continuous-batching solve service on one GPU.
//
// Requests (ground states, quenches) are queued by gpue_server_submit from any
// thread.  A worker thread admits them in arrival order whenever a slot and
// enough HBM (by the request's device footprint) are free, keeps every psi
// resident on the device, and drives each admitted request on its own context
// and HIP stream in chunks of `check_every` steps followed by an asynchronous
// energy evaluation.  Chunks of different requests are in flight at the same
// time; a request's next chunk is issued as soon as its previous one has
// completed (event polling, no global tick), so small and large problems
// progress independently.  A request's arithmetic does not depend on what else
// is running: its result is the same alone and in the mix.
//
// No reference counterpart: GPUE runs one problem per process (bin/run.sh
// starts several processes).
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "ctx_internal.h"

using namespace gpue;

namespace {

// counter-based uniform in [-1, 1) (splitmix64 finaliser of seed and index)
__device__ __forceinline__ double unif(unsigned long long seed, unsigned long long i) {
  unsigned long long z = seed * 0x9E3779B97F4A7C15ull + i + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// psi0 = exp(-r^2/2) (1 + 0.01 u) [* exp(i pi (u'+1))], on the context's grids
template <typename T>
__global__ void k_init_state(cx<T>* psi, const double* xg, const double* yg, const double* zg, int nx, int ny, int nz,
                             unsigned long long seed, int random_phase) {
  const long N = (long)nx * ny * nz;
  for (long gid = blockIdx.x * (long)blockDim.x + threadIdx.x; gid < N; gid += (long)gridDim.x * blockDim.x) {
    const int xi = (int)(gid % nx), yi = (int)((gid / nx) % ny), zi = (int)(gid / ((long)nx * ny));
    const double r2 = xg[xi] * xg[xi] + yg[yi] * yg[yi] + zg[zi] * zg[zi];
    const double a = exp(-r2 / 2) * (1.0 + 0.01 * unif(seed, 2 * (unsigned long long)gid));
    double re = a, im = 0.0;
    if (random_phase) {
      double s, c;
      sincospi(unif(seed, 2 * (unsigned long long)gid + 1) + 1.0, &s, &c);
      re = a * c;
      im = a * s;
    }
    psi[gid] = mk((T)re, (T)im);
  }
}

struct Req {
  gpue_request r;
  unsigned long long id;
  double t_submit;
};

struct Slot {
  Req req;
  gpue_ctx* ctx = nullptr;
  double* host = nullptr;          // pinned: energy terms read back after a chunk
  hipEvent_t ev = nullptr;
  hipEvent_t tev[2] = {};          // timing events: first GPU work, end of the latest chunk
  int cls = 0;                     // size class: the stream the request runs on
  bool busy = false;
  int gstate = 0;
  int step = 0;                    // step index within the current phase
  int imag_done = 0, real_done = 0;
  double e_prev = 0.0, E = 0.0, mu = 0.0;
  double bytes = 0.0;
  double t_start = 0.0;
  int status = GPUE_OK;
  bool finished = false;
};

}  // namespace

// A finished request's context, pinned buffer and event, kept for the next
// request of the same shape (its device memory stays counted in the budget).
struct Pooled {
  gpue_ctx* ctx = nullptr;
  double* host = nullptr;
  hipEvent_t ev = nullptr;
  hipEvent_t tev[2] = {};
  double bytes = 0.0;
  int cls = 0;
  long last_use = 0;
};

// Requests run on at most NSTREAMS streams, one per size class, since the device
// gives a process four hardware queues (more streams would share them in FIFO
// order anyway): large states keep one queue to themselves and small ones are
// never queued behind them.
constexpr int NSTREAMS = 4;
static int size_class(const gpue_params& p) {
  const double n = (double)p.xDim * (p.dimnum > 1 ? p.yDim : 1) * (p.dimnum > 2 ? p.zDim : 1);
  const double bytes = n * (p.precision == 32 ? 8.0 : 16.0);
  if (bytes >= 64.0 * 1024 * 1024) return 0;   // 256^3-class states
  if (bytes >= 8.0 * 1024 * 1024) return 1;    // 1024^2, 128^3-192^3
  if (bytes >= 1.0 * 1024 * 1024) return 2;    // 256^2-512^2, 64^3
  return 3;
}

struct gpue_server {
  int device = 0;
  double budget = 0.0, used = 0.0;
  int max_slots = 0;
  hipStream_t streams[NSTREAMS] = {};
  hipEvent_t ev0 = nullptr;        // timing event at creation: the origin of the device clock
  std::vector<Pooled> pool;
  long pool_clock = 0;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req> queue;
  std::vector<gpue_result> done;
  std::vector<Slot*> active;
  std::vector<Slot*> retired;      // finished, device memory not yet released
  std::thread worker;
  std::atomic<bool> stop{false};
  std::atomic<int> inflight{0};    // queued + active
  double clock() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

static double footprint(const gpue_params& p) {
  const double n = (double)p.xDim * (p.dimnum > 1 ? p.yDim : 1) * (p.dimnum > 2 ? p.zDim : 1);
  const double s = p.precision == 32 ? 8.0 : 16.0;
  return 2.0 * n * s + 4.0 * 1024 * 1024;   // psi, scratch psi, tables and partials
}

// Queue a chunk: nsteps of the current phase, then the energy terms into pinned memory.
static gpue_status issue_chunk(Slot* s, int nsteps) {
  gpue_ctx* c = s->ctx;
  gpue_status st = evolve_async(c, s->step, nsteps, s->gstate);
  if (st != GPUE_OK) return st;
  s->step += nsteps;
  if (s->gstate == 0) s->imag_done += nsteps;
  else s->real_done += nsteps;
  st = energy_launch(c, s->host);
  if (st != GPUE_OK) return st;
  if (hipEventRecord(s->ev, c->stream) != hipSuccess || hipEventRecord(s->tev[1], c->stream) != hipSuccess)
    return GPUE_EHIP;
  s->busy = true;
  return GPUE_OK;
}

// Free pooled contexts, least recently used first, until `need` more bytes fit
// the budget (hipFree waits for the whole device, so this happens only when the
// budget demands it).
static void shrink_pool(gpue_server* sv, double need) {
  while (!sv->pool.empty() && sv->used + need > sv->budget) {
    size_t k = 0;
    for (size_t j = 1; j < sv->pool.size(); ++j)
      if (sv->pool[j].last_use < sv->pool[k].last_use) k = j;
    Pooled& e = sv->pool[k];
    if (e.ev) (void)hipEventDestroy(e.ev);
    for (hipEvent_t t : e.tev)
      if (t) (void)hipEventDestroy(t);
    if (e.host) (void)hipHostFree(e.host);
    gpue_destroy(e.ctx);
    sv->used -= e.bytes;
    sv->pool.erase(sv->pool.begin() + k);
  }
}

// Admit a request: `reuse` is a pooled context of its shape taken out of the
// pool at admission (ctx null: none), else a new context is created.
static gpue_status admit(gpue_server* sv, const Req& q, const Pooled& reuse, Slot** out) {
  Slot* s = new Slot();
  s->req = q;
  s->bytes = footprint(q.r.params);
  gpue_params p = q.r.params;
  p.device = sv->device;
  s->cls = size_class(p);
  gpue_status st = GPUE_OK;
  if (reuse.ctx) {   // no device allocation
    s->ctx = reuse.ctx;
    s->host = reuse.host;
    s->ev = reuse.ev;
    s->tev[0] = reuse.tev[0];
    s->tev[1] = reuse.tev[1];
    st = ctx_reset(s->ctx, &p);
  } else {
    st = ctx_create(&p, sv->streams[s->cls], &s->ctx);
    if (st == GPUE_OK && (hipHostMalloc((void**)&s->host, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
                          hipEventCreateWithFlags(&s->ev, hipEventDisableTiming) != hipSuccess ||
                          hipEventCreate(&s->tev[0]) != hipSuccess || hipEventCreate(&s->tev[1]) != hipSuccess))
      st = GPUE_EHIP;
  }
  if (st == GPUE_OK && hipEventRecord(s->tev[0], s->ctx->stream) != hipSuccess) st = GPUE_EHIP;
  if (st == GPUE_OK) {
    gpue_ctx* c = s->ctx;
    const long N = c->N;
    int blocks = (int)std::min<long>((N + 255) / 256, 8192);
    if (c->prec == 64)
      hipLaunchKernelGGL(k_init_state<double>, dim3(blocks), dim3(256), 0, c->stream, (cx<double>*)c->psi, c->dgrid[0],
                         c->dgrid[1], c->dgrid[2], c->dims[0], c->dims[1], c->dims[2], q.r.seed, q.r.random_phase);
    else
      hipLaunchKernelGGL(k_init_state<float>, dim3(blocks), dim3(256), 0, c->stream, (cx<float>*)c->psi, c->dgrid[0],
                         c->dgrid[1], c->dgrid[2], c->dims[0], c->dims[1], c->dims[2], q.r.seed, q.r.random_phase);
    if (hipGetLastError() != hipSuccess) st = GPUE_EHIP;
    if (st == GPUE_OK) st = gpue_parsum(c);
    if (st == GPUE_OK && hipEventRecord(s->tev[1], c->stream) != hipSuccess) st = GPUE_EHIP;
  }
  s->t_start = sv->clock();
  s->status = st;
  *out = s;
  return st;
}

// Record the result.  The slot's device memory is released later, in
// release_retired: hipFree waits for the whole device, which would stall the
// other requests' streams here.
static void retire(gpue_server* sv, Slot* s) {
  gpue_result r;
  std::memset(&r, 0, sizeof(r));
  r.id = s->req.id;
  r.status = s->status;
  r.imag_steps = s->imag_done;
  r.real_steps = s->real_done;
  r.energy = s->E;
  r.mu = s->mu;
  r.t_submit = s->req.t_submit;
  r.t_start = s->t_start;
  if (s->ctx && s->req.r.psi_out && s->status == GPUE_OK) {
    // waits for this request's copy, not for the later work of the stream's other requests
    const gpue_status st = download_wfc_ev(s->ctx, s->req.r.psi_out, s->ev);
    if (st != GPUE_OK) r.status = st;
  }
  r.t_done = sv->clock();
  if (s->status == GPUE_OK && s->tev[0] && s->tev[1] && sv->ev0) {   // both recorded and complete by now
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, sv->ev0, s->tev[0]) == hipSuccess &&
        hipEventElapsedTime(&b, sv->ev0, s->tev[1]) == hipSuccess) {
      r.t_dev_start = 1e-3 * a;
      r.t_dev_done = 1e-3 * b;
    }
  }
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->done.push_back(r);
  }
  sv->retired.push_back(s);
  sv->inflight--;
}

// Finished requests' contexts go to the pool (their memory stays counted).
static void release_retired(gpue_server* sv) {
  for (Slot* s : sv->retired) {
    if (s->ctx && s->status == GPUE_OK) {
      Pooled e;
      e.ctx = s->ctx;
      e.host = s->host;
      e.ev = s->ev;
      e.tev[0] = s->tev[0];
      e.tev[1] = s->tev[1];
      e.bytes = s->bytes;
      e.cls = s->cls;
      e.last_use = ++sv->pool_clock;
      sv->pool.push_back(e);
    } else {   // a failed request's context is not reused
      if (s->ev) (void)hipEventDestroy(s->ev);
      for (hipEvent_t t : s->tev)
        if (t) (void)hipEventDestroy(t);
      if (s->host) (void)hipHostFree(s->host);
      if (s->ctx) gpue_destroy(s->ctx);
      sv->used -= s->bytes;
    }
    delete s;
  }
  sv->retired.clear();
}

// After a chunk: convergence / phase change / completion, then the next chunk.
static void advance(gpue_server* sv, Slot* s) {
  const gpue_request& r = s->req.r;
  gpue_ctx* c = s->ctx;
  double E = 0, M = 0;
  energy_combine(c, s->host, &E, &M);
  if (!std::isfinite(E)) {
    s->status = GPUE_ENONFINITE;
    s->finished = true;
    return;
  }
  const int every = r.check_every > 0 ? r.check_every : 100;
  if (s->gstate == 0) {
    const bool conv = s->imag_done > every && std::fabs(E - s->e_prev) < r.energy_tol * std::fabs(s->e_prev);
    s->e_prev = E;
    s->E = E;
    s->mu = M;
    if (conv || s->imag_done >= r.max_imag_steps) {
      if (r.real_steps > 0) {   // quench: new interaction, real time from the ground state
        gpue_set_g(c, c->p.g * r.quench_factor);
        s->gstate = 1;
        s->step = 0;
      } else {
        s->finished = true;
        return;
      }
    }
  } else {
    s->E = E;
    s->mu = M;
    if (s->real_done >= r.real_steps) {
      s->finished = true;
      return;
    }
  }
  const int left = s->gstate == 0 ? r.max_imag_steps - s->imag_done : r.real_steps - s->real_done;
  if (issue_chunk(s, std::max(1, std::min(every, left))) != GPUE_OK) {
    s->status = GPUE_EHIP;
    s->finished = true;
  }
}

static void worker_main(gpue_server* sv) {
  (void)hipSetDevice(sv->device);
  for (;;) {
    // admissions, in arrival order, while a slot and the HBM budget allow; a
    // pooled context of the request's shape is reused, other pooled contexts
    // are freed only when the budget needs their memory
    std::vector<Req> admitted;
    std::vector<Pooled> reuse;
    double shrink = 0.0;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      if (sv->active.empty() && sv->queue.empty()) {
        if (sv->stop) break;
        sv->cv.wait_for(lk, std::chrono::milliseconds(5));
        continue;
      }
      while (!sv->queue.empty() && (int)(sv->active.size() + admitted.size()) < sv->max_slots) {
        gpue_params qp = sv->queue.front().r.params;
        qp.device = sv->device;
        Pooled e;
        for (size_t k = 0; k < sv->pool.size(); ++k)
          if (ctx_same_shape(sv->pool[k].ctx, &qp)) {
            e = sv->pool[k];
            sv->pool.erase(sv->pool.begin() + k);
            break;
          }
        const double extra = e.ctx ? 0.0 : footprint(qp);   // a pooled context is already counted
        if (sv->used + extra > sv->budget) {
          if (e.ctx) sv->pool.push_back(e);
          shrink = extra;
          break;
        }
        sv->used += extra;
        admitted.push_back(sv->queue.front());
        reuse.push_back(e);
        sv->queue.pop_front();
      }
    }
    if (shrink > 0.0 && admitted.empty()) {
      const size_t before = sv->pool.size();
      shrink_pool(sv, shrink);
      if (sv->pool.size() != before) continue;   // room made: admit again
    }
    for (size_t k = 0; k < admitted.size(); ++k) {
      Slot* s = nullptr;
      admit(sv, admitted[k], reuse[k], &s);
      const gpue_request& r = s->req.r;
      const int every = r.check_every > 0 ? r.check_every : 100;
      if (s->status == GPUE_OK) {
        if (r.max_imag_steps > 0) {
          s->status = issue_chunk(s, std::min(every, r.max_imag_steps));
        } else if (r.real_steps > 0) {   // straight to real time
          gpue_set_g(s->ctx, s->ctx->p.g * r.quench_factor);
          s->gstate = 1;
          s->status = issue_chunk(s, std::min(every, r.real_steps));
        } else {
          s->finished = true;
        }
      }
      if (s->status != GPUE_OK) s->finished = true;
      sv->active.push_back(s);
    }
    bool progressed = !admitted.empty();
    for (Slot* s : sv->active) {
      if (s->finished || !s->busy) continue;
      hipError_t q = hipEventQuery(s->ev);
      if (q == hipErrorNotReady) continue;
      s->busy = false;
      progressed = true;
      if (q != hipSuccess) {
        s->status = GPUE_EHIP;
        s->finished = true;
        continue;
      }
      advance(sv, s);
    }
    for (size_t k = 0; k < sv->active.size();) {
      Slot* s = sv->active[k];
      if (s->finished && !s->busy) {
        retire(sv, s);
        sv->active.erase(sv->active.begin() + k);
        progressed = true;
      } else {
        ++k;
      }
    }
    release_retired(sv);
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  release_retired(sv);
}

extern "C" {

gpue_status gpue_server_create(int device, double budget, int max_slots, gpue_server** out) {
  if (!out || budget <= 0 || max_slots <= 0) return GPUE_EINVAL;
  gpue_server* sv = new gpue_server();
  sv->device = device;
  sv->budget = budget;
  sv->max_slots = max_slots;
  sv->t0 = std::chrono::steady_clock::now();
  if (hipSetDevice(device) != hipSuccess) {
    delete sv;
    return GPUE_EHIP;
  }
  for (int k = 0; k < NSTREAMS; ++k)
    if (hipStreamCreateWithFlags(&sv->streams[k], hipStreamNonBlocking) != hipSuccess) {
      for (int j = 0; j < k; ++j) (void)hipStreamDestroy(sv->streams[j]);
      delete sv;
      return GPUE_EHIP;
    }
  if (hipEventCreate(&sv->ev0) != hipSuccess || hipEventRecord(sv->ev0, sv->streams[0]) != hipSuccess ||
      hipEventSynchronize(sv->ev0) != hipSuccess) {
    for (int k = 0; k < NSTREAMS; ++k) (void)hipStreamDestroy(sv->streams[k]);
    delete sv;
    return GPUE_EHIP;
  }
  sv->worker = std::thread(worker_main, sv);
  *out = sv;
  return GPUE_OK;
}

gpue_status gpue_server_submit(gpue_server* sv, const gpue_request* r, unsigned long long id) {
  if (!sv || !r) return GPUE_EINVAL;
  if (footprint(r->params) > sv->budget) return GPUE_EBUDGET;
  Req q;
  q.r = *r;
  q.id = id;
  q.t_submit = sv->clock();
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->queue.push_back(q);
  }
  sv->inflight++;
  sv->cv.notify_one();
  return GPUE_OK;
}

int gpue_server_poll(gpue_server* sv, gpue_result* out, int max) {
  if (!sv || !out || max <= 0) return 0;
  std::lock_guard<std::mutex> lk(sv->mu);
  const int n = std::min<int>(max, (int)sv->done.size());
  for (int i = 0; i < n; ++i) out[i] = sv->done[i];
  sv->done.erase(sv->done.begin(), sv->done.begin() + n);
  return n;
}

gpue_status gpue_server_drain(gpue_server* sv, double timeout_s) {
  if (!sv) return GPUE_EINVAL;
  const double t_end = sv->clock() + timeout_s;
  while (sv->inflight.load() > 0) {
    if (sv->clock() > t_end) return GPUE_ETIMEOUT;
    std::this_thread::sleep_for(std::chrono::microseconds(200));
  }
  return GPUE_OK;
}

double gpue_server_clock(const gpue_server* sv) { return sv ? sv->clock() : 0.0; }

void gpue_server_destroy(gpue_server* sv) {
  if (!sv) return;
  sv->stop = true;
  sv->cv.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  shrink_pool(sv, 2.0 * sv->budget + 1.0);   // every pooled context
  for (int k = 0; k < NSTREAMS; ++k)
    if (sv->streams[k]) (void)hipStreamDestroy(sv->streams[k]);
  if (sv->ev0) (void)hipEventDestroy(sv->ev0);
  delete sv;
}

}  // extern "C"
