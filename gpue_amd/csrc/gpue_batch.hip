// gpue_batch.hip -- many independent problems advanced together.
//
// Instances with the same shape (precision, dimensions, grid sizes, kernel
// variants, gauge/no gauge) form a group.  For every sweep of the step one
// batched kernel covers every instance of the group: blocks [i*bpi, (i+1)*bpi)
// work on instance i, whose SweepArgs live in a device array.  GPUE's gauge
// order alternates with the step parity (src/evolution.cu:837,842).  When the
// odd-step sweep list is a prefix of the even-step one in kernel kind (the 2D
// and rotation-about-z cases: [FWD][K][INV(+A_x)][A_y(+norm)] and
// [FWD][K][INV][A_y][A_x+norm]) both parities share one launch per position,
// with per-instance flags, and the extra position is inactive for odd
// instances; otherwise each parity has its own sequence.  Every imaginary-time
// sequence ends with one batched norm finalize.  The groups run on at most
// four streams (the process's hardware queues; the groups with the most work
// get one each, the others share the last), with two consecutive steps of a
// group captured once per gstate as a HIP graph and replayed, so the groups of
// different shapes (the small ones latency bound) overlap on the device.
// Device memory is allocated only in gpue_batch_create.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "ctx_internal.h"

using namespace gpue;

namespace {

struct BSeq {                       // one launch sequence of a group for (gstate, phase)
  std::vector<Sweep> keys;          // kernel kind per position (flags come per instance)
  int mask = 3;                     // step parities it serves (bit p: parity p)
  std::vector<SweepArgs*> dargs;    // per position: device array [members]
  std::vector<std::vector<SweepArgs>> hargs;
  FinArgs* dfin = nullptr;          // imaginary time: device array [members]
  std::vector<FinArgs> hfin;
  bool any = true;                  // some member takes this sequence (else it is not launched)
};

struct Group {
  std::vector<int> members;
  gpue_ctx* proto;                  // first member: shape and variants
  std::vector<Sweep> sched[2][2];   // [gstate][parity]
  std::vector<BSeq> seqs[2][2];     // [gstate][phase]
  hipStream_t stream = nullptr;     // one of the batch's streams (not owned)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
};

bool same_shape(const gpue_ctx* a, const gpue_ctx* b) {
  return a->prec == b->prec && a->p.dimnum == b->p.dimnum && a->dims[0] == b->dims[0] && a->dims[1] == b->dims[1] &&
         a->dims[2] == b->dims[2] && a->var[0] == b->var[0] && a->var[1] == b->var[1] && a->var[2] == b->var[2] &&
         a->p.corotating == b->p.corotating && a->az_zero == b->az_zero && a->p.device == b->p.device;
}

bool same_kind(const Sweep& x, const Sweep& y) { return x.row == y.row && x.axis == y.axis && x.mode == y.mode; }

constexpr size_t kBatchStreams = 4;   // the process's hardware queues

// the shorter list's kinds are a prefix of the longer list's
bool mergeable(const std::vector<Sweep>& a, const std::vector<Sweep>& b) {
  const size_t n = std::min(a.size(), b.size());
  for (size_t k = 0; k < n; ++k)
    if (!same_kind(a[k], b[k])) return false;
  return true;
}

}  // namespace

struct gpue_batch {
  int device = 0;
  std::vector<gpue_ctx*> ctx;
  std::vector<int> active;
  std::vector<int> cls;            // parity class: the instance's step parity at phase 0
  std::vector<Group> groups;
  // at most four streams (default priority), owned here: the groups with the
  // most work get one each, the others share the last
  std::vector<hipStream_t> streams;
  std::vector<size_t> by_work;     // group indices, most work first (launch order of the replays)
  int phase = 0;
  // graph mode 1: one graph per pair of steps for the whole batch, captured on
  // `origin` with a branch per shape group (forked to the group streams and
  // joined back), so a pair of steps of every group is one launch
  int graph_mode = 0;
  hipStream_t origin = nullptr;
  hipEvent_t fork_ev = nullptr;
  std::vector<hipEvent_t> join_ev;   // per group
  hipGraph_t wgraph[2] = {nullptr, nullptr};
  hipGraphExec_t wexec[2] = {nullptr, nullptr};
  double* terms = nullptr;         // pinned: 6 energy terms per instance (gpue_batch_energy)
  std::string err;
};

static gpue_status bfail(gpue_batch* b, gpue_status s, const std::string& m) {
  if (b) b->err = m;
  return s;
}

#define BHIP(b, expr)                                                                           \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) return bfail((b), GPUE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Fill the host argument arrays of every sequence from the members' current
// state; returns true when the set of sequences that are launched changed
// (captured graphs hold the launches, so they are captured again).
static bool fill_args(gpue_batch* b) {
  bool pattern = false;
  for (Group& g : b->groups)
    for (int gs = 0; gs < 2; ++gs)
      for (int ph = 0; ph < 2; ++ph)
        for (BSeq& s : g.seqs[gs][ph]) {
          bool any = false;
          for (size_t m = 0; m < g.members.size(); ++m) {
            const int idx = g.members[m];
            gpue_ctx* c = b->ctx[idx];
            const int par = (b->cls[idx] + ph) % 2;   // the instance's step parity at this phase
            const bool in = b->active[idx] && ((s.mask >> par) & 1);
            any = any || in;
            const std::vector<Sweep>& sc = g.sched[gs][par];
            const Phys phys = make_phys(c, gs, 0.0);
            for (size_t k = 0; k < s.keys.size(); ++k) {
              const bool on = in && k < sc.size();
              const Sweep& sw = on ? sc[k] : s.keys[k];
              SweepArgs& a = s.hargs[k][m];
              std::memset(&a, 0, sizeof(a));
              // every context holds the tables of both plans for each of its
              // axes; a missing one would leave the sweep off rather than fault
              const auto twi = c->tw.find(2 * c->dims[sw.axis] + (c->var[sw.axis] & 1));
              a.active = on && twi != c->tw.end();
              a.in = c->psi;
              a.out = c->psi;
              a.tw = twi == c->tw.end() ? nullptr : twi->second;
              a.ph = phys;
              a.lines = c->dims[1] * c->dims[2];
              a.axis = sw.axis;
              a.gauge_axis = sw.axis;
              a.flags = sw.flags;
              a.invN = sw.invN;
              a.norm_in = c->results;
              a.partials = c->partials;
            }
            if (s.dfin) {
              // the partials come from the instance's F_NORM sweep (its last)
              const Sweep& last = sc.back();
              const int n = c->dims[last.axis];
              const int var = c->var[last.axis];
              FinArgs& f = s.hfin[m];
              std::memset(&f, 0, sizeof(f));
              f.active = in;
              f.nb = last.row ? sweep_blocks(c->prec, true, n, var, c->dims[1] * c->dims[2], 1)
                              : sweep_blocks(c->prec, false, n, var, c->dims[0], last.axis == 1 ? c->dims[2] : c->dims[1]);
              f.na = 1;
              f.mode = 1;
              f.scale = c->d[0] * c->d[1] * c->d[2];
              f.partials = c->partials;
              f.out = c->results;
            }
          }
          if (any != s.any) pattern = true;
          s.any = any;
        }
  return pattern;
}

static gpue_status upload_args(gpue_batch* b) {
  for (Group& g : b->groups)
    for (int gs = 0; gs < 2; ++gs)
      for (int ph = 0; ph < 2; ++ph)
        for (BSeq& s : g.seqs[gs][ph]) {
          for (size_t k = 0; k < s.keys.size(); ++k)
            BHIP(b, hipMemcpy(s.dargs[k], s.hargs[k].data(), s.hargs[k].size() * sizeof(SweepArgs),
                              hipMemcpyHostToDevice));
          if (s.dfin)
            BHIP(b, hipMemcpy(s.dfin, s.hfin.data(), s.hfin.size() * sizeof(FinArgs), hipMemcpyHostToDevice));
        }
  return GPUE_OK;
}

// One step (one phase) of group g for gstate, on the group's stream.
static gpue_status launch_group_phase(gpue_batch* b, Group& g, int gs, int ph) {
  {
    gpue_ctx* c = g.proto;
    for (BSeq& s : g.seqs[gs][ph]) {
      if (!s.any) continue;   // the step parity no member has
      for (size_t k = 0; k < s.keys.size(); ++k) {
        const Sweep& sw = s.keys[k];
        const int n = c->dims[sw.axis];
        const int lx = sw.row ? c->dims[1] * c->dims[2] : c->dims[0];
        const int others = sw.row ? 1 : (sw.axis == 1 ? c->dims[2] : c->dims[1]);
        bool handled = false;
        // FP32 real time: the kernels with FP64 phase factors, as the single-context path
        const bool x = c->prec == 32 && gs != 0 && (sw.row || col_phase_mode(sw.mode));
        hipError_t e = launch_sweep_batched(c->prec, sw.row, n, c->var[sw.axis], x ? (sw.mode | 16) : sw.mode, s.dargs[k],
                                            (int)g.members.size(), lx, others, g.stream, &handled);
        if (!handled || e != hipSuccess) return bfail(b, GPUE_EHIP, "batched sweep launch");
      }
      if (s.dfin) {
        hipError_t e = launch_finalize_batched(s.dfin, (int)g.members.size(), g.stream);
        if (e != hipSuccess) return bfail(b, GPUE_EHIP, "batched finalize launch");
      }
    }
  }
  return GPUE_OK;
}

static gpue_status launch_phase(gpue_batch* b, int gs, int ph) {
  for (Group& g : b->groups) {
    gpue_status st = launch_group_phase(b, g, gs, ph);
    if (st != GPUE_OK) return st;
  }
  return GPUE_OK;
}

static void drop_batch_graphs(gpue_batch* b) {
  for (Group& grp : b->groups)
    for (int g = 0; g < 2; ++g) {
      if (grp.gexec[g]) (void)hipGraphExecDestroy(grp.gexec[g]);
      if (grp.graph[g]) (void)hipGraphDestroy(grp.graph[g]);
      grp.gexec[g] = nullptr;
      grp.graph[g] = nullptr;
    }
  for (int g = 0; g < 2; ++g) {
    if (b->wexec[g]) (void)hipGraphExecDestroy(b->wexec[g]);
    if (b->wgraph[g]) (void)hipGraphDestroy(b->wgraph[g]);
    b->wexec[g] = nullptr;
    b->wgraph[g] = nullptr;
  }
}

static gpue_status sync_groups(gpue_batch* b) {
  for (hipStream_t st : b->streams)
    if (st) BHIP(b, hipStreamSynchronize(st));
  if (b->origin) BHIP(b, hipStreamSynchronize(b->origin));
  return GPUE_OK;
}

// Graph mode 1: capture a pair of steps of every group as one graph on
// `origin`: fork to each group's stream, launch its two phases there, join back.
static gpue_status build_whole_graph(gpue_batch* b, int gstate) {
  BHIP(b, hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
  gpue_status st = GPUE_OK;
  hipError_t e = hipEventRecord(b->fork_ev, b->origin);
  for (size_t k = 0; k < b->groups.size() && st == GPUE_OK && e == hipSuccess; ++k) {
    Group& grp = b->groups[k];
    e = hipStreamWaitEvent(grp.stream, b->fork_ev, 0);
    if (e == hipSuccess) st = launch_group_phase(b, grp, gstate, 0);
    if (st == GPUE_OK && e == hipSuccess) st = launch_group_phase(b, grp, gstate, 1);
    if (st == GPUE_OK && e == hipSuccess) e = hipEventRecord(b->join_ev[k], grp.stream);
    if (st == GPUE_OK && e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->join_ev[k], 0);
  }
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->origin, &g);
  if (st != GPUE_OK || e != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return st != GPUE_OK ? st : bfail(b, GPUE_EHIP, "whole-batch graph capture");
  }
  if (hipGraphInstantiate(&b->wexec[gstate], g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return bfail(b, GPUE_EHIP, "whole-batch graph instantiate");
  }
  b->wgraph[gstate] = g;
  return GPUE_OK;
}

extern "C" {

void gpue_batch_destroy(gpue_batch* b) {
  if (!b) return;
  for (hipStream_t st : b->streams)
    if (st) (void)hipStreamSynchronize(st);
  if (b->origin) (void)hipStreamSynchronize(b->origin);
  drop_batch_graphs(b);
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  for (hipEvent_t ev : b->join_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (b->origin) (void)hipStreamDestroy(b->origin);
  for (Group& g : b->groups)
    for (int gs = 0; gs < 2; ++gs)
      for (int ph = 0; ph < 2; ++ph)
        for (BSeq& s : g.seqs[gs][ph]) {
          for (auto* p : s.dargs)
            if (p) (void)hipFree(p);
          if (s.dfin) (void)hipFree(s.dfin);
        }
  for (hipStream_t st : b->streams)
    if (st) (void)hipStreamDestroy(st);
  if (b->terms) (void)hipHostFree(b->terms);
  delete b;
}

const char* gpue_batch_last_error(const gpue_batch* b) { return b ? b->err.c_str() : "null batch"; }

gpue_status gpue_batch_set_graph_mode(gpue_batch* b, int mode) {
  if (!b || mode < 0 || mode > 2) return GPUE_EINVAL;
  if (mode != b->graph_mode) {
    gpue_status st = sync_groups(b);
    if (st != GPUE_OK) return st;
    drop_batch_graphs(b);
    b->graph_mode = mode;
  }
  return GPUE_OK;
}

gpue_status gpue_batch_create(gpue_ctx* const* ctxs, int n, gpue_batch** out) {
  if (!ctxs || n <= 0 || !out) return GPUE_EINVAL;
  *out = nullptr;
  gpue_batch* b = new gpue_batch();
  for (int i = 0; i < n; ++i) {
    gpue_ctx* c = ctxs[i];
    if (!c || c->t_dep || c->ramp_type || c->p.device != ctxs[0]->p.device) {
      gpue_batch_destroy(b);
      return GPUE_EINVAL;   // time-dependent V or ramp and mixed devices are run per context
    }
    b->ctx.push_back(c);
    b->active.push_back(1);
    b->cls.push_back(c->next_step % 2);
  }
  b->device = ctxs[0]->p.device;
  if (hipSetDevice(b->device) != hipSuccess ||
      hipHostMalloc((void**)&b->terms, (size_t)n * 6 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    gpue_batch_destroy(b);
    return GPUE_EHIP;
  }
  for (int i = 0; i < n; ++i) {
    bool placed = false;
    for (Group& g : b->groups)
      if (same_shape(g.proto, b->ctx[i])) {
        g.members.push_back(i);
        placed = true;
        break;
      }
    if (!placed) {
      Group g;
      g.proto = b->ctx[i];
      g.members.push_back(i);
      b->groups.push_back(g);
    }
  }
  {
    // A process gets four hardware queues (GPU_MAX_HW_QUEUES); more streams
    // than that share queues in FIFO order wherever the runtime places them.
    // So the groups are ranked by work (instances x points), the first three
    // get a stream each and the rest share the last one, where their graphs
    // run one after another.
    std::vector<size_t> order(b->groups.size());
    std::iota(order.begin(), order.end(), size_t(0));
    auto work = [&](size_t k) { return (double)b->groups[k].members.size() * (double)b->groups[k].proto->N; };
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work(x) > work(y); });
    const size_t ns = std::min<size_t>(b->groups.size(), kBatchStreams);
    for (size_t i = 0; i < ns; ++i) {
      hipStream_t st = nullptr;
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        gpue_batch_destroy(b);
        return GPUE_EHIP;
      }
      b->streams.push_back(st);
    }
    for (size_t r = 0; r < order.size(); ++r) b->groups[order[r]].stream = b->streams[std::min(r, ns - 1)];
    b->by_work = order;
  }
  b->join_ev.assign(b->groups.size(), nullptr);
  if (hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) != hipSuccess) {
    gpue_batch_destroy(b);
    return GPUE_EHIP;
  }
  for (hipEvent_t& ev : b->join_ev)
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      gpue_batch_destroy(b);
      return GPUE_EHIP;
    }
  for (Group& g : b->groups)
    for (int gs = 0; gs < 2; ++gs) {
      // imaginary time: the norm slot holds 1.0 whenever nothing is pending, so
      // the renormalising first sweep is always safe
      for (int par = 0; par < 2; ++par) schedule(g.proto, gs, par, gs == 0, g.sched[gs][par]);
      const std::vector<Sweep>& s0 = g.sched[gs][0];
      const std::vector<Sweep>& s1 = g.sched[gs][1];
      for (int ph = 0; ph < 2; ++ph) {
        std::vector<BSeq>& seqs = g.seqs[gs][ph];
        if (mergeable(s0, s1)) {
          seqs.resize(1);
          seqs[0].keys = s0.size() >= s1.size() ? s0 : s1;
          seqs[0].mask = 3;
        } else {
          seqs.resize(2);
          for (int par = 0; par < 2; ++par) {
            seqs[par].keys = g.sched[gs][par];
            seqs[par].mask = 1 << par;
          }
        }
        const size_t M = g.members.size();
        for (BSeq& s : seqs) {
          s.hargs.assign(s.keys.size(), std::vector<SweepArgs>(M));
          s.dargs.assign(s.keys.size(), nullptr);
          for (size_t k = 0; k < s.keys.size(); ++k)
            if (hipMalloc((void**)&s.dargs[k], M * sizeof(SweepArgs)) != hipSuccess) {
              gpue_batch_destroy(b);
              return GPUE_ENOMEM;
            }
          if (gs == 0) {
            s.hfin.assign(M, FinArgs{});
            if (hipMalloc((void**)&s.dfin, M * sizeof(FinArgs)) != hipSuccess) {
              gpue_batch_destroy(b);
              return GPUE_ENOMEM;
            }
          }
        }
      }
    }
  fill_args(b);
  if (upload_args(b) != GPUE_OK) {
    gpue_batch_destroy(b);
    return GPUE_EHIP;
  }
  *out = b;
  return GPUE_OK;
}

gpue_status gpue_batch_set_active(gpue_batch* b, int index, int active) {
  if (!b || index < 0 || index >= (int)b->ctx.size()) return GPUE_EINVAL;
  b->active[index] = active ? 1 : 0;
  // graphs keep their argument pointers, only the contents change, unless a
  // sequence starts or stops being launched (every evolve call ends synchronised,
  // so no replay is in flight here)
  if (fill_args(b)) drop_batch_graphs(b);
  return upload_args(b);
}

// Advance every active instance by nsteps steps of gstate.
gpue_status gpue_batch_evolve(gpue_batch* b, int nsteps, int gstate) {
  if (!b || nsteps < 0 || (gstate != 0 && gstate != 1)) return GPUE_EINVAL;
  // settle per-instance state on the instances' own streams first
  bool changed = false;
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    gpue_ctx* c = b->ctx[i];
    if (!b->active[i]) continue;
    if (gstate == 1 && c->pending_norm) {
      if (apply_pending(c) != GPUE_OK) return bfail(b, GPUE_EHIP, c->err);
    }
    BHIP(b, hipStreamSynchronize(c->stream));
    const int cls = (c->next_step + b->phase) % 2;   // parity at phase 0 of the current pair
    if (cls != b->cls[i]) {
      b->cls[i] = cls;
      changed = true;
    }
  }
  if (changed) {
    if (fill_args(b)) drop_batch_graphs(b);
    gpue_status st = upload_args(b);
    if (st != GPUE_OK) return st;
  }
  int s = 0;
  if (b->phase == 1 && s < nsteps) {   // finish a pair started by an odd call
    gpue_status st = launch_phase(b, gstate, 1);
    if (st != GPUE_OK) return st;
    b->phase = 0;
    ++s;
  }
  if (nsteps - s >= 2 && b->graph_mode == 1) {
    if (!b->wexec[gstate]) {
      gpue_status st = build_whole_graph(b, gstate);
      if (st != GPUE_OK) return st;
    }
    // the whole graph runs on `origin` after the groups' earlier work, and their
    // later work waits for it
    for (size_t k = 0; k < b->groups.size(); ++k) {
      BHIP(b, hipEventRecord(b->join_ev[k], b->groups[k].stream));
      BHIP(b, hipStreamWaitEvent(b->origin, b->join_ev[k], 0));
    }
    for (; nsteps - s >= 2; s += 2) BHIP(b, hipGraphLaunch(b->wexec[gstate], b->origin));
    BHIP(b, hipEventRecord(b->fork_ev, b->origin));
    for (Group& grp : b->groups) BHIP(b, hipStreamWaitEvent(grp.stream, b->fork_ev, 0));
  }
  if (b->graph_mode == 2) {   // no graphs: every phase launched on its own (kernel traces)
    for (; nsteps - s >= 2; s += 2) {
      gpue_status st = launch_phase(b, gstate, 0);
      if (st == GPUE_OK) st = launch_phase(b, gstate, 1);
      if (st != GPUE_OK) return st;
    }
  }
  if (nsteps - s >= 2) {
    for (Group& grp : b->groups) {
      if (grp.gexec[gstate]) continue;
      BHIP(b, hipStreamBeginCapture(grp.stream, hipStreamCaptureModeThreadLocal));
      gpue_status st = launch_group_phase(b, grp, gstate, 0);
      if (st == GPUE_OK) st = launch_group_phase(b, grp, gstate, 1);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(grp.stream, &g);
      if (st != GPUE_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      if (e != hipSuccess) return bfail(b, GPUE_EHIP, "batch graph capture");
      if (hipGraphInstantiate(&grp.gexec[gstate], g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return bfail(b, GPUE_EHIP, "batch graph instantiate");
      }
      grp.graph[gstate] = g;
    }
    // pairs of steps, the groups' graphs interleaved across their streams, the
    // group with the most work (the critical path) first
    for (; nsteps - s >= 2; s += 2)
      for (size_t k : b->by_work) BHIP(b, hipGraphLaunch(b->groups[k].gexec[gstate], b->groups[k].stream));
  }
  if (s < nsteps) {
    gpue_status st = launch_phase(b, gstate, 0);
    if (st != GPUE_OK) return st;
    b->phase = 1;
    ++s;
  }
  {
    gpue_status st = sync_groups(b);
    if (st != GPUE_OK) return st;
  }
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    if (!b->active[i]) continue;
    b->ctx[i]->next_step += nsteps;
    if (gstate == 0) b->ctx[i]->pending_norm = true;
  }
  return GPUE_OK;
}

// Host time per step of `steps` steps of the whole batch (after two steps that
// capture any graph the current choice needs).
static gpue_status time_batch(gpue_batch* b, int gstate, int steps, double* us) {
  gpue_status st = gpue_batch_evolve(b, 2, gstate);
  if (st != GPUE_OK) return st;
  const auto t0 = std::chrono::steady_clock::now();
  st = gpue_batch_evolve(b, steps, gstate);   // returns after the groups' streams finished
  if (st != GPUE_OK) return st;
  *us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / steps;
  return GPUE_OK;
}

// Every member of group g takes variant v on axis a; the launch arguments and
// graphs of the batch follow.
static gpue_status set_group_variant(gpue_batch* b, Group& g, int a, int v) {
  for (int idx : g.members) {
    gpue_ctx* c = b->ctx[idx];
    int vars[3] = {c->var[0], c->var[1], c->var[2]};
    vars[a] = v;
    gpue_status st = gpue_set_variants(c, vars);
    if (st != GPUE_OK) return bfail(b, st, c->err);
  }
  fill_args(b);
  drop_batch_graphs(b);
  return upload_args(b);
}

gpue_status gpue_batch_autotune(gpue_batch* b, int gstate, int steps, double* us_per_step, int* graph_mode) {
  if (!b || steps < 2 || (gstate != 0 && gstate != 1)) return GPUE_EINVAL;
  gpue_status st = sync_groups(b);
  if (st != GPUE_OK) return st;
  // save every instance's state: psi into its scratch buffer, the step counter,
  // the pending norm and its value
  const size_t n = b->ctx.size();
  std::vector<int> next(n), pend(n), cls = b->cls;
  std::vector<double> nrm(n);
  const int phase = b->phase;
  for (size_t i = 0; i < n; ++i) {
    gpue_ctx* c = b->ctx[i];
    next[i] = c->next_step;
    pend[i] = c->pending_norm;
    BHIP(b, hipMemcpyAsync(c->tmp, c->psi, c->N * c->esz, hipMemcpyDeviceToDevice, c->stream));
    BHIP(b, hipMemcpyAsync(&nrm[i], c->results, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BHIP(b, hipStreamSynchronize(c->stream));
  }
  double best = 0.0;
  st = time_batch(b, gstate, steps, &best);
  // coordinate descent: groups with the most points first, then each axis
  std::vector<size_t> order(b->groups.size());
  for (size_t k = 0; k < order.size(); ++k) order[k] = k;
  std::sort(order.begin(), order.end(), [&](size_t x, size_t y) {
    return b->groups[x].proto->N * (long)b->groups[x].members.size() >
           b->groups[y].proto->N * (long)b->groups[y].members.size();
  });
  for (size_t k = 0; k < order.size() && st == GPUE_OK; ++k) {
    Group& g = b->groups[order[k]];
    for (int a = 0; a < g.proto->p.dimnum && st == GPUE_OK; ++a) {
      const int keep = g.proto->var[a];
      int pick = keep;
      // row sweeps: plan and twiddle loading (variants 0-3); strided: plan and tile width (0-5)
      const int nv = a == 0 ? 4 : N_VARIANTS;
      for (int v = 0; v < nv && st == GPUE_OK; ++v) {
        if (v == keep) continue;
        st = set_group_variant(b, g, a, v);
        double us = 0.0;
        if (st == GPUE_OK) st = time_batch(b, gstate, steps, &us);
        if (st == GPUE_OK && us < best) {
          best = us;
          pick = v;
        }
      }
      if (st == GPUE_OK && g.proto->var[a] != pick) st = set_group_variant(b, g, a, pick);
    }
  }
  // then the graph mode: per-group graphs or one graph for the whole batch
  if (st == GPUE_OK) {
    const int m0 = b->graph_mode == 2 ? 0 : b->graph_mode, m1 = 1 - m0;
    st = gpue_batch_set_graph_mode(b, m0);
    double us0 = best, us1 = 0.0;
    if (st == GPUE_OK) st = time_batch(b, gstate, steps, &us0);
    if (st == GPUE_OK) st = gpue_batch_set_graph_mode(b, m1);
    if (st == GPUE_OK) st = time_batch(b, gstate, steps, &us1);
    if (st == GPUE_OK && us0 <= us1) st = gpue_batch_set_graph_mode(b, m0);
    best = std::min(us0, us1);
  }
  // restore the states (also after a failure, as far as the device allows)
  gpue_status rs = sync_groups(b);
  for (size_t i = 0; i < n; ++i) {
    gpue_ctx* c = b->ctx[i];
    if (hipMemcpyAsync(c->psi, c->tmp, c->N * c->esz, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->results, &nrm[i], sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rs = bfail(b, GPUE_EHIP, "restoring an instance after tuning");
    c->next_step = next[i];
    c->pending_norm = pend[i];
  }
  b->cls = cls;
  b->phase = phase;
  fill_args(b);
  drop_batch_graphs(b);
  if (rs == GPUE_OK) rs = upload_args(b);
  if (st == GPUE_OK) st = rs;
  if (st == GPUE_OK && us_per_step) *us_per_step = best;
  if (st == GPUE_OK && graph_mode) *graph_mode = b->graph_mode;
  return st;
}

gpue_status gpue_batch_energy(gpue_batch* b, double* energy, double* mu) {
  if (!b || !energy || !mu) return GPUE_EINVAL;
  const size_t n = b->ctx.size();
  for (size_t i = 0; i < n; ++i) {   // queue every instance's evaluation first
    gpue_status st = energy_launch_graph(b->ctx[i], b->terms + 6 * i);
    if (st != GPUE_OK) return bfail(b, st, b->ctx[i]->err);
  }
  gpue_status out = GPUE_OK;
  for (size_t i = 0; i < n; ++i) {
    gpue_ctx* c = b->ctx[i];
    BHIP(b, hipStreamSynchronize(c->stream));
    energy_combine(c, b->terms + 6 * i, &energy[i], &mu[i]);
    if (!std::isfinite(energy[i]) && out == GPUE_OK) out = bfail(b, GPUE_ENONFINITE, "energy is not finite");
  }
  return out;
}

}  // extern "C"
