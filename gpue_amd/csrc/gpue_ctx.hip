// gpue_ctx.hip -- the C ABI (include/gpue_amd.h): per-problem context, the
// step schedule of fused sweeps, energy, renormalisation and vortex detection.
//
// Host control flow restates GPUE's evolve() (src/evolution.cu:260-908) with
// the host I/O left to the caller; device work is queued on one HIP stream
// per context and nothing is allocated inside the step loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ctx_internal.h"
#include "eqn.h"

namespace gpue {
void make_grids(int dimnum, int xDim, int yDim, int zDim, double xMax, double yMax, double zMax,
                std::vector<double>* g, double* d);
void vortex_mask(const double* x, int xDim, double radius, std::vector<unsigned char>& m);
void twiddles(int n, std::vector<double>& re, std::vector<double>& im);
}  // namespace gpue

using namespace gpue;

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
template <int N, int V>
static void radices_of(std::vector<int>& R) {
  R.assign(Plan<N, V>::R, Plan<N, V>::R + Plan<N, V>::NP);
}
static bool plan_radices(int n, int plan, std::vector<int>& R) {
  switch (n) {
#define GPUE_RAD(N_) \
  case N_: plan ? radices_of<N_, 1>(R) : radices_of<N_, 0>(R); return true;
    GPUE_RAD(16) GPUE_RAD(32) GPUE_RAD(48) GPUE_RAD(64) GPUE_RAD(96) GPUE_RAD(128) GPUE_RAD(192) GPUE_RAD(256)
    GPUE_RAD(384) GPUE_RAD(512) GPUE_RAD(1024) GPUE_RAD(2048) GPUE_RAD(4096)
#undef GPUE_RAD
    default: return false;
  }
}
// Entry order of the device pass tables (TwTab in fft_core.h): the forward
// radix order's passes, then the reverse order's; pass P with L > 1 gives
// L*(R-1) entries [t-1][k] = index t*k*N/(L*R) of the full-circle table.
static void pass_table(int n, const std::vector<int>& R, std::vector<int>& idx) {
  idx.clear();
  for (int ord = 0; ord < 2; ++ord) {
    std::vector<int> seq = R;
    if (ord) std::reverse(seq.begin(), seq.end());
    int L = 1;
    for (int r : seq) {
      if (L > 1) {
        const int step = n / (L * r);
        for (int t = 1; t < r; ++t)
          for (int k = 0; k < L; ++k) idx.push_back(t * k * step);
      }
      L *= r;
    }
  }
}
gpue_status gpue::fail(gpue_ctx* c, gpue_status s, const std::string& msg) {
  if (c) c->err = msg;
  return s;
}

#define HIPCHK(c, expr)                                                                      \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail((c), GPUE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

static double* dev_copy(const std::vector<double>& v, hipError_t* e) {
  double* p = nullptr;
  *e = hipMalloc((void**)&p, std::max<size_t>(1, v.size()) * sizeof(double));
  if (*e != hipSuccess) return nullptr;
  *e = hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice);
  return p;
}

Phys gpue::make_phys(const gpue_ctx* c, int gstate, double time) {
  const gpue_params& p = c->p;
  Phys f;
  std::memset(&f, 0, sizeof(f));
  f.xDim = c->dims[0];
  f.yDim = c->dims[1];
  f.zDim = c->dims[2];
  f.dimnum = p.dimnum;
  f.gstate = gstate;
  f.gpe = p.gpe;
  f.v_kind = c->v_kind;
  f.a_kind = c->a_kind;
  f.k_kind = c->k_kind;
  const double Dt = gstate == 0 ? p.gdt : p.dt;   // src/evolution.cu:344-351
  f.cV = Dt / (2 * p.hbar);
  f.cH = (0.5 * Dt) / p.hbar;
  f.g = p.g;
  f.cK = Dt / p.hbar;
  f.Dt = Dt;
  f.hb2m = p.hbar * p.hbar / (2 * p.mass);
  f.mass = p.mass;
  f.omega = p.omega;
  f.omegaX = p.omegaX;
  f.omegaY = p.omegaY;
  f.omegaZe = p.dimnum == 3 ? p.omegaZ : 0.0;   // items[5] (src/operators.cu:765-767)
  f.gammaY = p.gammaY;
  f.x0 = p.x0_shift;
  f.y0 = p.y0_shift;
  f.z0e = p.dimnum == 3 ? p.z0_shift : 0.0;     // items[8]
  for (int a = 0; a < 3; ++a) {
    (&f.xg)[a] = c->dgrid[a];
    (&f.pxg)[a] = c->dgrid[a + 3];
    f.vs[a] = c->vs[a];
    f.vsf[a] = c->vsf[a];
    f.kff[a] = c->kff[gstate][a];
    f.ks[a] = c->ks[a];
    f.kf[a] = c->kf[gstate][a];
  }
  f.Vt = c->Vt;
  f.Kt = c->Kt;
  for (int a = 0; a < 3; ++a) f.At[a] = c->At[a];
  for (int k = 0; k < 5; ++k) f.eq[k] = c->deq[k];
  for (int a = 0; a < 3; ++a) {
    f.vst[a] = c->vst[a];
    f.dxyz[a] = c->d[a];
  }
  // find_field's coordinates dx*i - xMax (src/dynamic.cu:613-615)
  f.org[0] = -p.xMax;
  f.org[1] = p.dimnum > 1 ? -p.yMax : 0.0;
  f.org[2] = p.dimnum > 2 ? -p.zMax : 0.0;
  f.t = time;
  f.eqk = c->t_dep ? 1 : 0;
  return f;
}

// -R ramp (src/evolution.cu:409-428, 811-832): GPUE raises the gauge operators
// to omega_0 in place every step, so after step i they carry the exponent
// times F_i = omega_0(0) * ... * omega_0(i): omega/(i+1) (type 1) or
// omega*(i+1)/numSteps (type 2).  The same F_i scales the gauge exponent here,
// in imaginary and real time (GPUE's element-wise pow of a complex phase is
// not a phase; the exponent scaling is what the imaginary-time code computes).
double gpue::ramp_scale(const gpue_ctx* c, int step) {
  if (!c->ramp_type || !c->p.corotating) return 1.0;
  const double w = c->p.omega;
  if (c->ramp_type == 1) return w / (double)(step + 1);
  return w * (double)(step + 1) / (double)c->ramp_steps;
}

// Sweeps of step i (src/evolution.cu:727-870): V/2, the forward transforms, K,
// the inverse transforms, V/2, then the gauge operators in GPUE's order (x -> y
// (-> z) on odd steps, (z ->) y -> x on even ones, src/evolution.cu:837,842).
// The transforms of different axes commute, so the step is led by the axis
// whose gauge operator comes first: its closing sweep [IFFT, V/2, FFT, A, IFFT]
// then applies that operator too, and a 2D gauge step takes 4 sweeps in either
// order (3D: 6, or 7 with a non-zero Az) instead of 5 when the order starts on
// another axis.  The K step runs on the innermost axis.  NORM is set on the
// last sweep of an imaginary-time step; SCALE_IN on the first when a norm is
// pending.
void gpue::schedule(const gpue_ctx* c, int gstate, int i, bool pending, std::vector<Sweep>& out) {
  out.clear();
  const gpue_params& p = c->p;
  const int n[3] = {c->dims[0], c->dims[1], c->dims[2]};
  const double invN = 1.0 / (double)c->N;
  const unsigned scale = pending ? F_SCALE_IN : 0u;
  const bool flip = (i % 2) != 0;
  if (p.dimnum == 1) {
    unsigned f = scale | F_HALFV_PRE | F_HALFV_POST | (p.corotating ? F_GAUGE_POST : 0u);
    out.push_back({true, 0, RM_FULL, f, invN});
    if (gstate == 0) out.back().flags |= F_NORM;
    return;
  }
  // gauge axes in GPUE's order for this step (the z operator is the identity when Az = 0)
  int G[3], ng = 0;
  if (p.corotating) {
    const bool zg = p.dimnum == 3 && !c->az_zero;
    if (flip) {
      G[ng++] = 0;
      G[ng++] = 1;
      if (zg) G[ng++] = 2;
    } else {
      if (zg) G[ng++] = 2;
      G[ng++] = 1;
      G[ng++] = 0;
    }
  }
  const int lead = ng ? G[0] : 0;
  int order[3], no = 0;   // forward transform order: the lead axis first, the K axis last
  order[no++] = lead;
  for (int a = 0; a < p.dimnum; ++a)
    if (a != lead) order[no++] = a;
  auto fwd = [&](int a, unsigned f) -> Sweep {
    return a == 0 ? Sweep{true, 0, RM_FWD, f, 0.0} : Sweep{false, a, f ? (int)CM_FWD_V : (int)CM_FWD, f, 0.0};
  };
  auto inv = [&](int a) -> Sweep { return a == 0 ? Sweep{true, 0, RM_INV, 0u, 0.0} : Sweep{false, a, CM_INV, 0u, 0.0}; };
  out.push_back(fwd(lead, scale | F_HALFV_PRE));
  for (int k = 1; k + 1 < no; ++k) out.push_back(fwd(order[k], 0u));
  const int kax = order[no - 1];
  out.push_back(kax == 0 ? Sweep{true, 0, RM_K, 0u, invN} : Sweep{false, kax, CM_K, 0u, invN});
  for (int k = no - 2; k >= 1; --k) out.push_back(inv(order[k]));
  // closing sweep of the lead axis: IFFT, V/2 and, with a gauge, that axis' operator
  const unsigned fc = F_HALFV_POST | (ng ? F_GAUGE_POST : 0u);
  const double invc = ng ? 1.0 / n[lead] : 0.0;
  out.push_back(lead == 0 ? Sweep{true, 0, RM_INV, fc, invc} : Sweep{false, lead, CM_INV_VG, fc, invc});
  for (int k = 1; k < ng; ++k)
    out.push_back(G[k] == 0 ? Sweep{true, 0, RM_GAUGE, 0u, 1.0 / n[0]} : Sweep{false, G[k], CM_GAUGE, 0u, 1.0 / n[G[k]]});
  if (gstate == 0) out.back().flags |= F_NORM;
}

// The sweep that opens a step and may rescale its input by a pending norm.
static bool opens_step(const Sweep& s) {
  return s.row ? s.mode == RM_FWD : s.mode == CM_FWD_V;
}

static int default_variant(const gpue_ctx* c, int axis);

// An opening sweep that can reduce the previous step's norm partials itself
// (part_norm needs workgroups of whole waves).
static bool takes_deferred_norm(const gpue_ctx* c, const Sweep& s) {
  const int var = c->t_dep ? default_variant(c, s.axis) : s.var >= 0 ? s.var : c->var[s.axis];
  return opens_step(s) && sweep_full_waves(c->prec, s.row, c->dims[s.axis], var);
}

static std::string sweep_name(const Sweep& s) {
  static const char* rm[] = {"row_fwd", "row_inv", "row_gauge", "row_full", "row_ksum", "row_inv_fwd",
                             "row_gauge_fwd", "row_k"};
  static const char* cm[] = {"fwd", "inv", "k", "gauge", "ksum", "fwdv", "invvg", "gauge_fwd"};
  std::string n = s.row ? rm[s.mode] : (std::string(s.axis == 1 ? "coly_" : "colz_") + cm[s.mode]);
  if (s.flags & F_HALFV_PRE) n += "+vpre";
  if (s.flags & F_HALFV_POST) n += "+vpost";
  if (s.flags & F_GAUGE_POST) n += "+gx";
  if (s.flags & F_NORM) n += "+norm";
  return n;
}

static int kind_index(gpue_ctx* c, const std::string& n) {
  for (size_t k = 0; k < c->kind_names.size(); ++k)
    if (c->kind_names[k] == n) return (int)k;
  c->kind_names.push_back(n);
  c->kind_n.push_back(0);
  c->kind_ms.push_back(0.0);
  return (int)c->kind_names.size() - 1;
}

static gpue_status timing_flush(gpue_ctx* c) {
  if (c->ev_used == 0) return GPUE_OK;
  HIPCHK(c, hipEventSynchronize(c->ev[2 * (c->ev_used - 1) + 1]));
  for (int k = 0; k < c->ev_used; ++k) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]));
    c->kind_ms[c->ev_kind[k]] += ms;
    c->kind_n[c->ev_kind[k]] += 1;
  }
  c->ev_used = 0;
  return GPUE_OK;
}

// Workgroups (= partials written) of a sweep.
static int sweep_nb(const gpue_ctx* c, const Sweep& s) {
  const int n = c->dims[s.axis];
  const int var = c->t_dep ? default_variant(c, s.axis) : s.var >= 0 ? s.var : c->var[s.axis];
  return s.row ? sweep_blocks(c->prec, true, n, var, c->dims[1] * c->dims[2], 1)
               : sweep_blocks(c->prec, false, n, var, c->dims[0], s.axis == 1 ? c->dims[2] : c->dims[1]);
}

// Deferred renormalisation: the norm sweep `last` of one imaginary step leaves
// its partials to `next`, the first sweep of the following step, which reduces
// them itself (part_norm, the same additions as the finalize) instead of a
// one-workgroup finalize kernel and its dispatch gap between the two.  Only
// for fewer than 4096 partials (every workgroup of `next` sums all of them)
// and when `next` is the row sweep that scales its input.
static bool defer_norm(const gpue_ctx* c, Sweep& last, Sweep& next) {
  if (c->p.dimnum < 2 || !(last.flags & F_NORM) || !(next.flags & F_SCALE_IN) || !takes_deferred_norm(c, next))
    return false;
  const int nb = sweep_nb(c, last);
  if (nb >= 4096) return false;
  last.flags |= F_DEFER;
  next.flags |= F_SCALE_PART;
  next.part_nb = nb;
  return true;
}

// A finalize left by an eager step whose successor did not take it.
static gpue_status settle_deferred(gpue_ctx* c) {
  if (c->defer_nb <= 0) return GPUE_OK;
  const double dV = c->d[0] * c->d[1] * c->d[2];
  hipError_t e = launch_finalize(c->partials, c->defer_nb, 1, 1, dV, c->results + 0, c->stream);
  c->defer_nb = 0;
  if (e != hipSuccess) return fail(c, GPUE_EHIP, std::string("finalize: ") + hipGetErrorString(e));
  return GPUE_OK;
}

// The variants every context starts with (no timing, so results never depend on
// it): for small, latency-bound grids narrow lines with up-front twiddle loads
// and quarter strided tiles (the autotuner's choice for 2D grids of
// 128^2-512^2), wide lines otherwise.  Contexts with time-dependent fields
// always run these (their EQ kernels exist for these variants only).
static int default_variant(const gpue_ctx* c, int axis) {
  const bool small = c->N <= (1L << 20);
  return axis == 0 ? (small ? 3 : 0) : (small ? 5 : 0);
}

gpue_status gpue::launch(gpue_ctx* c, const Sweep& s, const Phys& ph, const void* in, void* out) {
  const int n = c->dims[s.axis];
  const int var = c->t_dep ? default_variant(c, s.axis) : s.var >= 0 ? s.var : c->var[s.axis];
  SweepArgs a;
  std::memset(&a, 0, sizeof(a));
  a.active = 1;
  a.in = in;
  a.out = out;
  const auto twi = c->tw.find(2 * n + (var & 1));
  if (twi == c->tw.end()) return fail(c, GPUE_EUNSUPPORTED, "no twiddle table for n=" + std::to_string(n));
  a.tw = twi->second;
  a.ph = ph;
  a.lines = c->dims[1] * c->dims[2];
  a.axis = s.axis;
  a.gauge_axis = s.axis;
  a.flags = s.flags;
  a.invN = s.invN;
  a.norm_in = c->results + 0;
  a.partials = c->partials;
  if (s.flags & F_SCALE_PART) {
    a.norm_nb = s.part_nb;
    a.norm_scale = c->d[0] * c->d[1] * c->d[2];
    a.norm_out = c->results + 0;
  }
  int slot = -1;
  if (c->timing) {
    if (c->ev_used == TIMING_PAIRS) {
      gpue_status st = timing_flush(c);
      if (st != GPUE_OK) return st;
    }
    slot = c->ev_used++;
    c->ev_kind[slot] = kind_index(c, sweep_name(s));
    HIPCHK(c, hipEventRecord(c->ev[2 * slot], c->stream));
  }
  bool handled = false;
  hipError_t e = launch_sweep(c->prec, s.row, n, var, s.mode, a, c->stream, &handled);
  if (!handled) return fail(c, GPUE_EUNSUPPORTED, "no sweep kernel for n=" + std::to_string(n));
  if (e != hipSuccess) return fail(c, GPUE_EHIP, std::string("sweep launch: ") + hipGetErrorString(e));
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[2 * slot + 1], c->stream));
  // second stage of the fused reductions
  const bool ksum = s.mode == (s.row ? (int)RM_KSUM : (int)CM_KSUM);
  if (ksum || ((s.flags & F_NORM) && !(s.flags & F_DEFER))) {
    const int nb = sweep_nb(c, s);
    const double dV = c->d[0] * c->d[1] * c->d[2];
    // KSUM sums land in results[1] (kinetic), with F_ASUM in results[6 + axis],
    // with F_KAX in results[9 + axis]
    const int slot = (s.flags & F_ASUM) ? 6 + s.axis : (s.flags & F_KAX) ? 9 + s.axis : 1;
    e = ksum ? launch_finalize(c->partials, nb, 1, 0, s.invN, c->results + slot, c->stream)
             : launch_finalize(c->partials, nb, 1, 1, dV, c->results + 0, c->stream);
    if (e != hipSuccess) return fail(c, GPUE_EHIP, std::string("finalize: ") + hipGetErrorString(e));
  }
  return GPUE_OK;
}

gpue_status gpue::apply_pending(gpue_ctx* c) {
  if (!c->pending_norm) return GPUE_OK;
  hipError_t e = c->prec == 64 ? launch_apply_norm<double>(c->psi, c->N, c->results, c->stream)
                               : launch_apply_norm<float>(c->psi, c->N, c->results, c->stream);
  if (e == hipSuccess) e = launch_set_double(c->results, 1.0, c->stream);   // nothing pending: norm 1
  if (e != hipSuccess) return fail(c, GPUE_EHIP, hipGetErrorString(e));
  c->pending_norm = false;
  return GPUE_OK;
}

static bool valid_dim(int n) { return sweep_supported(n); }

// Any change of fields, parameters or kernel variants invalidates the captured
// graphs.  Replays may still be queued (evolve returns without waiting in real
// time), so the stream is drained before an executable graph is destroyed.
static void drop_energy_graph(gpue_ctx* c) {
  if (!c->egexec) return;
  (void)hipStreamSynchronize(c->stream);   // a replay may be queued (the context's own stream)
  (void)hipGraphExecDestroy(c->egexec);
  if (c->egraph) (void)hipGraphDestroy(c->egraph);
  c->egexec = nullptr;
  c->egraph = nullptr;
  c->egraph_dst = nullptr;
}

static void drop_graphs(gpue_ctx* c) {
  drop_energy_graph(c);
  if (c->graphs.empty()) return;
  if (c->gev) (void)hipEventSynchronize(c->gev);   // the last replay (graph launches record it)
  else if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (StepGraph& g : c->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  c->graphs.clear();
}

// Per-axis terms of kharmonic_V (src/operators.cu:994-1008) with the rotation
// gauge's 1/2 m |A|^2 (Ax = -y omega omegaX, Ay = x omega omegaY, Az = 0,
// src/operators.cu:621-647) and of simple_K (src/operators.cu:460-469).
static void sep_tables(const gpue_ctx* c, std::vector<double>* vs, std::vector<double>* ks) {
  const gpue_params& p = c->p;
  const double m = p.mass, hb2m = p.hbar * p.hbar / (2 * p.mass);
  const bool rot = c->a_kind == A_ROTATION;
  const double oZ = p.dimnum == 3 ? p.omegaZ : 0.0, z0 = p.dimnum == 3 ? p.z0_shift : 0.0;
  for (int a = 0; a < 3; ++a) {
    const std::vector<double>& x = c->grid[a];
    const std::vector<double>& k = c->grid[a + 3];
    vs[a].resize(x.size());
    ks[a].resize(k.size());
    for (size_t i = 0; i < x.size(); ++i) {
      double t, A = 0.0;
      if (a == 0) { t = p.omegaX * (x[i] + p.x0_shift); if (rot) A = x[i] * p.omega * p.omegaY; }
      else if (a == 1) { t = p.gammaY * p.omegaY * (x[i] + p.y0_shift); if (rot) A = -x[i] * p.omega * p.omegaX; }
      else { t = 1.0 * oZ * (x[i] + z0); }
      vs[a][i] = 0.5 * m * (t * t + A * A);
      ks[a][i] = hb2m * (k[i] * k[i]);
    }
  }
}

// A full V table in GPUE's kharmonic_V form for the current A tables.
static gpue_status materialise_V(gpue_ctx* c) {
  const gpue_params& p = c->p;
  std::vector<double> A[3];
  for (int a = 0; a < 3; ++a) {
    A[a].assign(c->N, 0.0);
    if (c->At[a]) HIPCHK(c, hipMemcpy(A[a].data(), c->At[a], c->N * sizeof(double), hipMemcpyDeviceToHost));
  }
  const double oZ = p.dimnum == 3 ? p.omegaZ : 0.0, z0 = p.dimnum == 3 ? p.z0_shift : 0.0;
  std::vector<double> V(c->N);
  for (long gid = 0; gid < c->N; ++gid) {
    long xi = gid % c->dims[0], yi = (gid / c->dims[0]) % c->dims[1], zi = gid / ((long)c->dims[0] * c->dims[1]);
    double Vx = p.omegaX * (c->grid[0][xi] + p.x0_shift);
    double Vy = p.gammaY * p.omegaY * (c->grid[1][yi] + p.y0_shift);
    double Vz = 1.0 * oZ * (c->grid[2][zi] + z0);
    V[gid] = 0.5 * p.mass * ((Vx * Vx + Vy * Vy + Vz * Vz) + (A[0][gid] * A[0][gid] + A[1][gid] * A[1][gid] +
                                                            A[2][gid] * A[2][gid]));
  }
  if (!c->Vt && hipMalloc((void**)&c->Vt, c->N * sizeof(double)) != hipSuccess) return fail(c, GPUE_ENOMEM, "V table");
  HIPCHK(c, hipMemcpy(c->Vt, V.data(), c->N * sizeof(double), hipMemcpyHostToDevice));
  c->v_kind = V_TABLE;
  return GPUE_OK;
}

// ---------------------------------------------------------------------------
// ABI
// ---------------------------------------------------------------------------
extern "C" {

gpue_status gpue_params_default(gpue_params* p) {
  if (!p) return GPUE_EINVAL;
  std::memset(p, 0, sizeof(*p));
  // src/parser.cu:14-67
  p->dimnum = 2;
  p->xDim = p->yDim = p->zDim = 256;
  p->omega = 0.0;
  p->gammaY = 1.0;
  p->gdt = 1e-4;
  p->dt = 1e-4;
  p->omegaX = p->omegaY = p->omegaZ = 6.283;
  p->gpe = 0;
  p->corotating = 0;
  p->a_kind = GPUE_A_ROTATION;
  p->precision = 64;
  // src/init.cu:98-137 with atoms = 1, interaction = 1, box_size < 0
  const double HBAR = 1.05457148e-34, PI = 3.141592653589793;
  const double mass = 1.4431607e-25, a_s = 4.76e-9;
  const int atoms = 1;
  p->hbar = HBAR;
  p->mass = mass;
  double gDenConst = atoms * 4 * HBAR * HBAR * PI * (a_s / mass);
  gDenConst *= std::sqrt(mass * (p->omegaZ) / (2 * PI * HBAR));   // 2D
  p->g = 1.0 * gDenConst;
  double Rxy = std::pow(15, 0.2) * std::pow(atoms * a_s * std::sqrt(mass * p->omegaZ / HBAR), 0.2);
  double a0x = std::sqrt(HBAR / (2 * mass * p->omegaX));
  double a0y = std::sqrt(HBAR / (2 * mass * p->omegaY));
  double a0z = std::sqrt(HBAR / (2 * mass * p->omegaZ));
  p->xMax = 6 * Rxy * a0x;
  p->yMax = 6 * Rxy * a0y;
  p->zMax = 6 * Rxy * a0z;
  return GPUE_OK;
}

const char* gpue_last_error(const gpue_ctx* c) { return c ? c->err.c_str() : "null context"; }

void gpue_destroy(gpue_ctx* c) {
  if (!c) return;
  if (c->stream && c->own_stream) {
    (void)hipStreamSynchronize(c->stream);
  } else if (c->stream) {   // a shared stream: wait for the work queued so far, not for later work
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess) {
      if (hipEventRecord(ev, c->stream) == hipSuccess) (void)hipEventSynchronize(ev);
      (void)hipEventDestroy(ev);
    } else {
      (void)hipStreamSynchronize(c->stream);
    }
  }
  void* bufs[] = {c->psi, c->tmp, c->partials, c->results, c->Vt, c->Kt, c->At[0], c->At[1], c->At[2],
                  c->vs[0], c->vs[1], c->vs[2], c->ks[0], c->ks[1], c->ks[2], c->kf[0][0], c->kf[0][1],
                  c->kf[0][2], c->kf[1][0], c->kf[1][1], c->kf[1][2], c->vsf[0], c->vsf[1], c->vsf[2],
                  c->kff[0][0], c->kff[0][1], c->kff[0][2], c->kff[1][0], c->kff[1][1], c->kff[1][2],
                  c->vmask, c->vcand, c->vmarker, c->vfound};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (int a = 0; a < 6; ++a)
    if (c->dgrid[a]) (void)hipFree(c->dgrid[a]);
  for (int k = 0; k < 5; ++k)
    if (c->deq[k]) (void)hipFree(c->deq[k]);
  for (int a = 0; a < 3; ++a)
    if (c->vst[a]) (void)hipFree(c->vst[a]);
  for (auto& kv : c->tw) (void)hipFree(kv.second);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  drop_graphs(c);
  if (c->gev) (void)hipEventDestroy(c->gev);
  if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"

// Upload a host table into *slot (allocated on first use, then reused).
template <typename TT>
static hipError_t put_table(TT** slot, const std::vector<TT>& v) {
  if (!*slot) {
    hipError_t e = hipMalloc((void**)slot, std::max<size_t>(1, v.size()) * sizeof(TT));
    if (e != hipSuccess) return e;
  }
  return hipMemcpy(*slot, v.data(), v.size() * sizeof(TT), hipMemcpyHostToDevice);
}

// The tables that follow from the physics parameters (grids, separable V and K,
// the K operator factors and their FP32 copies), for a new context or one that
// is given new parameters of the same shape.
static gpue_status fill_tables(gpue_ctx* c) {
  const gpue_params& p = c->p;
  c->a_kind = p.a_kind == GPUE_A_ROTATION ? A_ROTATION : A_ZERO;
  make_grids(p.dimnum, p.xDim, p.yDim, p.zDim, p.xMax, p.yMax, p.zMax, c->grid, c->d);
  std::vector<double> vsh[3], ksh[3];
  sep_tables(c, vsh, ksh);
  for (int a = 0; a < 6; ++a)
    if (put_table(&c->dgrid[a], c->grid[a]) != hipSuccess) return fail(c, GPUE_ENOMEM, "grid arrays");
  for (int a = 0; a < 3; ++a) {
    if (put_table(&c->vs[a], vsh[a]) != hipSuccess) return fail(c, GPUE_ENOMEM, "V tables");
    if (put_table(&c->ks[a], ksh[a]) != hipSuccess) return fail(c, GPUE_ENOMEM, "K tables");
    // K operator factors: exp(-K_a gdt/hbar) and exp(-i K_a dt/hbar) (src/operators.cu:1082,1102-1103)
    std::vector<double> fg(2 * ksh[a].size()), fe(2 * ksh[a].size());
    for (size_t i = 0; i < ksh[a].size(); ++i) {
      fg[2 * i] = std::exp(-ksh[a][i] * (p.gdt / p.hbar));
      fg[2 * i + 1] = 0.0;
      fe[2 * i] = std::cos(-ksh[a][i] * (p.dt / p.hbar));
      fe[2 * i + 1] = std::sin(-ksh[a][i] * (p.dt / p.hbar));
    }
    if (put_table(&c->kf[0][a], fg) != hipSuccess || put_table(&c->kf[1][a], fe) != hipSuccess)
      return fail(c, GPUE_ENOMEM, "K factors");
    if (c->prec == 32) {   // rounded once, to nearest
      const std::vector<float> v32(vsh[a].begin(), vsh[a].end()), g32(fg.begin(), fg.end()), e32(fe.begin(), fe.end());
      if (put_table(&c->vsf[a], v32) != hipSuccess || put_table(&c->kff[0][a], g32) != hipSuccess ||
          put_table(&c->kff[1][a], e32) != hipSuccess)
        return fail(c, GPUE_ENOMEM, "FP32 tables");
    }
  }
  return GPUE_OK;
}

static gpue_status check_params(const gpue_params& p, std::string& msg) {
  if (p.dimnum < 1 || p.dimnum > 3) return msg = "dimnum must be 1, 2 or 3", GPUE_EINVAL;
  if (p.precision != 64 && p.precision != 32) return msg = "precision must be 64 or 32", GPUE_EINVAL;
  const int dims[3] = {p.xDim, p.dimnum > 1 ? p.yDim : 1, p.dimnum > 2 ? p.zDim : 1};
  for (int a = 0; a < p.dimnum; ++a)
    if (!valid_dim(dims[a]))
      return msg = "unsupported grid size " + std::to_string(dims[a]) +
                   " (supported: 16 32 48 64 96 128 192 256 384 512 1024 2048 4096)",
             GPUE_EUNSUPPORTED;
  if (!(p.hbar > 0) || !(p.mass > 0) || !(p.xMax > 0) || (p.dimnum > 1 && !(p.yMax > 0)) ||
      (p.dimnum > 2 && !(p.zMax > 0)))
    return msg = "hbar, mass and box sizes must be positive", GPUE_EINVAL;
  if (p.a_kind != GPUE_A_ROTATION && p.a_kind != GPUE_A_CONSTANT) return msg = "bad a_kind", GPUE_EINVAL;
  return GPUE_OK;
}

bool gpue::ctx_same_shape(const gpue_ctx* c, const gpue_params* p) {
  return p->dimnum == c->p.dimnum && p->precision == c->prec && p->device == c->p.device && p->xDim == c->dims[0] &&
         (p->dimnum < 2 || p->yDim == c->dims[1]) && (p->dimnum < 3 || p->zDim == c->dims[2]);
}

static void clear_program(gpue_ctx* c, int k);

// New parameters for a context of the same shape (grid sizes, precision,
// device): the tables are refilled in place and every field returns to its
// default, as for a new context; psi is left as it is.  No device memory is
// allocated (the serving loop reuses finished requests' contexts this way).
gpue_status gpue::ctx_reset(gpue_ctx* c, const gpue_params* pin) {
  std::string msg;
  gpue_status st = check_params(*pin, msg);
  if (st != GPUE_OK) return fail(c, st, msg);
  if (!ctx_same_shape(c, pin)) return fail(c, GPUE_EINVAL, "reset to another shape");
  drop_graphs(c);   // waits for this context's last replay only
  for (int k = 0; k < 5; ++k) clear_program(c, k);
  gpue_params p = *pin;
  if (p.dimnum < 3) p.zDim = 1;
  if (p.dimnum < 2) p.yDim = 1;
  c->p = p;
  c->v_kind = V_SEP;
  c->k_kind = K_SEP;
  c->v_user = false;
  c->az_zero = true;
  c->ramp_type = 0;
  c->pending_norm = false;
  c->defer_nb = 0;
  c->next_step = 0;
  c->last_energy = 0.0;
  c->t_last = 0.0;
  c->vmask_valid = false;
  st = fill_tables(c);
  if (st != GPUE_OK) return st;
  HIPCHK(c, launch_set_double(c->results, 1.0, c->stream));
  return GPUE_OK;
}

gpue_status gpue::ctx_create(const gpue_params* pin, hipStream_t ext_stream, gpue_ctx** out) {
  if (!pin || !out) return GPUE_EINVAL;
  *out = nullptr;
  gpue_ctx* c = new gpue_ctx();
  c->p = *pin;
  gpue_params& p = c->p;
  auto bail = [&](gpue_status s, const std::string& m) {
    c->err = m;
    std::fprintf(stderr, "gpue_create: %s\n", m.c_str());
    gpue_destroy(c);
    return s;
  };
  {
    std::string msg;
    const gpue_status st = check_params(p, msg);
    if (st != GPUE_OK) return bail(st, msg);
  }
  if (p.dimnum < 3) p.zDim = 1;
  if (p.dimnum < 2) p.yDim = 1;
  c->dims[0] = p.xDim;
  c->dims[1] = p.yDim;
  c->dims[2] = p.zDim;
  c->prec = p.precision;
  c->esz = p.precision == 64 ? 16 : 8;
  c->N = (long)p.xDim * p.yDim * p.zDim;

  hipError_t e = hipSetDevice(p.device);
  if (e != hipSuccess) return bail(GPUE_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  if (ext_stream) {
    c->stream = ext_stream;
    c->own_stream = false;
  } else {
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(GPUE_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  if (hipMalloc(&c->psi, c->N * c->esz) != hipSuccess || hipMalloc(&c->tmp, c->N * c->esz) != hipSuccess)
    return bail(GPUE_ENOMEM, "psi buffers");
  {
    const gpue_status st = fill_tables(c);
    if (st != GPUE_OK) return bail(st, c->err);
  }
  for (int a = 0; a < p.dimnum; ++a) {
    const int n = c->dims[a];
    std::vector<double> re, im;
    twiddles(n, re, im);
    for (int plan = 0; plan < 2; ++plan) {
      if (c->tw.count(2 * n + plan)) continue;
      std::vector<int> R;
      if (!plan_radices(n, plan, R)) return bail(GPUE_EUNSUPPORTED, "no plan for n");
      std::vector<int> idx;
      pass_table(n, R, idx);
      void* t = nullptr;
      if (hipMalloc(&t, std::max<size_t>(idx.size(), 1) * 16) != hipSuccess) return bail(GPUE_ENOMEM, "twiddles");
      c->tw[2 * n + plan] = t;
      if (c->prec == 64) {
        std::vector<double> il(2 * idx.size());
        for (size_t j = 0; j < idx.size(); ++j) { il[2 * j] = re[idx[j]]; il[2 * j + 1] = im[idx[j]]; }
        e = hipMemcpy(t, il.data(), il.size() * sizeof(double), hipMemcpyHostToDevice);
      } else {
        // split twiddles: (hi.re, hi.im, lo.re, lo.im), lo = the rounded remainder
        std::vector<float> il(4 * idx.size());
        for (size_t j = 0; j < idx.size(); ++j) {
          const double xr = re[idx[j]], xi = im[idx[j]];
          const float hr = (float)xr, hi = (float)xi;
          il[4 * j] = hr;
          il[4 * j + 1] = hi;
          il[4 * j + 2] = (float)(xr - (double)hr);
          il[4 * j + 3] = (float)(xi - (double)hi);
        }
        e = hipMemcpy(t, il.data(), il.size() * sizeof(float), hipMemcpyHostToDevice);
      }
      if (e != hipSuccess) return bail(GPUE_EHIP, "twiddle upload");
    }
  }
  // reduction scratch: the largest grid of any reducing kernel, times 4 accumulators
  size_t nb = REDUCE_BLOCKS;
  for (int v = 0; v < N_VARIANTS; ++v) {
    nb = std::max<size_t>(nb, sweep_blocks(c->prec, true, p.xDim, v, p.yDim * p.zDim, 1));
    if (p.dimnum > 1) nb = std::max<size_t>(nb, sweep_blocks(c->prec, false, p.yDim, v, p.xDim, p.zDim));
    if (p.dimnum > 2) nb = std::max<size_t>(nb, sweep_blocks(c->prec, false, p.zDim, v, p.xDim, p.yDim));
  }
  for (int a = 0; a < 3; ++a) c->var[a] = default_variant(c, a);
  c->partials_n = nb * 4;
  if (hipMalloc((void**)&c->partials, c->partials_n * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->results, 16 * sizeof(double)) != hipSuccess)
    return bail(GPUE_ENOMEM, "reduction scratch");
  // on the context's own stream: other contexts' work on the device is not waited for
  if (hipMemsetAsync(c->results, 0, 16 * sizeof(double), c->stream) != hipSuccess ||
      launch_set_double(c->results, 1.0, c->stream) != hipSuccess ||
      hipMemsetAsync(c->psi, 0, c->N * c->esz, c->stream) != hipSuccess)
    return bail(GPUE_EHIP, "memset");
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(GPUE_EHIP, "stream sync");
  *out = c;
  return GPUE_OK;
}

extern "C" {

gpue_status gpue_create(const gpue_params* pin, gpue_ctx** out) { return ctx_create(pin, nullptr, out); }

gpue_status gpue_get_grid(gpue_ctx* c, int axis, double* out) {
  if (!c || !out || axis < 0 || axis > 5) return GPUE_EINVAL;
  std::memcpy(out, c->grid[axis].data(), c->grid[axis].size() * sizeof(double));
  return GPUE_OK;
}

gpue_status gpue_set_g(gpue_ctx* c, double g) {
  if (!c) return GPUE_EINVAL;
  c->p.g = g;
  drop_graphs(c);
  return GPUE_OK;
}

gpue_status gpue_sync(gpue_ctx* c) {
  if (!c) return GPUE_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GPUE_OK;
}

// The current A components as tables (the rotation gauge's or zero), before one
// of them is replaced by a table or a time-dependent equation.
static gpue_status materialise_A(gpue_ctx* c, int skip) {
  if (c->a_kind == A_TABLE || c->a_kind == A_EQN) return GPUE_OK;
  for (int a = 0; a < 3; ++a) {
    if (a == skip) continue;
    if (!c->At[a] && hipMalloc((void**)&c->At[a], c->N * sizeof(double)) != hipSuccess)
      return fail(c, GPUE_ENOMEM, "field table");
    std::vector<double> h(c->N, 0.0);
    if (c->a_kind == A_ROTATION) {
      for (long gid = 0; gid < c->N; ++gid) {
        long xi = gid % c->dims[0], yi = (gid / c->dims[0]) % c->dims[1];
        if (a == 0) h[gid] = -c->grid[1][yi] * c->p.omega * c->p.omegaX;
        if (a == 1) h[gid] = c->grid[0][xi] * c->p.omega * c->p.omegaY;
      }
    }
    HIPCHK(c, hipMemcpy(c->At[a], h.data(), c->N * sizeof(double), hipMemcpyHostToDevice));
  }
  c->a_kind = A_TABLE;
  return GPUE_OK;
}

// Drop field k's time-dependent program (a static value replaces it).
static void clear_program(gpue_ctx* c, int k) {
  if (c->deq[k]) (void)hipFree(c->deq[k]);
  c->deq[k] = nullptr;
  c->t_dep = false;
  for (int j = 0; j < 5; ++j) c->t_dep = c->t_dep || c->deq[j] != nullptr;
  if (k >= 2 && !c->deq[2] && !c->deq[3] && !c->deq[4] && c->a_kind == A_EQN) c->a_kind = A_TABLE;
}

static gpue_status set_table(gpue_ctx* c, int which, const double* src, bool from_device) {
  drop_graphs(c);
  double** slot = which == GPUE_FIELD_V ? &c->Vt : which == GPUE_FIELD_K ? &c->Kt : &c->At[which - GPUE_FIELD_AX];
  if (!*slot) {
    if (hipMalloc((void**)slot, c->N * sizeof(double)) != hipSuccess) return fail(c, GPUE_ENOMEM, "field table");
  }
  HIPCHK(c, hipMemcpyAsync(*slot, src, c->N * sizeof(double), from_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  clear_program(c, which);
  if (which == GPUE_FIELD_V) {
    c->v_kind = V_TABLE;
    c->v_user = true;
  } else if (which == GPUE_FIELD_K) {
    c->k_kind = K_TABLE;
  } else {
    // the other components keep their current values: materialise them as tables
    gpue_status st = materialise_A(c, which - GPUE_FIELD_AX);
    if (st != GPUE_OK) return st;
    if (which == GPUE_FIELD_AZ) {
      std::vector<double> h(c->N);
      HIPCHK(c, hipMemcpy(h.data(), c->At[2], c->N * sizeof(double), hipMemcpyDeviceToHost));
      c->az_zero = std::all_of(h.begin(), h.end(), [](double v) { return v == 0.0; });
    }
    // GPUE builds V after A (src/operators.cu:726-829): a harmonic V follows the new A
    if (!c->v_user) {
      if (c->a_kind == A_EQN) {
        c->v_kind = V_TRAP_A;   // another component still depends on time
        return GPUE_OK;
      }
      return materialise_V(c);
    }
  }
  return GPUE_OK;
}

gpue_status gpue_set_field(gpue_ctx* c, int which, const double* host_N) {
  if (!c || !host_N || which < GPUE_FIELD_V || which > GPUE_FIELD_AZ) return GPUE_EINVAL;
  return set_table(c, which, host_N, false);
}

static std::map<std::string, double> name_map(const char* const* names, const double* values, int n) {
  std::map<std::string, double> m;
  for (int i = 0; i < n; ++i)
    if (names && names[i]) m[names[i]] = values[i];
  return m;
}

gpue_status gpue_set_field_eqn(gpue_ctx* c, int which, const char* expr, const char* const* names,
                               const double* values, int n_names) {
  if (!c || !expr || which < GPUE_FIELD_V || which > GPUE_FIELD_AZ) return GPUE_EINVAL;
  if (n_names < 0 || (n_names > 0 && (!names || !values))) return GPUE_EINVAL;
  std::vector<EqnNodeC> nodes;
  bool td = false;
  std::string err;
  if (eqn_parse(expr, name_map(names, values, n_names), {}, nodes, td, err)) return fail(c, GPUE_EPARSE, err);
  EqnProgH prog;
  if (eqn_to_program(nodes.data(), (int)nodes.size(), prog, err)) return fail(c, GPUE_EPARSE, err);
  static_assert(sizeof(EqnProgH) == sizeof(EqnProg), "program layouts differ");
  EqnProg dp;
  std::memcpy(&dp, &prog, sizeof(dp));
  const gpue_params& p = c->p;
  const double xM = p.xMax, yM = p.dimnum > 1 ? p.yMax : 0.0, zM = p.dimnum > 2 ? p.zMax : 0.0;
  if (td) {
    // evaluated per point inside the step's sweeps at the step's time (the EQ
    // kernels); the field's table is only filled for the energy
    if (c->prec != 64)
      return fail(c, GPUE_EUNSUPPORTED, "time-dependent equation fields need an FP64 context (GPUE's precision)");
    if (dp.depth > EQ_REG_DEPTH)
      return fail(c, GPUE_EUNSUPPORTED, "time-dependent equation needs a stack deeper than " +
                                            std::to_string(EQ_REG_DEPTH));
    drop_graphs(c);
    const int k = which;   // GPUE_FIELD_V, _K, _AX, _AY, _AZ = 0..4
    double** slot = which == GPUE_FIELD_V ? &c->Vt : which == GPUE_FIELD_K ? &c->Kt : &c->At[which - GPUE_FIELD_AX];
    if (!*slot && hipMalloc((void**)slot, c->N * sizeof(double)) != hipSuccess) return fail(c, GPUE_ENOMEM, "field");
    if (which >= GPUE_FIELD_AX) {
      gpue_status st = materialise_A(c, which - GPUE_FIELD_AX);
      if (st != GPUE_OK) return st;
    }
    if (!c->deq[k] && hipMalloc((void**)&c->deq[k], sizeof(EqnProg)) != hipSuccess)
      return fail(c, GPUE_ENOMEM, "equation program");
    HIPCHK(c, hipMemcpy(c->deq[k], &dp, sizeof(EqnProg), hipMemcpyHostToDevice));
    c->t_dep = true;
    if (which == GPUE_FIELD_V) {
      c->v_kind = V_EQN;
      c->v_user = true;
    } else if (which == GPUE_FIELD_K) {
      c->k_kind = K_EQN;
    } else {
      c->a_kind = A_EQN;
      if (which == GPUE_FIELD_AZ) c->az_zero = false;
      if (!c->v_user) {   // V = trap + 1/2 m |A(t)|^2, per point
        if (!c->Vt && hipMalloc((void**)&c->Vt, c->N * sizeof(double)) != hipSuccess)
          return fail(c, GPUE_ENOMEM, "V");
        for (int a = 0; a < 3 && !c->vst[a]; ++a) {
          const gpue_params& q = c->p;
          std::vector<double> t(c->grid[a].size());
          const double oZ = q.dimnum == 3 ? q.omegaZ : 0.0, z0 = q.dimnum == 3 ? q.z0_shift : 0.0;
          for (size_t i = 0; i < t.size(); ++i) {
            const double x = c->grid[a][i];
            const double u = a == 0 ? q.omegaX * (x + q.x0_shift)
                                    : a == 1 ? q.gammaY * q.omegaY * (x + q.y0_shift) : 1.0 * oZ * (x + z0);
            t[i] = 0.5 * q.mass * (u * u);
          }
          hipError_t e = hipSuccess;
          c->vst[a] = dev_copy(t, &e);
          if (e != hipSuccess) return fail(c, GPUE_ENOMEM, "trap tables");
        }
        c->v_kind = V_TRAP_A;
      }
    }
    return GPUE_OK;
  }
  // static: evaluate once on the device (find_field), then use as a table
  double* buf = nullptr;
  if (hipMalloc((void**)&buf, c->N * sizeof(double)) != hipSuccess) return fail(c, GPUE_ENOMEM, "eqn field");
  hipError_t e = launch_eqn_field(dp, buf, c->dims[0], c->dims[1], c->dims[2], c->d[0], c->d[1], c->d[2], xM, yM, zM,
                                  0.0, c->stream);
  if (e != hipSuccess) {
    (void)hipFree(buf);
    return fail(c, GPUE_EHIP, hipGetErrorString(e));
  }
  gpue_status st = set_table(c, which, buf, true);
  (void)hipFree(buf);
  return st;
}

gpue_status gpue_upload_wfc(gpue_ctx* c, const double* il) {
  if (!c || !il) return GPUE_EINVAL;
  if (c->prec == 64) {
    HIPCHK(c, hipMemcpyAsync(c->psi, il, c->N * 16, hipMemcpyHostToDevice, c->stream));
  } else {
    std::vector<float> f(2 * c->N);
    for (long i = 0; i < 2 * c->N; ++i) f[i] = (float)il[i];
    HIPCHK(c, hipMemcpyAsync(c->psi, f.data(), c->N * 8, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, launch_set_double(c->results, 1.0, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->pending_norm = false;
  return GPUE_OK;
}

gpue_status gpue_download_wfc(gpue_ctx* c, double* il) {
  if (!c || !il) return GPUE_EINVAL;
  gpue_status st = apply_pending(c);
  if (st != GPUE_OK) return st;
  if (c->prec == 64) {
    HIPCHK(c, hipMemcpyAsync(il, c->psi, c->N * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> f(2 * c->N);
    HIPCHK(c, hipMemcpyAsync(f.data(), c->psi, c->N * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (long i = 0; i < 2 * c->N; ++i) il[i] = f[i];
  }
  return GPUE_OK;
}

}  // extern "C"

gpue_status gpue::download_wfc_ev(gpue_ctx* c, double* il, hipEvent_t ev) {
  gpue_status st = apply_pending(c);
  if (st != GPUE_OK) return st;
  if (c->prec == 64) {
    HIPCHK(c, hipMemcpyAsync(il, c->psi, c->N * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipEventRecord(ev, c->stream));
    HIPCHK(c, hipEventSynchronize(ev));
  } else {
    std::vector<float> f(2 * c->N);
    HIPCHK(c, hipMemcpyAsync(f.data(), c->psi, c->N * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipEventRecord(ev, c->stream));
    HIPCHK(c, hipEventSynchronize(ev));
    for (long i = 0; i < 2 * c->N; ++i) il[i] = f[i];
  }
  return GPUE_OK;
}

extern "C" {

gpue_status gpue_parsum(gpue_ctx* c) {
  if (!c) return GPUE_EINVAL;
  gpue_status st = apply_pending(c);
  if (st != GPUE_OK) return st;
  const double dV = c->d[0] * c->d[1] * c->d[2];
  int nb = 0;
  hipError_t e = c->prec == 64 ? launch_norm_sum<double>(c->psi, c->N, c->partials, &nb, c->stream)
                               : launch_norm_sum<float>(c->psi, c->N, c->partials, &nb, c->stream);
  if (e == hipSuccess) e = launch_finalize(c->partials, nb, 1, 1, dV, c->results, c->stream);
  if (e != hipSuccess) return fail(c, GPUE_EHIP, hipGetErrorString(e));
  c->pending_norm = true;
  return apply_pending(c);
}

}  // extern "C"

// Queue the energy terms (kinetic by Parseval, position terms, GPUE's
// forward-difference gauge term) into results[1..5]; copy results[0..5] to
// host_dst (pinned memory for an asynchronous caller) when given.
// The fields the energy kernels read: time-dependent ones are filled into their
// tables at the time of the last step (one kernel per energy evaluation).
static gpue_status energy_phys(gpue_ctx* c, Phys* out, bool fill) {
  Phys ph = make_phys(c, 0, c->t_last);
  if (c->t_dep) {
    if (fill) {
      hipError_t e = launch_fields_at(ph, (ph.v_kind == V_EQN || ph.v_kind == V_TRAP_A) ? c->Vt : nullptr,
                                      c->deq[1] ? c->Kt : nullptr, c->deq[2] ? c->At[0] : nullptr,
                                      c->deq[3] ? c->At[1] : nullptr, c->deq[4] ? c->At[2] : nullptr, c->N, c->stream);
      if (e != hipSuccess) return fail(c, GPUE_EHIP, std::string("field tables: ") + hipGetErrorString(e));
    }
    if (ph.v_kind == V_EQN || ph.v_kind == V_TRAP_A) ph.v_kind = V_TABLE;
    if (ph.k_kind == K_EQN) ph.k_kind = K_TABLE;
    if (ph.a_kind == A_EQN) ph.a_kind = A_TABLE;
    ph.eqk = 0;
  }
  *out = ph;
  return GPUE_OK;
}

gpue_status gpue::energy_launch(gpue_ctx* c, double* host_dst) {
  const gpue_params& p = c->p;
  Phys ph;
  gpue_status st = energy_phys(c, &ph, true);
  if (st != GPUE_OK) return st;
  // A separable K (GPUE's simple_K, src/operators.cu:460-469) is summed one axis
  // at a time: by Parseval along the other axes, sum_k K_a(k_a) |psi_hat|^2 / N
  // equals sum over the mixed representation (transformed along a only) of
  // K_a(k_a) |psi_a|^2 / n_a.  Each axis is then one read-only sweep of psi, and
  // a pending renormalisation is divided out of the sums instead of applied.
  const bool sep = ph.k_kind == K_SEP;
  if (sep) {
    for (int a = 0; a < p.dimnum && st == GPUE_OK; ++a)
      st = launch(c, {a == 0, a, a == 0 ? (int)RM_KSUM : (int)CM_KSUM, F_KAX, 1.0 / (double)c->dims[a]}, ph, c->psi,
                  c->psi);
  } else {
    // a K table: Parseval over the full transform of the normalised psi
    st = apply_pending(c);
    const double invN = 1.0 / (double)c->N;
    if (st == GPUE_OK && p.dimnum == 1) {
      st = launch(c, {true, 0, RM_KSUM, 0u, invN}, ph, c->psi, c->psi);
    } else if (st == GPUE_OK) {
      st = launch(c, {true, 0, RM_FWD, 0u, 0.0}, ph, c->psi, c->tmp);
      if (st == GPUE_OK && p.dimnum == 3) st = launch(c, {false, 1, CM_FWD, 0u, 0.0}, ph, c->tmp, c->tmp);
      if (st == GPUE_OK) st = launch(c, {false, p.dimnum == 2 ? 1 : 2, CM_KSUM, 0u, invN}, ph, c->tmp, c->tmp);
    }
  }
  if (st != GPUE_OK) return st;
  const int gauge = p.corotating && p.dimnum > 1;
  int nb = 0;
  hipError_t e = c->prec == 64 ? launch_energy_pos<double>(c->psi, ph, gauge, p.hbar, c->d[0], c->d[1], c->d[2],
                                                           c->partials, &nb, c->stream)
                               : launch_energy_pos<float>(c->psi, ph, gauge, p.hbar, c->d[0], c->d[1], c->d[2],
                                                          c->partials, &nb, c->stream);
  if (e == hipSuccess) e = launch_finalize(c->partials, nb, 4, 0, 1.0, c->results + 2, c->stream);
  // results[1..5] of the normalised psi: kinetic from the per-axis sums, and
  // every term divided by the pending norm (results[0], 1.0 when none)
  if (e == hipSuccess) e = launch_energy_fold(c->results, sep ? p.dimnum : 0, c->stream);
  if (e == hipSuccess && host_dst)
    e = hipMemcpyAsync(host_dst, c->results, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return fail(c, GPUE_EHIP, hipGetErrorString(e));
  return GPUE_OK;
}

// energy_launch as one graph, captured at the first call and replayed after
// (about eight launches become one; gpue_batch_energy evaluates every instance
// at each convergence check). Taken only where the launches depend on no host
// state: a separable K, static fields, no ramp, no kernel timing, and the context's own
// stream (a capture there sees no other caller's work); drop_graphs discards
// the graph whenever parameters, fields or variants change.
gpue_status gpue::energy_launch_graph(gpue_ctx* c, double* host_dst) {
  if (!host_dst || c->t_dep || c->ramp_type || c->timing || !c->own_stream) return energy_launch(c, host_dst);
  Phys ph;
  gpue_status st = energy_phys(c, &ph, false);
  if (st != GPUE_OK) return st;
  if (ph.k_kind != K_SEP) return energy_launch(c, host_dst);
  if (c->egexec && c->egraph_dst != host_dst) drop_energy_graph(c);
  if (!c->egexec) {
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    st = energy_launch(c, host_dst);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (st != GPUE_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return st != GPUE_OK ? st : fail(c, GPUE_EHIP, std::string("energy graph capture: ") + hipGetErrorString(e));
    }
    if (hipGraphInstantiate(&c->egexec, g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(g);
      c->egexec = nullptr;
      return fail(c, GPUE_EHIP, "energy graph instantiate");
    }
    c->egraph = g;
    c->egraph_dst = host_dst;
  }
  HIPCHK(c, hipGraphLaunch(c->egexec, c->stream));
  return GPUE_OK;
}

// E from the terms as src/split_op.cu:208-375 (energy_calc) sums them, with
// g|psi|^4/2; the chemical potential counts the interaction term once more.
void gpue::energy_combine(const gpue_ctx* c, const double* r, double* energy, double* mu) {
  const gpue_params& p = c->p;
  const double dV = c->d[0] * c->d[1] * c->d[2];
  const double kin = r[1], vt = r[2], in4 = r[3], lz = r[4];
  if (energy) *energy = p.gpe ? (kin + vt + 0.5 * p.g * in4 + lz) * dV : (kin + vt + lz) * dV;
  if (mu) *mu = p.gpe ? (kin + vt + p.g * in4 + lz) * dV : (kin + vt + lz) * dV;
}

extern "C" {

gpue_status gpue_energy(gpue_ctx* c, double* energy, double* mu) {
  if (!c) return GPUE_EINVAL;
  double r[6];
  gpue_status st = energy_launch(c, nullptr);
  if (st != GPUE_OK) return st;
  HIPCHK(c, hipMemcpyAsync(r, c->results, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double E = 0, M = 0;
  energy_combine(c, r, &E, &M);
  if (energy) *energy = E;
  if (mu) *mu = M;
  if (!std::isfinite(E)) return fail(c, GPUE_ENONFINITE, "energy is not finite");
  return GPUE_OK;
}

// GPUE's energy with its gauge term (+i hbar A.grad by forward differences,
// src/split_op.cu:312-360) replaced by the term of the Hamiltonian the step
// evolves: apply_gauge multiplies by exp(-A_a k_a Dt) between 1-D transforms
// along axis a (src/operators.cu:1089-1110, src/evolution.cu:163-257), i.e.
// hbar A_a k_a.  One Parseval sweep per gauge axis: FFT along a, then
// sum A_a[gid] k_a |psi_hat|^2 / n_a over the mixed representation.
gpue_status gpue_energy_consistent(gpue_ctx* c, double* energy, double* mu) {
  if (!c) return GPUE_EINVAL;
  gpue_status st = energy_launch(c, nullptr);
  if (st != GPUE_OK) return st;
  const gpue_params& p = c->p;
  const bool gauge = p.corotating && p.dimnum > 1 && c->a_kind != A_ZERO;
  // Az of the rotation gauge is 0; a table Az only when set and not all zero
  const int naxes = !gauge ? 0 : (p.dimnum == 3 && (c->a_kind == A_TABLE || c->a_kind == A_EQN) && !c->az_zero) ? 3 : 2;
  Phys ph;
  st = energy_phys(c, &ph, false);   // the tables energy_launch filled
  if (st != GPUE_OK) return st;
  for (int a = 0; a < naxes && st == GPUE_OK; ++a)
    st = launch(c, {a == 0, a, a == 0 ? (int)RM_KSUM : (int)CM_KSUM, F_ASUM, 1.0 / (double)c->dims[a]}, ph, c->psi,
                c->psi);
  if (st != GPUE_OK) return st;
  double r[9];
  HIPCHK(c, hipMemcpyAsync(r, c->results, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (p.corotating && p.dimnum > 1) {
    double s = 0.0;
    for (int a = 0; a < naxes; ++a) s += r[6 + a];
    r[4] = p.hbar * s / (r[0] * r[0]);   // psi may still carry a pending norm
  }
  double E = 0, M = 0;
  energy_combine(c, r, &E, &M);
  if (energy) *energy = E;
  if (mu) *mu = M;
  if (!std::isfinite(E)) return fail(c, GPUE_ENONFINITE, "energy is not finite");
  return GPUE_OK;
}

// Replays of a graph of `reps` (copy psi -> scratch, then optionally the sweep
// in place on the scratch buffer) pairs; microseconds per replay.
static gpue_status time_graph(gpue_ctx* c, const Sweep* sw, const Phys& ph, int reps, hipEvent_t e0, hipEvent_t e1,
                              double* us) {
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  gpue_status st = GPUE_OK;
  hipError_t e = hipSuccess;
  for (int r = 0; r < reps && st == GPUE_OK && e == hipSuccess; ++r) {
    e = launch_copy16(c->psi, c->tmp, c->N * c->esz, c->stream);
    if (e == hipSuccess && sw) st = launch(c, *sw, ph, c->tmp, c->tmp);
  }
  hipGraph_t g = nullptr;
  hipError_t ec = hipStreamEndCapture(c->stream, &g);
  if (st != GPUE_OK || e != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return st != GPUE_OK ? st : fail(c, GPUE_EHIP, "timing capture");
  }
  hipGraphExec_t ge = nullptr;
  e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipGraphLaunch(ge, c->stream);   // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
  for (int r = 0; r < 2 && e == hipSuccess; ++r) e = hipGraphLaunch(ge, c->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (ge) (void)hipGraphExecDestroy(ge);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(c, GPUE_EHIP, std::string("timing: ") + hipGetErrorString(e));
  *us = ms * 1000.0 / 2.0;
  return GPUE_OK;
}

// Time one sweep as the step runs it, in place; returns microseconds per
// launch, including the dispatch gap between dependent launches.
//  * States smaller than the 256 MiB last-level cache (MALL), and real time:
//    each of `reps` launches is preceded by a copy of psi into the scratch
//    buffer, and the same graph without the sweeps is timed and subtracted.
//    A launch on its own output would find its lines in the caches it just
//    wrote (the L2 of its XCD, the MALL), which no step does (at 512^2 this
//    read 1.5 us low per sweep, for the 128 MiB FP32 256^3 state 18 % low),
//    and real-time phases of an unnormalised transform applied over and over
//    grow into the slow range of sincos.
//  * States of 256 MiB and more in imaginary time (no data-dependent path):
//    `reps` launches in place on one copy.  Here the subtraction over-counts
//    (a copy after a sweep runs slower than after a copy): measured against
//    rocprofv3 kernel durations of the FP64 256^3 job, in place is within
//    2-5 %, the subtraction 7-25 % above.
static gpue_status time_sweep(gpue_ctx* c, Sweep sw, const Phys& ph, int reps, hipEvent_t e0, hipEvent_t e1,
                              double* us) {
  sw.flags &= ~F_NORM;   // no reduction slots are written
  if (c->N * c->esz >= ((size_t)256 << 20) && ph.gstate == 0) {
    HIPCHK(c, hipMemcpyAsync(c->tmp, c->psi, c->N * c->esz, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    gpue_status st = GPUE_OK;
    for (int r = 0; r < reps && st == GPUE_OK; ++r) st = launch(c, sw, ph, c->tmp, c->tmp);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (st != GPUE_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return st != GPUE_OK ? st : fail(c, GPUE_EHIP, "timing capture");
    }
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipGraphLaunch(ge, c->stream);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    for (int r = 0; r < 2 && e == hipSuccess; ++r) e = hipGraphLaunch(ge, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (ge) (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(c, GPUE_EHIP, std::string("timing: ") + hipGetErrorString(e));
    *us = ms * 1000.0 / (2.0 * reps);
    return GPUE_OK;
  }
  double t_pair = 0.0, t_copy = 0.0;
  gpue_status st = time_graph(c, &sw, ph, reps, e0, e1, &t_pair);
  if (st == GPUE_OK) st = time_graph(c, nullptr, ph, reps, e0, e1, &t_copy);
  if (st != GPUE_OK) return st;
  *us = std::max(t_pair - t_copy, 0.0) / reps;
  return GPUE_OK;
}

static gpue_status evolve_impl(gpue_ctx* c, int first_step, int nsteps, int gstate, int energy_every,
                               double energy_thr, double* energies, int max_energies, gpue_evolve_result* res,
                               bool check_norm);

// Whole steps with the variants v, imaginary plus real time: graph blocks of
// steps run as evolve runs them, on the working copy the caller swapped in;
// microseconds per step (imaginary + real).
static gpue_status step_cost(gpue_ctx* c, const int* v, hipEvent_t e0, hipEvent_t e1, double* us) {
  for (int a = 0; a < 3; ++a) c->var[a] = v[a];
  const int n = 2 * c->graph_steps;
  double total = 0.0;
  for (int gstate = 0; gstate < 2; ++gstate) {
    drop_graphs(c);
    gpue_status st = evolve_impl(c, 0, c->graph_steps, gstate, 0, 0.0, nullptr, 0, nullptr, false);   // capture
    if (st != GPUE_OK) return st;
    // the best of at least two timings and of about 8 ms of steps
    double best = 0.0, spent = 0.0;
    for (int rep = 0; rep < 8 && (rep < 2 || spent < 8.0); ++rep) {
      HIPCHK(c, hipEventRecord(e0, c->stream));
      st = evolve_impl(c, 0, n, gstate, 0, 0.0, nullptr, 0, nullptr, false);
      if (st != GPUE_OK) return st;
      HIPCHK(c, hipEventRecord(e1, c->stream));
      HIPCHK(c, hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
      spent += ms;
      const double us_step = ms * 1000.0 / n;
      if (rep == 0 || us_step < best) best = us_step;
    }
    total += best;
  }
  drop_graphs(c);
  *us = total;
  return GPUE_OK;
}

// Second stage of the autotuner: the per-sweep times miss how sweeps of one
// step interact (which data the previous sweep left in the caches), so the
// per-axis choice is refined by one pass of coordinate descent over whole
// graph-replayed steps on a copy of the state (psi, the pending norm and the
// step counter are restored).  A change is kept when it saves over 0.5 %.
static gpue_status refine_steps(gpue_ctx* c, hipEvent_t e0, hipEvent_t e1) {
  if (c->graph_steps <= 0 || c->t_dep || c->ramp_type) return GPUE_OK;
  const int next_step = c->next_step;
  HIPCHK(c, hipMemcpyAsync(c->tmp, c->psi, c->N * c->esz, hipMemcpyDeviceToDevice, c->stream));
  std::swap(c->psi, c->tmp);
  int best[3] = {c->var[0], c->var[1], c->var[2]};
  double best_us = 0.0;
  gpue_status st = step_cost(c, best, e0, e1, &best_us);
  for (int a = 0; a < c->p.dimnum && st == GPUE_OK; ++a) {
    const int nv = a == 0 ? 4 : N_VARIANTS;   // row sweeps: variants 0..3
    for (int v = 0; v < nv && st == GPUE_OK; ++v) {
      if (v == best[a]) continue;
      int cand[3] = {best[0], best[1], best[2]};
      cand[a] = v;
      double us = 0.0;
      st = step_cost(c, cand, e0, e1, &us);
      if (st == GPUE_OK && us < 0.995 * best_us) {
        best_us = us;
        best[a] = v;
      }
    }
  }
  std::swap(c->psi, c->tmp);
  drop_graphs(c);   // captured on the copy (also after a failed trial)
  for (int a = 0; a < 3; ++a) c->var[a] = best[a];
  c->next_step = next_step;
  c->pending_norm = false;   // as at entry (the tuner settles the pending norm first)
  if (launch_set_double(c->results, 1.0, c->stream) != hipSuccess) return fail(c, GPUE_EHIP, "reset norm");
  return st;
}

// Autotuner: for every sweep of the imaginary- and real-time schedules, time
// each kernel variant (graph replays of back-to-back launches, so launch
// overhead does not decide) and keep, per axis, the variant with the smallest
// summed time; then refine over whole steps (refine_steps).
static gpue_status autotune(gpue_ctx* c, int reps) {
  if (reps <= 0) reps = 10;
  gpue_status st0 = apply_pending(c);
  if (st0 != GPUE_OK) return st0;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  double cost[3][N_VARIANTS] = {};
  std::vector<Sweep> sched;
  const bool timing = c->timing;
  c->timing = false;
  gpue_status st = GPUE_OK;
  for (int gstate = 0; gstate < 2 && st == GPUE_OK; ++gstate) {
    for (int parity = 0; parity < 2 && st == GPUE_OK; ++parity) {
      schedule(c, gstate, parity, true, sched);
      Phys ph = make_phys(c, gstate, 0.0);
      for (Sweep s : sched) {
        for (int v = 0; v < N_VARIANTS && st == GPUE_OK; ++v) {
          if (s.row && v > 3) break;   // row sweeps: plan and twiddle loading only
          s.var = v;
          double us = 0.0;
          st = time_sweep(c, s, ph, reps, e0, e1, &us);
          cost[s.axis][v] += us;
        }
      }
    }
  }
  c->timing = timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != GPUE_OK) return st;
  for (int a = 0; a < 3; ++a) {
    int best = c->var[a];
    for (int v = 0; v < N_VARIANTS; ++v)
      if (cost[a][v] > 0 && cost[a][v] < cost[a][best]) best = v;
    c->var[a] = best;
  }
  drop_graphs(c);
  c->timing = false;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  st = refine_steps(c, e0, e1);
  c->timing = timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != GPUE_OK) return st;
  drop_graphs(c);
  c->tuned = true;
  return GPUE_OK;
}

}  // extern "C"

// Sweeps of steps first..first+n-1 with, in real time, the sweep that ends a
// step fused with the one that starts the next when both run along the same
// axis (no norm stands between them, and the fields do not change): the pair's
// arithmetic is unchanged, one read and one write of psi are saved.  With a
// gauge this holds at every step boundary (GPUE's gauge order alternates, so
// a step ends on the axis the next one leads with): 2D real-time steps take 3
// sweeps.
void gpue::schedule_block(const gpue_ctx* c, int gstate, int first, int n, std::vector<Sweep>& out,
                          std::vector<int>& step_of) {
  std::vector<Sweep> one;
  out.clear();
  step_of.clear();
  for (int s = 0; s < n; ++s) {
    schedule(c, gstate, first + s, gstate == 0, one);
    for (const Sweep& sw : one) {
      out.push_back(sw);
      step_of.push_back(first + s);
    }
  }
  if (gstate != 1 || c->t_dep || c->ramp_type) return;
  std::vector<Sweep> fused;
  std::vector<int> fstep;
  for (size_t k = 0; k < out.size(); ++k) {
    const Sweep& a = out[k];
    if (k + 1 < out.size() && step_of[k + 1] == step_of[k] + 1 && opens_step(out[k + 1]) &&
        out[k + 1].row == a.row && out[k + 1].axis == a.axis && !(out[k + 1].flags & F_SCALE_IN)) {
      int mode = -1;
      if (a.row && a.mode == RM_INV) mode = RM_INV_FWD;
      if (a.row && a.mode == RM_GAUGE) mode = RM_GAUGE_FWD;
      if (!a.row && a.mode == CM_GAUGE) mode = CM_GAUGE_FWD;
      if (mode >= 0) {
        Sweep m = a;
        m.mode = mode;
        m.flags |= out[k + 1].flags & F_HALFV_PRE;
        fused.push_back(m);
        fstep.push_back(step_of[k]);
        ++k;
        continue;
      }
    }
    fused.push_back(a);
    fstep.push_back(step_of[k]);
  }
  out.swap(fused);
  step_of.swap(fstep);
}

extern "C" {

// Capture `len` steady-state steps (starting on an even step, norm pending in
// imaginary time) into a graph for this gstate; returns the cache entry.
static gpue_status build_graph(gpue_ctx* c, int gstate, int len, StepGraph** out) {
  std::vector<Sweep> sched;
  std::vector<int> step_of;
  const double Dt = gstate == 0 ? c->p.gdt : c->p.dt;
  schedule_block(c, gstate, 0, len, sched, step_of);
  if (gstate == 0)
    for (size_t k = 0; k + 1 < sched.size(); ++k)
      if (step_of[k + 1] == step_of[k] + 1) defer_norm(c, sched[k], sched[k + 1]);
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  gpue_status st = GPUE_OK;
  for (size_t k = 0; k < sched.size() && st == GPUE_OK; ++k) {
    const Phys ph = make_phys(c, gstate, Dt * step_of[k]);
    st = launch(c, sched[k], ph, c->psi, c->psi);
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (st != GPUE_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return fail(c, GPUE_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  hipGraphExec_t ge = nullptr;
  e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(c, GPUE_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  }
  if ((int)c->graphs.size() >= MAX_STEP_GRAPHS) {   // evict the least recently used block
    size_t lru = 0;
    for (size_t k = 1; k < c->graphs.size(); ++k)
      if (c->graphs[k].last_use < c->graphs[lru].last_use) lru = k;
    // its last replay may still be queued: wait for the latest replay of any block
    if (c->gev) HIPCHK(c, hipEventSynchronize(c->gev));
    (void)hipGraphExecDestroy(c->graphs[lru].exec);
    (void)hipGraphDestroy(c->graphs[lru].graph);
    c->graphs.erase(c->graphs.begin() + lru);
  }
  StepGraph sg;
  sg.gstate = gstate;
  sg.len = len;
  sg.graph = g;
  sg.exec = ge;
  c->graphs.push_back(sg);
  *out = &c->graphs.back();
  return GPUE_OK;
}

static gpue_status step_graph(gpue_ctx* c, int gstate, int len, StepGraph** out) {
  for (StepGraph& g : c->graphs)
    if (g.gstate == gstate && g.len == len) {
      *out = &g;
      return GPUE_OK;
    }
  return build_graph(c, gstate, len, out);
}

// Steps of a captured block starting at step s of an evolve call (0: the step
// is launched on its own): up to G steps (an even count, from an even step) when
// the fields do not depend on time, ending before the next energy evaluation and
// at the end of the call.
static int block_len(const gpue_ctx* c, int first_step, int nsteps, int s, int energy_every) {
  const int G = c->graph_steps;
  const int i = first_step + s;
  if (!(G > 0 && !c->timing && !c->t_dep && !c->ramp_type && (i % 2) == 0)) return 0;
  int L = std::min(G, nsteps - s);
  if (energy_every > 0) {
    const int nxt = ((i + energy_every - 1) / energy_every) * energy_every;   // first cadence step >= i
    L = std::min(L, nxt - i);
  }
  L &= ~1;
  return L >= 2 ? L : 0;
}

static gpue_status evolve_impl(gpue_ctx* c, int first_step, int nsteps, int gstate, int energy_every,
                               double energy_thr, double* energies, int max_energies, gpue_evolve_result* res,
                               bool check_norm) {
  gpue_evolve_result r;
  std::memset(&r, 0, sizeof(r));
  // real time starts from the renormalised state: settle a norm left pending
  // by imaginary time (GPUE divides at the end of every imaginary step)
  if (gstate == 1) {
    gpue_status st = apply_pending(c);
    if (st != GPUE_OK) return st;
  }
  std::vector<Sweep> sched;
  const double Dt = gstate == 0 ? c->p.gdt : c->p.dt;
  int s;
  gpue_status st0 = settle_deferred(c);
  if (st0 != GPUE_OK) return st0;
  // steady state: captured blocks of steps (block_len) are replayed; the
  // imaginary-time graph always rescales on entry (results[0] holds the
  // pending norm or 1.0, and x*1.0 is exact)
  for (s = 0; s < nsteps; ++s) {
    const int i = first_step + s;
    const int L = block_len(c, first_step, nsteps, s, energy_every);
    if (L > 0) {
      StepGraph* sg = nullptr;
      gpue_status st = step_graph(c, gstate, L, &sg);
      if (st != GPUE_OK) return st;
      sg->last_use = ++c->graph_clock;
      HIPCHK(c, hipGraphLaunch(sg->exec, c->stream));
      if (!c->gev) HIPCHK(c, hipEventCreateWithFlags(&c->gev, hipEventDisableTiming));
      HIPCHK(c, hipEventRecord(c->gev, c->stream));
      if (gstate == 0) c->pending_norm = true;
      s += L - 1;
      continue;
    }
    Phys ph = make_phys(c, gstate, Dt * i);   // time = Dt*i (src/evolution.cu:408)
    if (c->ramp_type) ph.Dt = Dt * ramp_scale(c, i);
    c->t_last = Dt * i;
    schedule(c, gstate, i, c->pending_norm, sched);
    if (c->defer_nb > 0) {   // the previous step left its norm partials to this one
      if (takes_deferred_norm(c, sched[0]) && (sched[0].flags & F_SCALE_IN)) {
        sched[0].flags |= F_SCALE_PART;
        sched[0].part_nb = c->defer_nb;
        c->defer_nb = 0;
      } else {
        gpue_status st = settle_deferred(c);
        if (st != GPUE_OK) return st;
      }
    }
    // leave this step's norm to the next one when that is an eager step too
    if (gstate == 0 && c->p.dimnum >= 2 && s + 1 < nsteps && block_len(c, first_step, nsteps, s + 1, energy_every) == 0 &&
        !(energy_every > 0 && i % energy_every == 0) && (sched.back().flags & F_NORM)) {
      const int nb = sweep_nb(c, sched.back());
      if (nb < 4096) {
        sched.back().flags |= F_DEFER;
        c->defer_nb = nb;
      }
    }
    for (const Sweep& sw : sched) {
      gpue_status st = launch(c, sw, ph, c->psi, c->psi);
      if (st != GPUE_OK) return st;
    }
    if (gstate == 0) c->pending_norm = true;
    if (energy_every > 0 && i % energy_every == 0) {
      double E = 0, M = 0;
      gpue_status st = gpue_energy(c, &E, &M);
      if (st != GPUE_OK) return st;
      if (energies && r.n_energies < max_energies) energies[r.n_energies] = E;
      r.n_energies++;
      r.energy = E;
      r.mu = M;
      const double old = (i != 0) ? c->last_energy : 0.0;
      c->last_energy = E;
      if (i != 0 && std::fabs(old - E) < energy_thr * old && gstate == 0) {   // src/evolution.cu:899
        r.early_stop = 1;
        ++s;
        break;
      }
    }
  }
  {
    gpue_status st = settle_deferred(c);   // (no step defers past the last one)
    if (st != GPUE_OK) return st;
  }
  r.steps_done = s;
  c->next_step = first_step + s;
  if (c->timing) {
    gpue_status st = timing_flush(c);
    if (st != GPUE_OK) return st;
  }
  if (check_norm && gstate == 0 && c->pending_norm) {
    double nrm = 0;
    HIPCHK(c, hipMemcpyAsync(&nrm, c->results, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!std::isfinite(nrm) || nrm == 0.0) return fail(c, GPUE_ENONFINITE, "norm is not finite");
  }
  if (res) *res = r;
  return GPUE_OK;
}

gpue_status gpue_evolve(gpue_ctx* c, int first_step, int nsteps, int gstate, int energy_every, double energy_thr,
                        double* energies, int max_energies, gpue_evolve_result* res) {
  if (!c || nsteps < 0 || (gstate != 0 && gstate != 1) || first_step < 0) return GPUE_EINVAL;
  return evolve_impl(c, first_step, nsteps, gstate, energy_every, energy_thr, energies, max_energies, res, true);
}


// Capture, without launching anything, every step block an evolve call with
// these arguments replays, so the call itself spends no time on captures.
gpue_status gpue_prepare_evolve(gpue_ctx* c, int first_step, int nsteps, int gstate, int energy_every) {
  if (!c || nsteps < 0 || (gstate != 0 && gstate != 1) || first_step < 0) return GPUE_EINVAL;
  for (int s = 0; s < nsteps;) {
    const int L = block_len(c, first_step, nsteps, s, energy_every);
    if (L == 0) {
      ++s;
      continue;
    }
    StepGraph* sg = nullptr;
    gpue_status st = step_graph(c, gstate, L, &sg);
    if (st != GPUE_OK) return st;
    sg->last_use = ++c->graph_clock;
    s += L;
  }
  return GPUE_OK;
}

}  // extern "C"

// Queue nsteps steps without waiting for them (no energy cadence, no norm check).
gpue_status gpue::evolve_async(gpue_ctx* c, int first_step, int nsteps, int gstate) {
  if (!c || nsteps < 0 || (gstate != 0 && gstate != 1) || first_step < 0) return GPUE_EINVAL;
  return evolve_impl(c, first_step, nsteps, gstate, 0, 0.0, nullptr, 0, nullptr, false);
}

extern "C" {

gpue_status gpue_vortex_marker_2d(gpue_ctx* c, double radius, int* marker_out, int* found) {
  if (!c || !marker_out) return GPUE_EINVAL;
  if (c->p.dimnum != 2 || c->dims[0] != c->dims[1])
    return fail(c, GPUE_EUNSUPPORTED, "findVortex needs a square 2D grid (src/tracker.cu:111-113)");
  gpue_status st = apply_pending(c);
  if (st != GPUE_OK) return st;
  const int xDim = c->dims[0];
  const size_t np = (size_t)(xDim - 1) * (xDim - 1);
  if (!c->vmask || !c->vcand || !c->vmarker || !c->vfound) {   // all four, or none after a failure
    if ((!c->vmask && hipMalloc((void**)&c->vmask, np) != hipSuccess) ||
        (!c->vcand && hipMalloc((void**)&c->vcand, np * sizeof(int)) != hipSuccess) ||
        (!c->vmarker && hipMalloc((void**)&c->vmarker, (size_t)xDim * xDim * sizeof(int)) != hipSuccess) ||
        (!c->vfound && hipMalloc((void**)&c->vfound, 64) != hipSuccess)) {
      c->vmask_valid = false;
      return fail(c, GPUE_ENOMEM, "vortex scratch");
    }
  }
  if (!c->vmask_valid || radius != c->vmask_radius) {
    std::vector<unsigned char> m;
    vortex_mask(c->grid[0].data(), xDim, radius, m);
    c->vmask_valid = false;
    HIPCHK(c, hipMemcpy(c->vmask, m.data(), np, hipMemcpyHostToDevice));
    c->vmask_radius = radius;
    c->vmask_valid = true;
  }
  hipError_t e = c->prec == 64 ? launch_vortex<double>(c->psi, c->vmask, xDim, c->vcand, c->vmarker, c->vfound, c->stream)
                               : launch_vortex<float>(c->psi, c->vmask, xDim, c->vcand, c->vmarker, c->vfound, c->stream);
  if (e != hipSuccess) return fail(c, GPUE_EHIP, hipGetErrorString(e));
  int f = 0;
  HIPCHK(c, hipMemcpyAsync(marker_out, c->vmarker, (size_t)xDim * xDim * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&f, c->vfound, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (found) *found = f;
  return GPUE_OK;
}

gpue_status gpue_vortex_list_2d(gpue_ctx* c, double radius, int* ijw, int max_v, int* n_v) {
  if (!c) return GPUE_EINVAL;
  if (c->p.dimnum != 2) return fail(c, GPUE_EUNSUPPORTED, "2D only");
  const int xDim = c->dims[0];
  std::vector<int> m((size_t)xDim * xDim);
  int found = 0;
  gpue_status st = gpue_vortex_marker_2d(c, radius, m.data(), &found);
  if (st != GPUE_OK) return st;
  int n = 0;   // vortPos (src/tracker.cu:212-229): raster order, |marker| >= 1
  for (int i = 0; i < xDim; ++i)
    for (int j = 0; j < xDim; ++j) {
      int v = m[(size_t)i * xDim + j];
      if (std::abs(v) >= 1) {
        if (ijw && n < max_v) {
          ijw[3 * n] = i;
          ijw[3 * n + 1] = j;
          ijw[3 * n + 2] = v;
        }
        ++n;
      }
    }
  if (n_v) *n_v = n;
  return GPUE_OK;
}

gpue_status gpue_set_timing(gpue_ctx* c, int on) {
  if (!c) return GPUE_EINVAL;
  if (on && c->ev.empty()) {   // event pairs are created on first use
    c->ev_kind.resize(TIMING_PAIRS);
    c->ev.assign(2 * TIMING_PAIRS, nullptr);
    for (auto& ev : c->ev) HIPCHK(c, hipEventCreate(&ev));
  }
  c->timing = on != 0;
  c->ev_used = 0;
  c->kind_names.clear();
  c->kind_n.clear();
  c->kind_ms.clear();
  return GPUE_OK;
}

gpue_status gpue_timing_get(gpue_ctx* c, int kind, const char** name, long* launches, double* ms) {
  if (!c || kind < 0) return GPUE_EINVAL;
  if (kind >= (int)c->kind_names.size()) return GPUE_EINVAL;
  if (name) *name = c->kind_names[kind].c_str();
  if (launches) *launches = c->kind_n[kind];
  if (ms) *ms = c->kind_ms[kind];
  return GPUE_OK;
}

// Per-kernel cost: every distinct sweep of the gstate schedules (both parities)
// is timed by time_sweep (graph replays in place on copies of psi in the
// scratch buffer; the state is untouched, the norm partials are not written).
gpue_status gpue_sweep_profile(gpue_ctx* c, int gstate, int reps, int max_kinds, int* n_kinds, const char** names,
                               double* us_per_launch, double* launches_per_step) {
  if (!c || (gstate != 0 && gstate != 1) || reps <= 0 || max_kinds < 0 || !n_kinds) return GPUE_EINVAL;
  gpue_status st = apply_pending(c);
  if (st != GPUE_OK) return st;
  std::vector<Sweep> sched, kinds;
  std::vector<int> step_of;
  std::vector<double> count;
  c->prof_names.clear();
  // the sweeps of an even/odd step pair as the step loop launches them (with
  // the real-time step boundary fused when steps are replayed from graphs)
  if (c->graph_steps > 0) {
    schedule_block(c, gstate, 0, 2, sched, step_of);
  } else {
    std::vector<Sweep> one;
    sched.clear();
    for (int par = 0; par < 2; ++par) {
      schedule(c, gstate, par, gstate == 0, one);
      sched.insert(sched.end(), one.begin(), one.end());
    }
  }
  {
    for (const Sweep& sw : sched) {
      const std::string nm = sweep_name(sw);
      size_t k = 0;
      while (k < c->prof_names.size() && c->prof_names[k] != nm) ++k;
      if (k == c->prof_names.size()) {
        c->prof_names.push_back(nm);
        kinds.push_back(sw);
        count.push_back(0.0);
      }
      count[k] += 0.5;
    }
  }
  const bool timing = c->timing;
  c->timing = false;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  const Phys ph = make_phys(c, gstate, 0.0);
  *n_kinds = (int)kinds.size();
  for (size_t k = 0; k < kinds.size() && st == GPUE_OK; ++k) {
    double us = 0.0;
    st = time_sweep(c, kinds[k], ph, reps, e0, e1, &us);
    if (st == GPUE_OK && (int)k < max_kinds) {
      if (names) names[k] = c->prof_names[k].c_str();
      if (us_per_launch) us_per_launch[k] = us;
      if (launches_per_step) launches_per_step[k] = count[k];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->timing = timing;
  return st;
}

gpue_status gpue_math_eval(int device, int fn, const double* x, double* y, long n) {
  if (!x || !y || n < 0 || fn < 0 || fn > 7) return GPUE_EINVAL;
  if (n == 0) return GPUE_OK;
  if (hipSetDevice(device) != hipSuccess) return GPUE_EHIP;
  double* d = nullptr;
  if (hipMalloc((void**)&d, 2 * n * sizeof(double)) != hipSuccess) return GPUE_ENOMEM;
  hipError_t e = hipMemcpy(d, x, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_math_eval(fn, d, d + n, n, nullptr);
  if (e == hipSuccess) e = hipMemcpy(y, d + n, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? GPUE_OK : GPUE_EHIP;
}

gpue_status gpue_step_bytes(gpue_ctx* c, int gstate, double* bytes, double* sweeps) {
  if (!c || (gstate != 0 && gstate != 1)) return GPUE_EINVAL;
  // the sweeps the steady state runs: a captured block when graphs are on (real
  // time fuses the sweeps at step boundaries), else a step of each parity
  const int G = c->graph_steps > 0 && !c->t_dep && !c->ramp_type ? c->graph_steps : 2;
  std::vector<Sweep> blk;
  std::vector<int> step_of;
  if (G == c->graph_steps) {
    schedule_block(c, gstate, 0, G, blk, step_of);
  } else {
    std::vector<Sweep> a, b;
    schedule(c, gstate, 0, false, a);
    schedule(c, gstate, 1, false, b);
    blk.resize(a.size() + b.size());
  }
  const double P = (double)blk.size() / (double)G;
  if (sweeps) *sweeps = P;
  if (bytes) *bytes = 2.0 * (double)c->esz * (double)c->N * P;
  return GPUE_OK;
}

gpue_status gpue_eqn_parse(const char* expr, const char* const* names, const double* values, int n_names,
                           gpue_eqn_node* nodes, int max_nodes, int* n_nodes, int* time_dependent) {
  if (!expr || (nodes && max_nodes < 0)) return GPUE_EINVAL;
  if (n_names < 0 || (n_names > 0 && (!names || !values))) return GPUE_EINVAL;
  std::vector<EqnNodeC> out;
  bool td = false;
  std::string err;
  if (eqn_parse(expr, name_map(names, values, n_names), {}, out, td, err)) {
    std::fprintf(stderr, "gpue_eqn_parse: %s\n", err.c_str());
    return GPUE_EPARSE;
  }
  static_assert(sizeof(gpue_eqn_node) == sizeof(EqnNodeC), "node layouts differ");
  if (n_nodes) *n_nodes = (int)out.size();
  if (time_dependent) *time_dependent = td ? 1 : 0;
  if (nodes) std::memcpy(nodes, out.data(), std::min<size_t>(max_nodes, out.size()) * sizeof(EqnNodeC));
  return (int)out.size() > max_nodes && nodes ? GPUE_EINVAL : GPUE_OK;
}

double gpue_eqn_eval(const gpue_eqn_node* nodes, int n, double x, double y, double z, double t) {
  return eqn_eval(reinterpret_cast<const EqnNodeC*>(nodes), n, x, y, z, t);
}

gpue_status gpue_set_ramp(gpue_ctx* c, int type, int num_steps) {
  if (!c || type < 0 || type > 2 || (type == 2 && num_steps <= 0)) return GPUE_EINVAL;
  c->ramp_type = type;
  c->ramp_steps = num_steps;
  drop_graphs(c);
  return GPUE_OK;
}

gpue_status gpue_set_graph_steps(gpue_ctx* c, int steps) {
  if (!c || steps < 0 || (steps % 2) != 0) return GPUE_EINVAL;
  c->graph_steps = steps;
  drop_graphs(c);
  return GPUE_OK;
}

gpue_status gpue_autotune(gpue_ctx* c, int reps, int* variants) {
  if (!c) return GPUE_EINVAL;
  gpue_status st = autotune(c, reps);
  if (variants)
    for (int a = 0; a < 3; ++a) variants[a] = c->var[a];
  return st;
}

gpue_status gpue_set_variants(gpue_ctx* c, const int* variants) {
  if (!c || !variants) return GPUE_EINVAL;
  for (int a = 0; a < 3; ++a) {
    if (variants[a] < 0 || variants[a] >= N_VARIANTS) return fail(c, GPUE_EINVAL, "variant out of range");
  }
  for (int a = 0; a < 3; ++a) c->var[a] = variants[a];
  drop_graphs(c);
  return GPUE_OK;
}

gpue_status gpue_get_variants(const gpue_ctx* c, int* variants) {
  if (!c || !variants) return GPUE_EINVAL;
  for (int a = 0; a < 3; ++a) variants[a] = c->var[a];
  return GPUE_OK;
}

}  // extern "C"
