// ctx_internal.h -- the context behind the opaque gpue_ctx handle, shared by
// the single-problem ABI (gpue_ctx.hip) and the batch runtime (gpue_batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "gpue_amd.h"
#include "misc_kernels.h"

namespace gpue {
struct Sweep {
  bool row;
  int axis;
  int mode;
  unsigned flags;
  double invN;
  int var = -1;   // kernel variant; -1: the context's choice for this axis
  int part_nb = 0;   // F_SCALE_PART: partials left by the previous sweep
};

constexpr int TIMING_PAIRS = 256;

// A captured block of `len` steps of one gstate (replayed from an even step).
struct StepGraph {
  int gstate = 0, len = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  long last_use = 0;
};
constexpr int MAX_STEP_GRAPHS = 8;

}  // namespace gpue

struct gpue_ctx {
  gpue_params p;
  int prec = 64;
  size_t esz = 16;   // bytes per complex element
  long N = 0;
  int dims[3] = {1, 1, 1};
  double d[3] = {1, 1, 1};
  std::vector<double> grid[6];
  hipStream_t stream = nullptr;
  bool own_stream = true;      // false: a stream the serving loop shares among its requests
  void* psi = nullptr;
  void* tmp = nullptr;
  double* dgrid[6] = {};
  std::map<int, void*> tw;   // 2*N + plan -> pass twiddle table (TwTab) in the working precision
  int var[3] = {0, 0, 0};      // kernel variant per axis (row sweeps: axis 0)
  bool tuned = false;
  // steady-state step graphs: blocks of up to graph_steps consecutive steps (even
  // count, starting at an even step, norm pending) captured once per (gstate,
  // length) and replayed; shorter blocks cover the tail of an evolve call and the
  // steps between energy evaluations
  int graph_steps = 8;
  std::vector<gpue::StepGraph> graphs;
  hipEvent_t gev = nullptr;    // recorded after every graph replay (a shared stream is not waited for)
  // the energy evaluation of gpue_batch_energy as one graph (energy_launch_graph),
  // with the host buffer its terms are copied to; dropped with the step graphs
  hipGraph_t egraph = nullptr;
  hipGraphExec_t egexec = nullptr;
  double* egraph_dst = nullptr;
  long graph_clock = 0;        // use counter for the least-recently-used eviction
  double* partials = nullptr;
  size_t partials_n = 0;
  double* results = nullptr;   // [0] norm  [1] kinetic  [2..5] position terms
  int v_kind = gpue::V_SEP, a_kind = gpue::A_ROTATION, k_kind = gpue::K_SEP;
  bool v_user = false;         // V supplied by the caller (table or equation)
  double* vs[3] = {};          // separable V / K value tables
  double* ks[3] = {};
  double* kf[2][3] = {};       // per-axis K operator factors for gstate 0 / 1
  float* vsf[3] = {};          // FP32 contexts: vs and kf rounded to FP32
  float* kff[2][3] = {};
  double* Vt = nullptr;
  double* Kt = nullptr;
  double* At[3] = {};
  bool az_zero = true;
  // time-dependent equation fields (V, K, Ax, Ay, Az): device programs evaluated
  // per point inside the sweeps; tables of the same fields are filled at time
  // t_last only for the energy
  gpue::EqnProg* deq[5] = {};
  bool t_dep = false;          // some field depends on time (steps run one by one)
  double t_last = 0.0;         // time of the last step, Dt * i
  double* vst[3] = {};         // the harmonic trap's per-axis terms (V_TRAP_A)
  int ramp_type = 0;           // -R: gauge exponent scaled per step (0 off, 1, 2)
  int ramp_steps = 0;          // numSteps of the ramped evolve call
  bool pending_norm = false;   // results[0] holds a norm not yet applied (else it holds 1.0)
  int next_step = 0;           // step index after the last evolve (batch parity)
  int defer_nb = 0;            // eager steps: a finalize left to the next step's first sweep
  double last_energy = 0.0;
  std::string err;
  // vortex scratch
  unsigned char* vmask = nullptr;
  double vmask_radius = 0.0;
  bool vmask_valid = false;   // vmask holds the mask of vmask_radius
  int* vcand = nullptr;
  int* vmarker = nullptr;
  int* vfound = nullptr;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_kind;
  int ev_used = 0;
  std::vector<std::string> kind_names;
  std::vector<long> kind_n;
  std::vector<double> kind_ms;
  std::vector<std::string> prof_names;   // names handed out by gpue_sweep_profile
};


namespace gpue {
gpue_status fail(gpue_ctx* c, gpue_status s, const std::string& msg);
Phys make_phys(const gpue_ctx* c, int gstate, double time);
double ramp_scale(const gpue_ctx* c, int step);
void schedule(const gpue_ctx* c, int gstate, int i, bool pending, std::vector<Sweep>& out);
void schedule_block(const gpue_ctx* c, int gstate, int first, int n, std::vector<Sweep>& out,
                    std::vector<int>& step_of);
gpue_status launch(gpue_ctx* c, const Sweep& s, const Phys& ph, const void* in, void* out);
gpue_status apply_pending(gpue_ctx* c);
gpue_status evolve_async(gpue_ctx* c, int first_step, int nsteps, int gstate);
gpue_status energy_launch(gpue_ctx* c, double* host_dst);
gpue_status energy_launch_graph(gpue_ctx* c, double* host_dst);
// psi to the host, waiting on `ev` (recorded after the copy) rather than the stream.
gpue_status download_wfc_ev(gpue_ctx* c, double* interleaved, hipEvent_t ev);
// A context on a given stream (null: its own), and new parameters for one of the
// same shape (no device allocation).
gpue_status ctx_create(const gpue_params* p, hipStream_t ext_stream, gpue_ctx** out);
gpue_status ctx_reset(gpue_ctx* c, const gpue_params* p);
bool ctx_same_shape(const gpue_ctx* c, const gpue_params* p);
void energy_combine(const gpue_ctx* c, const double* terms, double* energy, double* mu);
}  // namespace gpue
