// misc_kernels.h -- launchers for the pointwise kernels around the sweeps.
#pragma once
#include "sweeps.h"

namespace gpue {

constexpr int REDUCE_BLOCKS = 1024;   // grid of the standalone reductions

struct EqnProg;
hipError_t launch_eqn_field(const EqnProg& prog, double* out, int xDim, int yDim, int zDim, double dx, double dy,
                            double dz, double xMax, double yMax, double zMax, double t, hipStream_t s);
hipError_t launch_fields_at(const Phys& p, double* V, double* K, double* A0, double* A1, double* A2, long N,
                            hipStream_t s);
template <typename T>
hipError_t launch_apply_norm(void* psi, long N, const double* norm, hipStream_t s);
template <typename T>
hipError_t launch_norm_sum(const void* psi, long N, double* partials, int* nb, hipStream_t s);
hipError_t launch_finalize(const double* partials, int nb, int na, int mode, double scale, double* out, hipStream_t s);
template <typename T>
hipError_t launch_energy_pos(const void* psi, const Phys& p, int gauge, double hbar, double dx, double dy, double dz,
                             double* partials, int* nb, hipStream_t s);
template <typename T>
hipError_t launch_vortex(const void* psi, const unsigned char* mask, int xDim, int* cand, int* marker, int* found,
                         hipStream_t s);

// sweep dispatch (sweep_inst_*.hip)
hipError_t launch_sweep(int prec, bool row, int N, int var, int mode, const SweepArgs& a, hipStream_t s,
                        bool* handled);
int sweep_blocks(int prec, bool row, int N, int var, int lines_or_xdim, int others);
bool sweep_full_waves(int prec, bool row, int N, int var);
hipError_t launch_sweep_batched(int prec, bool row, int N, int var, int mode, const SweepArgs* d, int ninst, int lx,
                                int others, hipStream_t s, bool* handled);

// Second stage of a batched reduction: one workgroup per instance.
struct FinArgs {
  int active;
  int nb, na, mode;
  double scale;
  const double* partials;
  double* out;
};
hipError_t launch_finalize_batched(const FinArgs* d, int ninst, hipStream_t s);
hipError_t launch_set_double(double* p, double v, hipStream_t s);
hipError_t launch_energy_fold(double* results, int naxes, hipStream_t s);
hipError_t launch_copy16(const void* in, void* out, size_t bytes, hipStream_t s);
hipError_t launch_math_eval(int fn, const double* x, double* y, long n, hipStream_t s);
// variant = plan (bit 0: 0 wide lines, 1 narrow lines) + 2 * narrow, where the
// strided sweeps' tiles are 2^narrow times narrower (narrow 0..2) and, for
// narrow > 0, every sweep loads its transforms' twiddles up front (row sweeps
// have variants 0..3 only)
constexpr int N_VARIANTS = 6;
bool sweep_supported(int N);

}  // namespace gpue
