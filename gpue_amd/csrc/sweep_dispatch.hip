// sweep_dispatch.hip -- maps (precision, N, variant) to the per-N translation units.
#include "misc_kernels.h"
#include "sweep_launch.h"

namespace gpue {

#define GPUE_DECL1(NAME_)                                                                         \
  hipError_t NAME_(bool row, int narrow, int mode, const SweepArgs& a, hipStream_t s);            \
  hipError_t NAME_##_b(bool row, int narrow, int mode, const SweepArgs* d, int ninst, int lx, int others, \
                       hipStream_t s);
#define GPUE_DECL(N_)                                                                     \
  GPUE_DECL1(sweep_f64_##N_##_v0) GPUE_DECL1(sweep_f32_##N_##_v0)                         \
  GPUE_DECL1(sweep_f64_##N_##_v1) GPUE_DECL1(sweep_f32_##N_##_v1)
GPUE_DECL(16) GPUE_DECL(32) GPUE_DECL(48) GPUE_DECL(64) GPUE_DECL(96) GPUE_DECL(128) GPUE_DECL(192) GPUE_DECL(256)
GPUE_DECL(384) GPUE_DECL(512) GPUE_DECL(1024) GPUE_DECL(2048)
GPUE_DECL(4096)
#undef GPUE_DECL

bool sweep_supported(int N) {
  switch (N) {
    case 16: case 32: case 48: case 64: case 96: case 128: case 192: case 256: case 384: case 512: case 1024:
    case 2048: case 4096:
      return true;
    default: return false;
  }
}

// Kernel variant: bit 0 the plan (0 wide lines, 1 narrow lines); var >> 1
// narrows the strided sweeps' tiles by 2^(var >> 1); var >> 1 > 0 also loads
// each transform's twiddles before its first pass (row and strided sweeps).
hipError_t launch_sweep(int prec, bool row, int N, int var, int mode, const SweepArgs& a, hipStream_t s,
                        bool* handled) {
  *handled = true;
  const int pv = var & 1, nw = var >> 1;
#define GPUE_CASE(N_)                                                                         \
  case N_:                                                                                    \
    if (pv == 0)                                                                              \
      return prec == 64 ? sweep_f64_##N_##_v0(row, nw, mode, a, s) : sweep_f32_##N_##_v0(row, nw, mode, a, s); \
    return prec == 64 ? sweep_f64_##N_##_v1(row, nw, mode, a, s) : sweep_f32_##N_##_v1(row, nw, mode, a, s);
  switch (N) {
    GPUE_CASE(16) GPUE_CASE(32) GPUE_CASE(48) GPUE_CASE(64) GPUE_CASE(96) GPUE_CASE(128) GPUE_CASE(192) GPUE_CASE(256)
    GPUE_CASE(384) GPUE_CASE(512) GPUE_CASE(1024) GPUE_CASE(2048)
    GPUE_CASE(4096)
    default: *handled = false; return hipErrorInvalidValue;
  }
#undef GPUE_CASE
}

hipError_t launch_sweep_batched(int prec, bool row, int N, int var, int mode, const SweepArgs* d, int ninst, int lx,
                                int others, hipStream_t s, bool* handled) {
  *handled = true;
  const int pv = var & 1, nw = var >> 1;
#define GPUE_CASE(N_)                                                                                  \
  case N_:                                                                                             \
    if (pv == 0)                                                                                       \
      return prec == 64 ? sweep_f64_##N_##_v0_b(row, nw, mode, d, ninst, lx, others, s)            \
                        : sweep_f32_##N_##_v0_b(row, nw, mode, d, ninst, lx, others, s);           \
    return prec == 64 ? sweep_f64_##N_##_v1_b(row, nw, mode, d, ninst, lx, others, s)              \
                      : sweep_f32_##N_##_v1_b(row, nw, mode, d, ninst, lx, others, s);
  switch (N) {
    GPUE_CASE(16) GPUE_CASE(32) GPUE_CASE(48) GPUE_CASE(64) GPUE_CASE(96) GPUE_CASE(128) GPUE_CASE(192) GPUE_CASE(256)
    GPUE_CASE(384) GPUE_CASE(512) GPUE_CASE(1024) GPUE_CASE(2048)
    GPUE_CASE(4096)
    default: *handled = false; return hipErrorInvalidValue;
  }
#undef GPUE_CASE
}

// Whether the sweep's workgroups are whole waves (64 lanes or a multiple): the
// deferred norm reduction inside a sweep needs full waves.
bool sweep_full_waves(int prec, bool row, int N, int var) {
#define GPUE_WG(N_, V_) (prec == 64 ? ColCfg<double, N_, V_>::WG : ColCfg<float, N_, V_>::WG)
#define GPUE_CASE(N_)                                                                              \
  case N_:                                                                                         \
    if (row) return ((var & 1) ? RowCfg<N_, 1>::WG : RowCfg<N_, 0>::WG) % 64 == 0;              \
    switch (var) {                                                                                 \
      case 0: return GPUE_WG(N_, 0) % 64 == 0;                                                     \
      case 1: return GPUE_WG(N_, 1) % 64 == 0;                                                     \
      case 2: return GPUE_WG(N_, 2) % 64 == 0;                                                     \
      case 3: return GPUE_WG(N_, 3) % 64 == 0;                                                     \
      case 4: return GPUE_WG(N_, 4) % 64 == 0;                                                     \
      default: return GPUE_WG(N_, 5) % 64 == 0;                                                    \
    }
  switch (N) {
    GPUE_CASE(16) GPUE_CASE(32) GPUE_CASE(48) GPUE_CASE(64) GPUE_CASE(96) GPUE_CASE(128) GPUE_CASE(192) GPUE_CASE(256)
    GPUE_CASE(384) GPUE_CASE(512) GPUE_CASE(1024) GPUE_CASE(2048)
    GPUE_CASE(4096)
    default: return false;
  }
#undef GPUE_CASE
#undef GPUE_WG
}

int sweep_blocks(int prec, bool row, int N, int var, int lines_or_xdim, int others) {
#define GPUE_COLB(N_, V_) \
  (prec == 64 ? col_blocks<double, N_, V_>(lines_or_xdim, others) : col_blocks<float, N_, V_>(lines_or_xdim, others))
#define GPUE_CASE(N_)                                                                              \
  case N_:                                                                                         \
    if (row) return (var & 1) ? row_blocks<N_, 1>(lines_or_xdim) : row_blocks<N_, 0>(lines_or_xdim); \
    switch (var) {                                                                                 \
      case 0: return GPUE_COLB(N_, 0);                                                             \
      case 1: return GPUE_COLB(N_, 1);                                                             \
      case 2: return GPUE_COLB(N_, 2);                                                             \
      case 3: return GPUE_COLB(N_, 3);                                                             \
      case 4: return GPUE_COLB(N_, 4);                                                             \
      default: return GPUE_COLB(N_, 5);                                                            \
    }
  switch (N) {
    GPUE_CASE(16) GPUE_CASE(32) GPUE_CASE(48) GPUE_CASE(64) GPUE_CASE(96) GPUE_CASE(128) GPUE_CASE(192) GPUE_CASE(256)
    GPUE_CASE(384) GPUE_CASE(512) GPUE_CASE(1024) GPUE_CASE(2048)
    GPUE_CASE(4096)
    default: return 0;
  }
#undef GPUE_CASE
#undef GPUE_COLB
}

}  // namespace gpue
