// sweep_launch.h -- grid shapes and launchers for the sweep kernels.
#pragma once
#include "sweep_kernels.h"

namespace gpue {

template <int N, int V>
inline int row_blocks(int lines) { return (lines + RowCfg<N, V>::LPW - 1) / RowCfg<N, V>::LPW; }
template <typename T, int N, int CV>
inline int col_blocks(int xDim, int others) {
  return ((xDim + ColCfg<T, N, CV>::TC - 1) / ColCfg<T, N, CV>::TC) * others;
}

// FP32 real-time steps run the kernels whose mode carries bit 16 (phase factors
// and their products in FP64, sweep_kernels.h); every other launch the plain ones.
// Kernels with time-dependent equation fields evaluated per point (bit 32 of
// the mode): only for the default variants (wide lines, or narrow lines with
// quarter strided tiles), which contexts with such fields always use; no
// bridge modes (their steps run one by one).
// FP64 only: GPUE's fields are FP64, and FP32 contexts reject time-dependent ones.
template <typename T, int RV>
constexpr bool row_eq_variant() { return sizeof(T) == 8 && (RV == 0 || RV == 3); }
template <typename T, int CV>
constexpr bool col_eq_variant() { return sizeof(T) == 8 && (CV == 0 || CV == 5); }

template <typename T, int N, int V>
hipError_t launch_row_eq(int mode, const SweepArgs& a, hipStream_t s) {
  if constexpr (row_eq_variant<T, V>()) {
    dim3 grid(row_blocks<N, V>(a.lines)), block(RowCfg<N, V>::WG);
#define GPUE_ROW_EQ(M_)                                                                          \
  hipLaunchKernelGGL((row_sweep<T, N, V, M_ | 32>), grid, block, 0, s, a);                       \
  break;
    switch (mode) {
      case RM_FWD: GPUE_ROW_EQ(RM_FWD)
      case RM_INV: GPUE_ROW_EQ(RM_INV)
      case RM_GAUGE: GPUE_ROW_EQ(RM_GAUGE)
      case RM_FULL: GPUE_ROW_EQ(RM_FULL)
      case RM_K: GPUE_ROW_EQ(RM_K)
      default: return hipErrorInvalidValue;
    }
#undef GPUE_ROW_EQ
    return hipGetLastError();
  } else {
    (void)mode; (void)a; (void)s;
    return hipErrorInvalidValue;
  }
}

template <typename T, int N, int V>
hipError_t launch_col_eq(int mode, const SweepArgs& a, hipStream_t s) {
  if constexpr (col_eq_variant<T, V>()) {
    const int others = a.axis == 1 ? a.ph.zDim : a.ph.yDim;
    dim3 grid(col_blocks<T, N, V>(a.ph.xDim, others)), block(ColCfg<T, N, V>::WG);
#define GPUE_COL_EQ(M_)                                                                          \
  hipLaunchKernelGGL((col_sweep<T, N, V, M_ | 32>), grid, block, 0, s, a);                       \
  break;
    switch (mode) {
      case CM_K: GPUE_COL_EQ(CM_K)
      case CM_GAUGE: GPUE_COL_EQ(CM_GAUGE)
      case CM_FWD_V: GPUE_COL_EQ(CM_FWD_V)
      case CM_INV_VG: GPUE_COL_EQ(CM_INV_VG)
      default: return hipErrorInvalidValue;
    }
#undef GPUE_COL_EQ
    return hipGetLastError();
  } else {
    (void)mode; (void)a; (void)s;
    return hipErrorInvalidValue;
  }
}

template <typename T, int N, int V>
hipError_t launch_row_t(int mode, const SweepArgs& a, hipStream_t s) {
  if (a.ph.eqk && mode != RM_KSUM) return launch_row_eq<T, N, V>(mode, a, s);
  dim3 grid(row_blocks<N, V>(a.lines)), block(RowCfg<N, V>::WG);
  // closing sweeps with no gauge operator take the kernels compiled without it
  const bool ng = !(a.flags & F_GAUGE_POST);
  if constexpr (sizeof(T) == 4) {
    if (a.ph.gstate != 0 && mode != RM_KSUM) {
      switch (mode) {
        case RM_FWD: hipLaunchKernelGGL((row_sweep<T, N, V, RM_FWD | 16>), grid, block, 0, s, a); break;
        case RM_INV:
          if (ng) hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV | 16 | 64>), grid, block, 0, s, a);
          else hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV | 16>), grid, block, 0, s, a);
          break;
        case RM_GAUGE: hipLaunchKernelGGL((row_sweep<T, N, V, RM_GAUGE | 16>), grid, block, 0, s, a); break;
        case RM_FULL: hipLaunchKernelGGL((row_sweep<T, N, V, RM_FULL | 16>), grid, block, 0, s, a); break;
        case RM_INV_FWD:
          if (ng) hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV_FWD | 16 | 64>), grid, block, 0, s, a);
          else hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV_FWD | 16>), grid, block, 0, s, a);
          break;
        case RM_GAUGE_FWD: hipLaunchKernelGGL((row_sweep<T, N, V, RM_GAUGE_FWD | 16>), grid, block, 0, s, a); break;
        case RM_K: hipLaunchKernelGGL((row_sweep<T, N, V, RM_K | 16>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
      }
      return hipGetLastError();
    }
  }
  switch (mode) {
    case RM_FWD: hipLaunchKernelGGL((row_sweep<T, N, V, RM_FWD>), grid, block, 0, s, a); break;
    case RM_INV:
      if (ng) hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV | 64>), grid, block, 0, s, a);
      else hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV>), grid, block, 0, s, a);
      break;
    case RM_GAUGE: hipLaunchKernelGGL((row_sweep<T, N, V, RM_GAUGE>), grid, block, 0, s, a); break;
    case RM_FULL: hipLaunchKernelGGL((row_sweep<T, N, V, RM_FULL>), grid, block, 0, s, a); break;
    case RM_KSUM: hipLaunchKernelGGL((row_sweep<T, N, V, RM_KSUM>), grid, block, 0, s, a); break;
    case RM_INV_FWD:
      if (ng) hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV_FWD | 64>), grid, block, 0, s, a);
      else hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV_FWD>), grid, block, 0, s, a);
      break;
    case RM_GAUGE_FWD: hipLaunchKernelGGL((row_sweep<T, N, V, RM_GAUGE_FWD>), grid, block, 0, s, a); break;
    case RM_K: hipLaunchKernelGGL((row_sweep<T, N, V, RM_K>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, int N, int V>
hipError_t launch_col_t(int mode, const SweepArgs& a, hipStream_t s) {
  if (a.ph.eqk && (mode == CM_K || mode == CM_GAUGE || mode == CM_FWD_V || mode == CM_INV_VG))
    return launch_col_eq<T, N, V>(mode, a, s);
  const int others = a.axis == 1 ? a.ph.zDim : a.ph.yDim;
  dim3 grid(col_blocks<T, N, V>(a.ph.xDim, others)), block(ColCfg<T, N, V>::WG);
  if constexpr (sizeof(T) == 4) {
    if (a.ph.gstate != 0 && col_phase_mode(mode)) {
      switch (mode) {
        case CM_K: hipLaunchKernelGGL((col_sweep<T, N, V, CM_K | 16>), grid, block, 0, s, a); break;
        case CM_GAUGE: hipLaunchKernelGGL((col_sweep<T, N, V, CM_GAUGE | 16>), grid, block, 0, s, a); break;
        case CM_FWD_V: hipLaunchKernelGGL((col_sweep<T, N, V, CM_FWD_V | 16>), grid, block, 0, s, a); break;
        case CM_INV_VG: hipLaunchKernelGGL((col_sweep<T, N, V, CM_INV_VG | 16>), grid, block, 0, s, a); break;
        case CM_GAUGE_FWD: hipLaunchKernelGGL((col_sweep<T, N, V, CM_GAUGE_FWD | 16>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
      }
      return hipGetLastError();
    }
  }
  switch (mode) {
    case CM_FWD: hipLaunchKernelGGL((col_sweep<T, N, V, CM_FWD>), grid, block, 0, s, a); break;
    case CM_INV: hipLaunchKernelGGL((col_sweep<T, N, V, CM_INV>), grid, block, 0, s, a); break;
    case CM_K: hipLaunchKernelGGL((col_sweep<T, N, V, CM_K>), grid, block, 0, s, a); break;
    case CM_GAUGE: hipLaunchKernelGGL((col_sweep<T, N, V, CM_GAUGE>), grid, block, 0, s, a); break;
    case CM_KSUM: hipLaunchKernelGGL((col_sweep<T, N, V, CM_KSUM>), grid, block, 0, s, a); break;
    case CM_FWD_V: hipLaunchKernelGGL((col_sweep<T, N, V, CM_FWD_V>), grid, block, 0, s, a); break;
    case CM_INV_VG: hipLaunchKernelGGL((col_sweep<T, N, V, CM_INV_VG>), grid, block, 0, s, a); break;
    case CM_GAUGE_FWD: hipLaunchKernelGGL((col_sweep<T, N, V, CM_GAUGE_FWD>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Batched: `ninst` problems of one shape, arguments in device memory.
template <typename T, int N, int V>
hipError_t launch_row_b_t(int mode, const SweepArgs* d, int ninst, int lines, hipStream_t s) {
  const int bpi = row_blocks<N, V>(lines);
  dim3 grid(bpi * ninst), block(RowCfg<N, V>::WG);
  if constexpr (sizeof(T) == 4) {
    if (mode & 16) {   // FP32 real time (see launch_row_t)
      switch (mode & 15) {
        case RM_FWD: hipLaunchKernelGGL((row_sweep_b<T, N, V, RM_FWD | 16>), grid, block, 0, s, d, bpi); break;
        case RM_INV: hipLaunchKernelGGL((row_sweep_b<T, N, V, RM_INV | 16>), grid, block, 0, s, d, bpi); break;
        case RM_GAUGE: hipLaunchKernelGGL((row_sweep_b<T, N, V, RM_GAUGE | 16>), grid, block, 0, s, d, bpi); break;
        case RM_FULL: hipLaunchKernelGGL((row_sweep_b<T, N, V, RM_FULL | 16>), grid, block, 0, s, d, bpi); break;
        case RM_K: hipLaunchKernelGGL((row_sweep_b<T, N, V, RM_K | 16>), grid, block, 0, s, d, bpi); break;
        default: return hipErrorInvalidValue;
      }
      return hipGetLastError();
    }
  }
  switch (mode) {
    case RM_FWD: hipLaunchKernelGGL((row_sweep_b<T, N, V, RM_FWD>), grid, block, 0, s, d, bpi); break;
    case RM_INV: hipLaunchKernelGGL((row_sweep_b<T, N, V, RM_INV>), grid, block, 0, s, d, bpi); break;
    case RM_GAUGE: hipLaunchKernelGGL((row_sweep_b<T, N, V, RM_GAUGE>), grid, block, 0, s, d, bpi); break;
    case RM_FULL: hipLaunchKernelGGL((row_sweep_b<T, N, V, RM_FULL>), grid, block, 0, s, d, bpi); break;
    case RM_K: hipLaunchKernelGGL((row_sweep_b<T, N, V, RM_K>), grid, block, 0, s, d, bpi); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, int N, int V>
hipError_t launch_col_b_t(int mode, const SweepArgs* d, int ninst, int xDim, int others, hipStream_t s) {
  const int bpi = col_blocks<T, N, V>(xDim, others);
  dim3 grid(bpi * ninst), block(ColCfg<T, N, V>::WG);
  if constexpr (sizeof(T) == 4) {
    if (mode & 16) {   // FP32 real time (see launch_row_t)
      switch (mode & 15) {
        case CM_K: hipLaunchKernelGGL((col_sweep_b<T, N, V, CM_K | 16>), grid, block, 0, s, d, bpi); break;
        case CM_GAUGE: hipLaunchKernelGGL((col_sweep_b<T, N, V, CM_GAUGE | 16>), grid, block, 0, s, d, bpi); break;
        case CM_FWD_V: hipLaunchKernelGGL((col_sweep_b<T, N, V, CM_FWD_V | 16>), grid, block, 0, s, d, bpi); break;
        case CM_INV_VG: hipLaunchKernelGGL((col_sweep_b<T, N, V, CM_INV_VG | 16>), grid, block, 0, s, d, bpi); break;
        default: return hipErrorInvalidValue;
      }
      return hipGetLastError();
    }
  }
  switch (mode) {
    case CM_FWD: hipLaunchKernelGGL((col_sweep_b<T, N, V, CM_FWD>), grid, block, 0, s, d, bpi); break;
    case CM_INV: hipLaunchKernelGGL((col_sweep_b<T, N, V, CM_INV>), grid, block, 0, s, d, bpi); break;
    case CM_K: hipLaunchKernelGGL((col_sweep_b<T, N, V, CM_K>), grid, block, 0, s, d, bpi); break;
    case CM_GAUGE: hipLaunchKernelGGL((col_sweep_b<T, N, V, CM_GAUGE>), grid, block, 0, s, d, bpi); break;
    case CM_FWD_V: hipLaunchKernelGGL((col_sweep_b<T, N, V, CM_FWD_V>), grid, block, 0, s, d, bpi); break;
    case CM_INV_VG: hipLaunchKernelGGL((col_sweep_b<T, N, V, CM_INV_VG>), grid, block, 0, s, d, bpi); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gpue

// One translation unit per (precision, N, plan) keeps the build parallel; it
// holds the strided sweeps of every tile width (narrow = 0, 1, 2) and the row
// sweeps with per-pass or up-front twiddle loads (narrow = 0 or > 0).
#define GPUE_SWEEP_TU(T_, N_, V_, NAME_)                                                          \
  namespace gpue {                                                                                \
  hipError_t NAME_(bool row, int narrow, int mode, const SweepArgs& a, hipStream_t s) {           \
    if (row) return narrow ? launch_row_t<T_, N_, V_ + 2>(mode, a, s) : launch_row_t<T_, N_, V_>(mode, a, s); \
    if (narrow == 1) return launch_col_t<T_, N_, V_ + 2>(mode, a, s);                             \
    if (narrow == 2) return launch_col_t<T_, N_, V_ + 4>(mode, a, s);                             \
    return launch_col_t<T_, N_, V_>(mode, a, s);                                                  \
  }                                                                                               \
  hipError_t NAME_##_b(bool row, int narrow, int mode, const SweepArgs* d, int ninst, int lx,     \
                       int others, hipStream_t s) {                                               \
    if (row)                                                                                      \
      return narrow ? launch_row_b_t<T_, N_, V_ + 2>(mode, d, ninst, lx, s)                       \
                    : launch_row_b_t<T_, N_, V_>(mode, d, ninst, lx, s);                          \
    if (narrow == 1) return launch_col_b_t<T_, N_, V_ + 2>(mode, d, ninst, lx, others, s);        \
    if (narrow == 2) return launch_col_b_t<T_, N_, V_ + 4>(mode, d, ninst, lx, others, s);        \
    return launch_col_b_t<T_, N_, V_>(mode, d, ninst, lx, others, s);                             \
  }                                                                                               \
  }
