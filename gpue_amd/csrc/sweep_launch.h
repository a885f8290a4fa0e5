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
template <typename T, int N, int V>
hipError_t launch_row_t(int mode, const SweepArgs& a, hipStream_t s) {
  dim3 grid(row_blocks<N, V>(a.lines)), block(RowCfg<N, V>::WG);
  if constexpr (sizeof(T) == 4) {
    if (a.ph.gstate != 0 && mode != RM_KSUM) {
      switch (mode) {
        case RM_FWD: hipLaunchKernelGGL((row_sweep<T, N, V, RM_FWD | 16>), grid, block, 0, s, a); break;
        case RM_INV: hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV | 16>), grid, block, 0, s, a); break;
        case RM_GAUGE: hipLaunchKernelGGL((row_sweep<T, N, V, RM_GAUGE | 16>), grid, block, 0, s, a); break;
        case RM_FULL: hipLaunchKernelGGL((row_sweep<T, N, V, RM_FULL | 16>), grid, block, 0, s, a); break;
        case RM_INV_FWD: hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV_FWD | 16>), grid, block, 0, s, a); break;
        case RM_GAUGE_FWD: hipLaunchKernelGGL((row_sweep<T, N, V, RM_GAUGE_FWD | 16>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
      }
      return hipGetLastError();
    }
  }
  switch (mode) {
    case RM_FWD: hipLaunchKernelGGL((row_sweep<T, N, V, RM_FWD>), grid, block, 0, s, a); break;
    case RM_INV: hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV>), grid, block, 0, s, a); break;
    case RM_GAUGE: hipLaunchKernelGGL((row_sweep<T, N, V, RM_GAUGE>), grid, block, 0, s, a); break;
    case RM_FULL: hipLaunchKernelGGL((row_sweep<T, N, V, RM_FULL>), grid, block, 0, s, a); break;
    case RM_KSUM: hipLaunchKernelGGL((row_sweep<T, N, V, RM_KSUM>), grid, block, 0, s, a); break;
    case RM_INV_FWD: hipLaunchKernelGGL((row_sweep<T, N, V, RM_INV_FWD>), grid, block, 0, s, a); break;
    case RM_GAUGE_FWD: hipLaunchKernelGGL((row_sweep<T, N, V, RM_GAUGE_FWD>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, int N, int V>
hipError_t launch_col_t(int mode, const SweepArgs& a, hipStream_t s) {
  const int others = a.axis == 1 ? a.ph.zDim : a.ph.yDim;
  dim3 grid(col_blocks<T, N, V>(a.ph.xDim, others)), block(ColCfg<T, N, V>::WG);
  if constexpr (sizeof(T) == 4) {
    if (a.ph.gstate != 0 && (mode == CM_K || mode == CM_GAUGE)) {
      if (mode == CM_K) hipLaunchKernelGGL((col_sweep<T, N, V, CM_K | 16>), grid, block, 0, s, a);
      else hipLaunchKernelGGL((col_sweep<T, N, V, CM_GAUGE | 16>), grid, block, 0, s, a);
      return hipGetLastError();
    }
  }
  switch (mode) {
    case CM_FWD: hipLaunchKernelGGL((col_sweep<T, N, V, CM_FWD>), grid, block, 0, s, a); break;
    case CM_INV: hipLaunchKernelGGL((col_sweep<T, N, V, CM_INV>), grid, block, 0, s, a); break;
    case CM_K: hipLaunchKernelGGL((col_sweep<T, N, V, CM_K>), grid, block, 0, s, a); break;
    case CM_GAUGE: hipLaunchKernelGGL((col_sweep<T, N, V, CM_GAUGE>), grid, block, 0, s, a); break;
    case CM_KSUM: hipLaunchKernelGGL((col_sweep<T, N, V, CM_KSUM>), grid, block, 0, s, a); break;
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
