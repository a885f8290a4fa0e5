// sweep_launch.h -- grid shapes and launchers for the sweep kernels.
#pragma once
#include "sweep_kernels.h"

namespace gpue {

template <int N>
inline int row_blocks(int lines) { return (lines + RowCfg<N>::LPW - 1) / RowCfg<N>::LPW; }
template <int N>
inline int col_blocks(int xDim, int others) { return ((xDim + ColCfg<N>::TC - 1) / ColCfg<N>::TC) * others; }

template <typename T, int N>
hipError_t launch_row_t(int mode, const SweepArgs& a, hipStream_t s) {
  dim3 grid(row_blocks<N>(a.lines)), block(RowCfg<N>::WG);
  switch (mode) {
    case RM_FWD: hipLaunchKernelGGL((row_sweep<T, N, RM_FWD>), grid, block, 0, s, a); break;
    case RM_INV: hipLaunchKernelGGL((row_sweep<T, N, RM_INV>), grid, block, 0, s, a); break;
    case RM_GAUGE: hipLaunchKernelGGL((row_sweep<T, N, RM_GAUGE>), grid, block, 0, s, a); break;
    case RM_FULL: hipLaunchKernelGGL((row_sweep<T, N, RM_FULL>), grid, block, 0, s, a); break;
    case RM_KSUM: hipLaunchKernelGGL((row_sweep<T, N, RM_KSUM>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, int N>
hipError_t launch_col_t(int mode, const SweepArgs& a, hipStream_t s) {
  const int others = a.axis == 1 ? a.ph.zDim : a.ph.yDim;
  dim3 grid(col_blocks<N>(a.ph.xDim, others)), block(ColCfg<N>::WG);
  switch (mode) {
    case CM_FWD: hipLaunchKernelGGL((col_sweep<T, N, CM_FWD>), grid, block, 0, s, a); break;
    case CM_INV: hipLaunchKernelGGL((col_sweep<T, N, CM_INV>), grid, block, 0, s, a); break;
    case CM_K: hipLaunchKernelGGL((col_sweep<T, N, CM_K>), grid, block, 0, s, a); break;
    case CM_GAUGE: hipLaunchKernelGGL((col_sweep<T, N, CM_GAUGE>), grid, block, 0, s, a); break;
    case CM_KSUM: hipLaunchKernelGGL((col_sweep<T, N, CM_KSUM>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gpue

// One translation unit per (precision, N) keeps the build parallel.
#define GPUE_SWEEP_TU(T_, N_, NAME_)                                                              \
  namespace gpue {                                                                                \
  hipError_t NAME_(bool row, int mode, const SweepArgs& a, hipStream_t s) {                       \
    return row ? launch_row_t<T_, N_>(mode, a, s) : launch_col_t<T_, N_>(mode, a, s);             \
  }                                                                                               \
  }
