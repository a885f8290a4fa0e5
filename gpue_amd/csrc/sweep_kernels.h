// sweep_kernels.h -- row (x, contiguous) and strided (y/z) sweep kernels.
#pragma once
#include "sweeps.h"

namespace gpue {

template <int N>
struct RowCfg {
  static constexpr int TPL = Plan<N>::TPL;
  static constexpr int WG = TPL >= 128 ? TPL : 128;
  static constexpr int LPW = WG / TPL;
};

template <int N>
struct ColCfg {
  static constexpr int TPL = Plan<N>::TPL;
  static constexpr int TCr = 256 / TPL;
  static constexpr int TC = TCr > 64 ? 64 : (TCr < 1 ? 1 : TCr);
  static constexpr int WG = TPL * TC;
};

template <typename T>
__device__ __forceinline__ cx<T> div_norm(cx<T> v, double nrm) {
  // scalarDiv_wfcNorm (src/kernels.cu:479-486): psi / sqrt(sum * dV)
  return mk((T)((double)v.x / nrm), (T)((double)v.y / nrm));
}

template <typename T>
__device__ __forceinline__ double abs2d(cx<T> v) {
  return (double)v.x * (double)v.x + (double)v.y * (double)v.y;
}

// ---------------------------------------------------------------------------
// Row sweep: lines are contiguous x-rows, line = y + yDim*z.
// ---------------------------------------------------------------------------
template <typename T, int N, int MODE>
__global__ void __launch_bounds__(RowCfg<N>::WG) row_sweep(SweepArgs a) {
  using P = Plan<N>;
  constexpr int TPL = P::TPL, E = P::E, LPW = RowCfg<N>::LPW;
  constexpr int FOOT = LdsRow::footprint(N);
  __shared__ cx<T> lds[P::NP > 1 ? LPW * FOOT : 1];
  __shared__ double sh[64];
  const Phys& p = a.ph;
  const unsigned flags = a.flags;
  const int tid = threadIdx.x % TPL, lw = threadIdx.x / TPL;
  const int line = blockIdx.x * LPW + lw;
  const bool active = line < a.lines;
  const int yi = active ? line % p.yDim : 0, zi = active ? line / p.yDim : 0;
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + (long)(active ? line : 0) * N;
  cx<T>* __restrict__ out = (cx<T>*)a.out + (long)(active ? line : 0) * N;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  const LdsRow la{lw * FOOT};
  constexpr int RF0 = Seq<N, ORD_FWD>::first(), RFL = Seq<N, ORD_FWD>::last();
  cx<T> v[E];
  double acc = 0.0;

  if constexpr (MODE == RM_INV) {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = in[bpos<N, RFL>(tid, r)];
    }
    fft_regs<T, N, ORD_REV, +1>(v, tid, tw, lds, la, active);
    if (flags & F_HALFV_POST) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = op_halfV(p, v[r], bpos<N, RF0>(tid, r), yi, zi);
    }
    if (flags & F_GAUGE_POST) {
      fft_regs<T, N, ORD_FWD, -1>(v, tid, tw, lds, la, active);
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = op_A(p, v[r], 0, bpos<N, RFL>(tid, r), yi, zi, a.invN);
      fft_regs<T, N, ORD_REV, +1>(v, tid, tw, lds, la, active);
    }
  } else {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = in[bpos<N, RF0>(tid, r)];
    }
    if constexpr (MODE == RM_FWD || MODE == RM_FULL) {
      if (flags & F_SCALE_IN) {
        const double nrm = *a.norm_in;
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = div_norm(v[r], nrm);
      }
      if (flags & F_HALFV_PRE) {
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = op_halfV(p, v[r], bpos<N, RF0>(tid, r), yi, zi);
      }
    }
    fft_regs<T, N, ORD_FWD, -1>(v, tid, tw, lds, la, active);
    if constexpr (MODE == RM_GAUGE) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = op_A(p, v[r], 0, bpos<N, RFL>(tid, r), yi, zi, a.invN);
      fft_regs<T, N, ORD_REV, +1>(v, tid, tw, lds, la, active);
    } else if constexpr (MODE == RM_FULL) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = op_K(p, v[r], bpos<N, RFL>(tid, r), yi, zi, a.invN);
      fft_regs<T, N, ORD_REV, +1>(v, tid, tw, lds, la, active);
      if (flags & F_HALFV_POST) {
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = op_halfV(p, v[r], bpos<N, RF0>(tid, r), yi, zi);
      }
      if (flags & F_GAUGE_POST) {
        fft_regs<T, N, ORD_FWD, -1>(v, tid, tw, lds, la, active);
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = op_A(p, v[r], 0, bpos<N, RFL>(tid, r), yi, zi, a.invN);
        fft_regs<T, N, ORD_REV, +1>(v, tid, tw, lds, la, active);
      }
    } else if constexpr (MODE == RM_KSUM) {
      if (active) {
#pragma unroll
        for (int r = 0; r < E; ++r) {
          const int xi = bpos<N, RFL>(tid, r);
          acc += K_val(p, xi, yi, zi, gid_of(p, xi, yi, zi)) * abs2d(v[r]);
        }
      }
    }
  }

  // store: RM_FWD leaves the forward layout (last radix), the others the first radix
  if constexpr (MODE == RM_FWD) {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) out[bpos<N, RFL>(tid, r)] = v[r];
    }
  } else if constexpr (MODE != RM_KSUM) {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        out[bpos<N, RF0>(tid, r)] = v[r];
        if (flags & F_NORM) acc += abs2d(v[r]);
      }
    }
  }
  if (MODE == RM_KSUM || (flags & F_NORM)) {
    double pv[1] = {acc};
    block_partial<1>(pv, a.partials, sh);
  }
}

// ---------------------------------------------------------------------------
// Strided sweep: lines along y (axis 1, stride xDim) or z (axis 2, stride
// xDim*yDim); a workgroup owns TC adjacent x-lines, lanes vary fastest in x.
// ---------------------------------------------------------------------------
template <typename T, int N, int MODE>
__global__ void __launch_bounds__(ColCfg<N>::WG) col_sweep(SweepArgs a) {
  using P = Plan<N>;
  constexpr int TC = ColCfg<N>::TC, E = P::E;
  __shared__ cx<T> lds[P::NP > 1 ? TC * N : 1];
  __shared__ double sh[64];
  const Phys& p = a.ph;
  const unsigned flags = a.flags;
  const int c = threadIdx.x % TC, tid = threadIdx.x / TC;
  const int tiles = (p.xDim + TC - 1) / TC;
  const int xt = blockIdx.x % tiles, o = blockIdx.x / tiles;
  const int xi = xt * TC + c;
  const bool active = xi < p.xDim;
  const long es = (a.axis == 1) ? (long)p.xDim : (long)p.xDim * p.yDim;
  const long base = (a.axis == 1) ? (long)(active ? xi : 0) + (long)o * p.xDim * p.yDim
                                  : (long)(active ? xi : 0) + (long)o * p.xDim;
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + base;
  cx<T>* __restrict__ out = (cx<T>*)a.out + base;
  const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
  const LdsCol la{c, TC};
  constexpr int RF0 = Seq<N, ORD_FWD>::first(), RFL = Seq<N, ORD_FWD>::last();
  cx<T> v[E];
  double acc = 0.0;
  // coordinates of element e along the axis
  auto yz = [&](int e, int& yi, int& zi) {
    if (a.axis == 1) { yi = e; zi = o; } else { yi = o; zi = e; }
  };

  if constexpr (MODE == CM_INV) {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = in[bpos<N, RFL>(tid, r) * es];
    }
    fft_regs<T, N, ORD_REV, +1>(v, tid, tw, lds, la, active);
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) out[bpos<N, RF0>(tid, r) * es] = v[r];
    }
  } else {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = in[bpos<N, RF0>(tid, r) * es];
    }
    fft_regs<T, N, ORD_FWD, -1>(v, tid, tw, lds, la, active);
    if constexpr (MODE == CM_FWD) {
      if (active) {
#pragma unroll
        for (int r = 0; r < E; ++r) out[bpos<N, RFL>(tid, r) * es] = v[r];
      }
    } else if constexpr (MODE == CM_KSUM) {
      if (active) {
#pragma unroll
        for (int r = 0; r < E; ++r) {
          int yi, zi;
          yz(bpos<N, RFL>(tid, r), yi, zi);
          acc += K_val(p, xi, yi, zi, gid_of(p, xi, yi, zi)) * abs2d(v[r]);
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        int yi, zi;
        yz(bpos<N, RFL>(tid, r), yi, zi);
        if constexpr (MODE == CM_K) v[r] = op_K(p, v[r], xi, yi, zi, a.invN);
        else v[r] = op_A(p, v[r], a.axis, xi, yi, zi, a.invN);
      }
      fft_regs<T, N, ORD_REV, +1>(v, tid, tw, lds, la, active);
      if (active) {
#pragma unroll
        for (int r = 0; r < E; ++r) {
          out[bpos<N, RF0>(tid, r) * es] = v[r];
          if (flags & F_NORM) acc += abs2d(v[r]);
        }
      }
    }
  }
  if (MODE == CM_KSUM || (flags & F_NORM)) {
    double pv[1] = {acc};
    block_partial<1>(pv, a.partials, sh);
  }
}

}  // namespace gpue
