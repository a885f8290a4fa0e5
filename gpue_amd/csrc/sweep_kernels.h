// sweep_kernels.h -- row (x, contiguous) and strided (y/z) sweep kernels.
#pragma once
#include "sweeps.h"

namespace gpue {

// Workgroup shapes.  Row sweeps: LPW lines of TPL lanes.  Strided sweeps: TC
// adjacent x-lines (lanes vary fastest in x for coalescing) of TPL lanes.
// RV is the row-sweep variant: bit 0 the plan, bit 1 twiddles loaded before
// the first pass of each transform (fft_regs PRE).
template <int N, int RV>
struct RowCfg {
  static constexpr int V = RV & 1;
  static constexpr int TPL = Plan<N, V>::TPL;
  static constexpr int WGr = V == 0 ? 128 : 64;
  static constexpr int WG = TPL >= WGr ? TPL : WGr;
  static constexpr int LPW = WG / TPL;
};

// Strided sweeps read TC adjacent x-lines per row of the tile: at least one
// 128-byte cache line (8 complex doubles, 16 complex floats) and up to 256
// lanes' worth for short transforms, within 1024 lanes per workgroup.
// Measured at 256^3: FP64 8-column tiles run the strided FFT at copy speed
// (2-4 columns lose 40-70 %); FP32 8-column tiles fetch every line twice.
//
// CV is the strided-sweep variant: bit 0 the plan, bits 1-2 narrow the tile
// to a half or a quarter of that width. Small grids have few tiles (64 of
// them for 512 FP64 columns), so narrower tiles put more CUs to work; the
// tiles that share a 128-byte line then run on one XCD (col_sweep), so the
// line is fetched into one L2 only.  Narrowed FP64 tiles also load each
// transform's twiddles before its first pass (these launches are latency
// bound and have registers to spare); FP32 tiles keep per-pass loads, since
// the up-front tables (hi + lo parts) cost 60-100 more VGPRs and FP32 tiles
// serve batches whose sweeps are limited by occupancy (82 VGPRs: 5 waves per
// SIMD; 141-180 with the tables: 2-3).
template <typename T, int N, int CV>
struct ColCfg {
  static constexpr int PV = CV & 1, NARROW = CV >> 1;
  static constexpr int TPL = Plan<N, PV>::TPL;
  static constexpr int TCmin = 128 / (2 * (int)sizeof(T));
  static constexpr int TCa = 256 / TPL > TCmin ? 256 / TPL : TCmin;
  static constexpr int TCb = TCa > 64 ? 64 : TCa;
  static constexpr int TCw = TPL * TCb > 1024 ? 1024 / TPL : TCb;
  // the tile's LDS (TC lines of N points) within 128 KiB of the 160 KiB per workgroup
  static constexpr int TCm = 131072 / (N * 2 * (int)sizeof(T));
  static constexpr int TC0 = TCw < TCm ? TCw : TCm;
  static constexpr int TC = (TC0 >> NARROW) > 0 ? (TC0 >> NARROW) : 1;
  static constexpr int G = TC0 / TC;   // tiles per default-width tile
  static constexpr int WG = TPL * TC;
  static_assert(TC0 >= 1 && G >= 1 && WG <= 1024, "strided tile does not fit a workgroup");
};

// scalarDiv_wfcNorm (src/kernels.cu:479-486) divides by sqrt(sum * dV); the
// fused form multiplies by its reciprocal (within 1 ulp of the quotient).
template <typename T>
__device__ __forceinline__ cx<T> scale_norm(cx<T> v, T inv) {
  return mk(v.x * inv, v.y * inv);
}

template <typename T>
__device__ __forceinline__ double abs2d(cx<T> v) {
  return (double)v.x * (double)v.x + (double)v.y * (double)v.y;
}

// ---------------------------------------------------------------------------
// Row sweep: lines are contiguous x-rows, line = y + yDim*z.
// ---------------------------------------------------------------------------
template <typename T, int N, int RV, int MODE_>
__device__ __forceinline__ void row_body(const SweepArgs& a, int bid, int nblk) {
  // bit 16 of the mode: FP32 real-time kernels forming phase factors in FP64;
  // bit 32: time-dependent equation fields evaluated per point (EQ kernels)
  constexpr int MODE = MODE_ & 15;
  constexpr bool X = (MODE_ & 16) != 0;
  constexpr bool EQ = (MODE_ & 32) != 0;
  // bit 64: RM_INV / RM_INV_FWD compiled without the gauge branch, for sweeps
  // whose flags carry no F_GAUGE_POST (its three-transform path sets the
  // register count of the whole kernel, and so the waves per SIMD)
  constexpr bool NG = (MODE_ & 64) != 0;
  constexpr int V = RV & 1;
  constexpr bool PRE = (RV >> 1) != 0;
  using P = Plan<N, V>;
  constexpr int TPL = P::TPL, E = P::E, LPW = RowCfg<N, RV>::LPW;
  constexpr int FOOT = LdsRow::footprint(N);
  __shared__ cx<T> lds[P::NP > 1 ? LPW * FOOT : 1];
  __shared__ double sh[64];
  const Phys& p = a.ph;
  const unsigned flags = a.flags;
  const int tid = threadIdx.x % TPL, lw = threadIdx.x / TPL;
  const int line = bid * LPW + lw;
  const bool active = line < a.lines;
  const int yi = active ? line % p.yDim : 0, zi = active ? line / p.yDim : 0;
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + (long)(active ? line : 0) * N;
  cx<T>* __restrict__ out = (cx<T>*)a.out + (long)(active ? line : 0) * N;
  const auto* __restrict__ tw = (const typename TwE<T>::type*)a.tw;
  const LdsRow la{lw * FOOT};
  constexpr int RF0 = Seq<N, V, ORD_FWD>::first(), RFL = Seq<N, V, ORD_FWD>::last();
  cx<T> v[E];
  double acc = 0.0;
  TwRegs<T, N, V> twr;   // PRE: every twiddle of the sweep, loaded after psi
  // grid indices of register slot r in the boundary layouts of the first / last radix
  auto pos0 = [&](int r) { return Pt{bpos<N, TPL, RF0>(tid, r), yi, zi}; };
  auto posL = [&](int r) { return Pt{bpos<N, TPL, RFL>(tid, r), yi, zi}; };

  if constexpr (MODE == RM_INV || MODE == RM_INV_FWD) {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = in[bpos<N, TPL, RFL>(tid, r)];
    }
    if constexpr (PRE) twr.load(tid, tw);
    fft_regs_tw<T, N, V, ORD_REV, +1, PRE>(v, tid, tw, twr, lds, la, active);
    if (flags & F_HALFV_POST) halfV_all<T, E, X, EQ>(p, v, pos0);
    if (!NG && (flags & F_GAUGE_POST)) {
      fft_regs_tw<T, N, V, ORD_FWD, -1, PRE>(v, tid, tw, twr, lds, la, active);
      if (!gauge_rot_lin<T, N, TPL, RFL, E, 0, X, EQ>(p, v, tid, yi, a.invN)) gauge_all<T, E, 0, X, EQ>(p, v, posL, a.invN);
      fft_regs_tw<T, N, V, ORD_REV, +1, PRE>(v, tid, tw, twr, lds, la, active);
    }
  } else {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = in[bpos<N, TPL, RF0>(tid, r)];
    }
    if constexpr (PRE) twr.load(tid, tw);
    if constexpr (MODE == RM_FWD || MODE == RM_FULL) {
      if (flags & F_SCALE_IN) {
        // a deferred renormalisation: the norm reduced here from the previous
        // sweep's partials by every wave (all lanes active; psi is already in
        // flight), stored once by block 0
        double nrm;
        if (MODE == RM_FWD && (flags & F_SCALE_PART)) {
          nrm = part_norm(a.partials, a.norm_nb, a.norm_scale);
          if (bid == 0 && threadIdx.x == 0) *a.norm_out = nrm;
        } else {
          nrm = *a.norm_in;
        }
        const T inv = (T)(1.0 / nrm);
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = scale_norm(v[r], inv);
      }
      if (flags & F_HALFV_PRE) halfV_all<T, E, X, EQ>(p, v, pos0);
    }
    fft_regs_tw<T, N, V, ORD_FWD, -1, PRE>(v, tid, tw, twr, lds, la, active);
    if constexpr (MODE == RM_GAUGE || MODE == RM_GAUGE_FWD) {
      if (!gauge_rot_lin<T, N, TPL, RFL, E, 0, X, EQ>(p, v, tid, yi, a.invN)) gauge_all<T, E, 0, X, EQ>(p, v, posL, a.invN);
      fft_regs_tw<T, N, V, ORD_REV, +1, PRE>(v, tid, tw, twr, lds, la, active);
    } else if constexpr (MODE == RM_K) {
      kstep_all<T, E, X, EQ>(p, v, posL, a.invN);
      fft_regs_tw<T, N, V, ORD_REV, +1, PRE>(v, tid, tw, twr, lds, la, active);
    } else if constexpr (MODE == RM_FULL) {
      kstep_all<T, E, X, EQ>(p, v, posL, a.invN);
      fft_regs_tw<T, N, V, ORD_REV, +1, PRE>(v, tid, tw, twr, lds, la, active);
      if (flags & F_HALFV_POST) halfV_all<T, E, X, EQ>(p, v, pos0);
      if (flags & F_GAUGE_POST) {
        fft_regs_tw<T, N, V, ORD_FWD, -1, PRE>(v, tid, tw, twr, lds, la, active);
        if (!gauge_rot_lin<T, N, TPL, RFL, E, 0, X, EQ>(p, v, tid, yi, a.invN)) gauge_all<T, E, 0, X, EQ>(p, v, posL, a.invN);
        fft_regs_tw<T, N, V, ORD_REV, +1, PRE>(v, tid, tw, twr, lds, la, active);
      }
    } else if constexpr (MODE == RM_KSUM) {
      if (active) {
#pragma unroll
        for (int r = 0; r < E; ++r) {
          const int xi = bpos<N, TPL, RFL>(tid, r);
          const long gid = gid_of(p, xi, yi, zi);
          const double w = (flags & F_ASUM)  ? A_val(p, 0, xi, yi, zi, gid) * p.pxg[xi]
                           : (flags & F_KAX) ? p.ks[0][xi]
                                             : K_val(p, xi, yi, zi, gid);
          acc += w * abs2d(v[r]);
        }
      }
    }
  }

  // bridge modes continue with the next step's first row sweep
  if constexpr (MODE == RM_INV_FWD || MODE == RM_GAUGE_FWD) {
    if (flags & F_HALFV_PRE) halfV_all<T, E, X, EQ>(p, v, pos0);
    fft_regs_tw<T, N, V, ORD_FWD, -1, PRE>(v, tid, tw, twr, lds, la, active);
  }

  // store: forward sweeps leave the forward layout (last radix), the others the first radix
  if constexpr (MODE == RM_FWD || MODE == RM_INV_FWD || MODE == RM_GAUGE_FWD) {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) out[bpos<N, TPL, RFL>(tid, r)] = v[r];
    }
  } else if constexpr (MODE != RM_KSUM) {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        out[bpos<N, TPL, RF0>(tid, r)] = v[r];
        if (flags & F_NORM) acc += abs2d(v[r]);
      }
    }
  }
  if (MODE == RM_KSUM || (flags & F_NORM)) {
    double pv[1] = {acc};
    block_partial<1>(pv, a.partials, sh, bid, nblk);
  }
}

// ---------------------------------------------------------------------------
// Strided sweep: lines along y (axis 1, stride xDim) or z (axis 2, stride
// xDim*yDim); a workgroup owns TC adjacent x-lines, lanes vary fastest in x.
// ---------------------------------------------------------------------------
template <typename T, int N, int CV, int MODE_>
__device__ __forceinline__ void col_body(const SweepArgs& a, int bid, int nblk) {
  // bit 16 of the mode: FP32 real-time kernels forming phase factors in FP64;
  // bit 32: time-dependent equation fields evaluated per point (EQ kernels)
  constexpr int MODE = MODE_ & 15;
  constexpr bool X = (MODE_ & 16) != 0;
  constexpr bool EQ = (MODE_ & 32) != 0;
  constexpr int V = CV & 1;
  constexpr bool PRE = (CV >> 1) != 0 && sizeof(T) == 8;
  using P = Plan<N, V>;
  constexpr int TC = ColCfg<T, N, CV>::TC, E = P::E, TPL = P::TPL;
  __shared__ cx<T> lds[P::NP > 1 ? TC * N : 1];
  __shared__ double sh[64];
  const Phys& p = a.ph;
  const unsigned flags = a.flags;
  const int c = threadIdx.x % TC, tid = threadIdx.x / TC;
  const int tiles = (p.xDim + TC - 1) / TC;
  const int xt = bid % tiles, o = bid / tiles;
  const int xi = xt * TC + c;
  const bool active = xi < p.xDim;
  const int xq = active ? xi : 0;   // table index: lanes past xDim (a last partial tile) read column 0
  const long es = (a.axis == 1) ? (long)p.xDim : (long)p.xDim * p.yDim;
  const long base = (a.axis == 1) ? (long)(active ? xi : 0) + (long)o * p.xDim * p.yDim
                                  : (long)(active ? xi : 0) + (long)o * p.xDim;
  // a lane's points lie at (tid + bpos(0, r)) * es from the tile's base:
  // one per-lane pointer, then per register a uniform offset bpos(0, r) * es
  // (scalar arithmetic) instead of a 64-bit multiply per access
  const cx<T>* __restrict__ in = (const cx<T>*)a.in + base + (long)tid * es;
  cx<T>* __restrict__ out = (cx<T>*)a.out + base + (long)tid * es;
  const auto* __restrict__ tw = (const typename TwE<T>::type*)a.tw;
  const LdsCol la{c, TC};
  constexpr int RF0 = Seq<N, V, ORD_FWD>::first(), RFL = Seq<N, V, ORD_FWD>::last();
  cx<T> v[E];
  double acc = 0.0;
  TwRegs<T, N, V> twr;   // PRE: every twiddle of the sweep, loaded after psi
  // coordinates of element e along the axis
  auto yz = [&](int e, int& yi, int& zi) {
    if (a.axis == 1) { yi = e; zi = o; } else { yi = o; zi = e; }
  };
  // grid indices of register slot r in the boundary layouts of the first / last radix
  auto pos0 = [&](int r) {
    const int e = bpos<N, TPL, RF0>(tid, r);
    return a.axis == 1 ? Pt{xq, e, o} : Pt{xq, o, e};
  };
  auto posL = [&](int r) {
    const int e = bpos<N, TPL, RFL>(tid, r);
    return a.axis == 1 ? Pt{xq, e, o} : Pt{xq, o, e};
  };
  auto gauge_here = [&](cx<T>* w) {
    if (a.axis == 1) {
      if (!gauge_rot_lin<T, N, TPL, RFL, E, 1, X, EQ>(p, w, tid, xq, a.invN)) gauge_all<T, E, 1, X, EQ>(p, w, posL, a.invN);
    } else {
      gauge_all<T, E, 2, X, EQ>(p, w, posL, a.invN);
    }
  };

  if constexpr (MODE == CM_INV || MODE == CM_INV_VG) {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = in[bpos<N, TPL, RFL>(0, r) * es];
    }
    if constexpr (PRE) twr.load(tid, tw);
    fft_regs_tw<T, N, V, ORD_REV, +1, PRE>(v, tid, tw, twr, lds, la, active);
    if constexpr (MODE == CM_INV_VG) {
      // the closing sweep of a step that leads with this axis: V/2, then this
      // axis' gauge operator
      if (flags & F_HALFV_POST) halfV_all<T, E, X, EQ>(p, v, pos0);
      if (flags & F_GAUGE_POST) {
        fft_regs_tw<T, N, V, ORD_FWD, -1, PRE>(v, tid, tw, twr, lds, la, active);
        gauge_here(v);
        fft_regs_tw<T, N, V, ORD_REV, +1, PRE>(v, tid, tw, twr, lds, la, active);
      }
    }
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        out[bpos<N, TPL, RF0>(0, r) * es] = v[r];
        if (MODE == CM_INV_VG && (flags & F_NORM)) acc += abs2d(v[r]);
      }
    }
  } else {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = in[bpos<N, TPL, RF0>(0, r) * es];
    }
    if constexpr (PRE) twr.load(tid, tw);
    if constexpr (MODE == CM_FWD_V) {
      // the opening sweep of a step that leads with this axis
      if (flags & F_SCALE_IN) {
        // a deferred renormalisation (as in the row sweep): the norm reduced
        // here from the previous sweep's partials by every wave, stored once
        double nrm;
        if (flags & F_SCALE_PART) {
          nrm = part_norm(a.partials, a.norm_nb, a.norm_scale);
          if (bid == 0 && threadIdx.x == 0) *a.norm_out = nrm;
        } else {
          nrm = *a.norm_in;
        }
        const T inv = (T)(1.0 / nrm);
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = scale_norm(v[r], inv);
      }
      if (flags & F_HALFV_PRE) halfV_all<T, E, X, EQ>(p, v, pos0);
    }
    auto epos = [&](int r) { return bpos<N, TPL, RFL>(tid, r); };
    fft_regs_tw<T, N, V, ORD_FWD, -1, PRE>(v, tid, tw, twr, lds, la, active);
    if constexpr (MODE == CM_FWD || MODE == CM_FWD_V) {
      if (active) {
#pragma unroll
        for (int r = 0; r < E; ++r) out[bpos<N, TPL, RFL>(0, r) * es] = v[r];
      }
    } else if constexpr (MODE == CM_KSUM) {
      if (active) {
#pragma unroll
        for (int r = 0; r < E; ++r) {
          int yi, zi;
          yz(bpos<N, TPL, RFL>(tid, r), yi, zi);
          const long gid = gid_of(p, xi, yi, zi);
          const double w = (flags & F_ASUM)
                               ? A_val(p, a.axis, xi, yi, zi, gid) * (a.axis == 1 ? p.pyg[yi] : p.pzg[zi])
                           : (flags & F_KAX) ? (a.axis == 1 ? p.ks[1][yi] : p.ks[2][zi])
                                             : K_val(p, xi, yi, zi, gid);
          acc += w * abs2d(v[r]);
        }
      }
    } else {
      if constexpr (MODE == CM_K) {
        if (a.axis == 1) kstep_col<T, E, 1, X, EQ>(p, v, xq, o, epos, a.invN);
        else kstep_col<T, E, 2, X, EQ>(p, v, xq, o, epos, a.invN);
      } else {
        gauge_here(v);
      }
      fft_regs_tw<T, N, V, ORD_REV, +1, PRE>(v, tid, tw, twr, lds, la, active);
      if constexpr (MODE == CM_GAUGE_FWD) {
        // bridge: the next step's V/2 and its first transform along this axis
        if (flags & F_HALFV_PRE) halfV_all<T, E, X, EQ>(p, v, pos0);
        fft_regs_tw<T, N, V, ORD_FWD, -1, PRE>(v, tid, tw, twr, lds, la, active);
        if (active) {
#pragma unroll
          for (int r = 0; r < E; ++r) out[bpos<N, TPL, RFL>(0, r) * es] = v[r];
        }
      } else {
        if (active) {
#pragma unroll
          for (int r = 0; r < E; ++r) {
            out[bpos<N, TPL, RF0>(0, r) * es] = v[r];
            if (flags & F_NORM) acc += abs2d(v[r]);
          }
        }
      }
    }
  }
  if (MODE == CM_KSUM || (flags & F_NORM)) {
    double pv[1] = {acc};
    block_partial<1>(pv, a.partials, sh, bid, nblk);
  }
}

// One problem per launch (arguments by value) ...
template <typename T, int N, int V, int MODE>
__global__ void __launch_bounds__((RowCfg<N, V>::WG)) row_sweep(SweepArgs a) {
  row_body<T, N, V, MODE>(a, blockIdx.x, gridDim.x);
}
// Workgroups are dealt to the 8 XCDs round robin (block b runs on XCD b % 8).
// With narrowed tiles, the G tiles of one default-width tile (one 128-byte
// line per row) are given blocks on the same XCD: block b = x + 8q takes
// tile G*(x + 8*(q/G)) + q%G. Blocks past the last whole group of 8G keep
// their own index, so the map is a permutation of [0, nblk).
template <int G>
__device__ __forceinline__ int xcd_tile(int b, int nblk) {
  if constexpr (G == 1) {
    return b;
  } else {
    const int full = nblk - nblk % (8 * G);
    if (b >= full) return b;
    const int x = b & 7, q = b >> 3;
    return G * (x + 8 * (q / G)) + q % G;
  }
}

template <typename T, int N, int CV, int MODE>
__global__ void __launch_bounds__((ColCfg<T, N, CV>::WG)) col_sweep(SweepArgs a) {
  col_body<T, N, CV, MODE>(a, xcd_tile<ColCfg<T, N, CV>::G>(blockIdx.x, gridDim.x), gridDim.x);
}

// ... or many problems of one shape per launch: blocks [i*bpi, (i+1)*bpi) work
// on instance i, whose arguments are read from device memory.
template <typename T, int N, int V, int MODE>
__global__ void __launch_bounds__((RowCfg<N, V>::WG)) row_sweep_b(const SweepArgs* __restrict__ args, int bpi) {
  const SweepArgs& a = args[blockIdx.x / bpi];
  if (!a.active) return;
  row_body<T, N, V, MODE>(a, blockIdx.x % bpi, bpi);
}
template <typename T, int N, int CV, int MODE>
__global__ void __launch_bounds__((ColCfg<T, N, CV>::WG)) col_sweep_b(const SweepArgs* __restrict__ args, int bpi) {
  const SweepArgs& a = args[blockIdx.x / bpi];
  if (!a.active) return;
  col_body<T, N, CV, MODE>(a, blockIdx.x % bpi, bpi);
}

}  // namespace gpue
