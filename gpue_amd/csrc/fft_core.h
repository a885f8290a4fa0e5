// fft_core.h -- register/LDS Stockham FFT building blocks for gfx950 (wave64).
//
// A line of N complex points is held by TPL = N/E threads, E points per thread.
// A transform is a sequence of Stockham passes with radices R_p (each R_p | E).
// Pass p, after L = R_0...R_{p-1} points are combined, m = N/R_p:
//     a_t = x[j + t*m] * w^(t*k),  k = j mod L,  w = exp(sign*2*pi*i/(L*R_p))
//     y[(j/L)*L*R_p + k + s*L] = sum_t a_t * exp(sign*2*pi*i*s*t/R_p)
// Thread `tid` owns butterflies j = tid + TPL*q (q < E/R_p), so its registers
// hold v[q*R_p + t].  Before the first pass (and after the last one) thread
// `tid` holds points  tid + TPL*q + t*(N/R)  -- the "boundary layout" of the
// boundary radix R.  A transform whose first radix equals the previous
// transform's last radix can therefore start from registers: FFT -> pointwise
// op -> IFFT chains never touch LDS at the boundary.  Inner passes exchange
// through LDS.  The radix sequence is walked forwards (ORD_FWD) or backwards
// (ORD_REV); chained transforms alternate the order.
//
// This replaces the cuFFT Z2Z plans of GPUE (src/init.cu:309-319,
// src/ds.cu:11-79): unnormalised, sign -1 forward, +1 inverse.
#pragma once
#include <hip/hip_runtime.h>

namespace gpue {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
// multiply by sign*i
template <typename T, int SIGN>
__device__ __forceinline__ cx<T> mul_si(cx<T> a) {
  if constexpr (SIGN < 0) return mk(a.y, -a.x);
  else return mk(-a.y, a.x);
}

// cos/sin(2*pi*m/48) for m = 0..47 -- covers every radix used here (2,3,4,8,12,16)
struct Tw48 {
  static constexpr double c[48] = {
      1.0, 0.99144486137381041114, 0.96592582628906828675, 0.92387953251128675613,
      0.86602540378443864676, 0.79335334029123516458, 0.70710678118654752440, 0.60876142900872063942,
      0.5, 0.38268343236508977173, 0.25881904510252076235, 0.13052619222005159155,
      0.0, -0.13052619222005159155, -0.25881904510252076235, -0.38268343236508977173,
      -0.5, -0.60876142900872063942, -0.70710678118654752440, -0.79335334029123516458,
      -0.86602540378443864676, -0.92387953251128675613, -0.96592582628906828675, -0.99144486137381041114,
      -1.0, -0.99144486137381041114, -0.96592582628906828675, -0.92387953251128675613,
      -0.86602540378443864676, -0.79335334029123516458, -0.70710678118654752440, -0.60876142900872063942,
      -0.5, -0.38268343236508977173, -0.25881904510252076235, -0.13052619222005159155,
      0.0, 0.13052619222005159155, 0.25881904510252076235, 0.38268343236508977173,
      0.5, 0.60876142900872063942, 0.70710678118654752440, 0.79335334029123516458,
      0.86602540378443864676, 0.92387953251128675613, 0.96592582628906828675, 0.99144486137381041114};
  static __device__ __forceinline__ constexpr double cosv(int m) { return c[((m % 48) + 48) % 48]; }
  static __device__ __forceinline__ constexpr double sinv(int m) { return c[((m - 12) % 48 + 48) % 48]; }
};

// FP32 twiddles are split into hi + lo floats (hi = the rounded value, lo = the
// rounded remainder): the product a*(hi + lo) is formed small terms first, so
// the twiddle's own rounding (a systematic error, the same at every step) drops
// to ~eps^2 and only the random rounding of the product remains.  Measured on a
// model of the 3D imaginary-time step, twiddles rounded once to FP32 give a
// 7.5e-6 relative energy error after 200 steps, exact ones 3e-7.
__device__ __forceinline__ cx<float> cmul_split(cx<float> a, float hx, float hy, float lx, float ly) {
  float re = fmaf(-a.y, ly, a.x * lx);
  float im = fmaf(a.y, lx, a.x * ly);
  re = fmaf(-a.y, hy, re);
  im = fmaf(a.y, hx, im);
  return mk(fmaf(a.x, hx, re), fmaf(a.x, hy, im));
}

// multiply a by exp(SIGN*2*pi*i*m/R), with the exact special cases folded
template <typename T, int R, int SIGN>
__device__ __forceinline__ cx<T> twc(cx<T> a, int m) {
  m = ((m % R) + R) % R;
  if (m == 0) return a;
  if (2 * m == R) return mk(-a.x, -a.y);
  if (4 * m == R) return mul_si<T, SIGN>(a);
  if (4 * m == 3 * R) return mul_si<T, -SIGN>(a);
  const int m48 = m * (48 / R);
  const double cd = Tw48::cosv(m48), sd = SIGN * Tw48::sinv(m48);
  if constexpr (sizeof(T) == 4) {
    if ((8 * m) % R == 0) {
      // odd multiple of 45 degrees: w = c*(sx + i*sy) with c = sqrt(1/2) split,
      // sx, sy = +-1: one exact sign pattern, then two split scalings
      const float sx = cd > 0 ? 1.0f : -1.0f, sy = sd > 0 ? 1.0f : -1.0f;
      const float ch = 0.70710677f, cl = (float)(0.70710678118654752440 - (double)0.70710677f);
      const float u = sx * a.x - sy * a.y, w = sx * a.y + sy * a.x;
      return mk(fmaf(u, ch, u * cl), fmaf(w, ch, w * cl));
    }
    const float ch = (float)cd, sh = (float)sd;
    return cmul_split(a, ch, sh, (float)(cd - (double)ch), (float)(sd - (double)sh));
  } else {
    const T c = (T)cd, s = (T)sd;
    return mk(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// In-place DFT of R points in registers, natural order in and out.
template <typename T, int R, int SIGN>
struct DFT;

template <typename T, int SIGN>
struct DFT<T, 2, SIGN> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    cx<T> t0 = a[0] + a[1], t1 = a[0] - a[1];
    a[0] = t0; a[1] = t1;
  }
};

template <typename T, int SIGN>
struct DFT<T, 3, SIGN> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    const T c = (T)-0.5;
    const double sd = SIGN * 0.86602540378443864676;
    const T s = (T)sd;
    cx<T> sum = a[1] + a[2], dif = a[1] - a[2];
    cx<T> b0 = a[0] + sum;
    cx<T> m = mk(a[0].x + c * sum.x, a[0].y + c * sum.y);
    cx<T> r;   // i*s*dif (s split for FP32, as the twiddles)
    if constexpr (sizeof(T) == 4) {
      const float sl = (float)(sd - (double)s);
      r = mk(fmaf(-s, dif.y, -sl * dif.y), fmaf(s, dif.x, sl * dif.x));
    } else {
      r = mk(-s * dif.y, s * dif.x);
    }
    a[0] = b0; a[1] = m + r; a[2] = m - r;
  }
};

template <typename T, int SIGN>
struct DFT<T, 4, SIGN> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    cx<T> s0 = a[0] + a[2], s1 = a[0] - a[2];
    cx<T> s2 = a[1] + a[3], s3 = mul_si<T, SIGN>(a[1] - a[3]);
    a[0] = s0 + s2; a[2] = s0 - s2; a[1] = s1 + s3; a[3] = s1 - s3;
  }
};

// R = P*Q:  t = t1 + P*t2,  s = Q*s1 + s2
//   b[Q s1 + s2] = sum_t1 wP^(s1 t1) * w^(s2 t1) * sum_t2 a[t1 + P t2] * wQ^(s2 t2)
template <typename T, int P, int Q, int SIGN>
__device__ __forceinline__ void dft_compose(cx<T>* a) {
  constexpr int R = P * Q;
  cx<T> c[P][Q];
#pragma unroll
  for (int t1 = 0; t1 < P; ++t1) {
    cx<T> col[Q];
#pragma unroll
    for (int t2 = 0; t2 < Q; ++t2) col[t2] = a[t1 + P * t2];
    DFT<T, Q, SIGN>::run(col);
#pragma unroll
    for (int s2 = 0; s2 < Q; ++s2) c[t1][s2] = twc<T, R, SIGN>(col[s2], s2 * t1);
  }
#pragma unroll
  for (int s2 = 0; s2 < Q; ++s2) {
    cx<T> row[P];
#pragma unroll
    for (int t1 = 0; t1 < P; ++t1) row[t1] = c[t1][s2];
    DFT<T, P, SIGN>::run(row);
#pragma unroll
    for (int s1 = 0; s1 < P; ++s1) a[Q * s1 + s2] = row[s1];
  }
}

template <typename T, int SIGN>
struct DFT<T, 8, SIGN> {
  static __device__ __forceinline__ void run(cx<T>* a) { dft_compose<T, 4, 2, SIGN>(a); }
};
template <typename T, int SIGN>
struct DFT<T, 12, SIGN> {
  static __device__ __forceinline__ void run(cx<T>* a) { dft_compose<T, 4, 3, SIGN>(a); }
};
template <typename T, int SIGN>
struct DFT<T, 16, SIGN> {
  static __device__ __forceinline__ void run(cx<T>* a) { dft_compose<T, 4, 4, SIGN>(a); }
};

// ---------------------------------------------------------------------------
// Plans: E points per thread and the radix sequence for each supported N.
// Variant 0 ("wide") holds 12-24 points per thread and needs the fewest LDS
// exchanges; variant 1 ("narrow") holds 4-12 points per thread, so a line is
// spread over more lanes -- more parallelism for small, latency-bound grids.
// The autotuner picks one per sweep (gpue_ctx.hip).
// ---------------------------------------------------------------------------
template <int N, int V>
struct Plan;
#define GPUE_PLAN(N_, V_, E_, NP_, ...)                                 \
  template <>                                                           \
  struct Plan<N_, V_> {                                                 \
    static constexpr int N = N_, E = E_, TPL = N_ / E_, NP = NP_;       \
    static constexpr int R[NP_] = {__VA_ARGS__};                        \
  };
GPUE_PLAN(16, 0, 16, 1, 16)
GPUE_PLAN(16, 1, 4, 2, 4, 4)
GPUE_PLAN(32, 0, 16, 2, 16, 2)
GPUE_PLAN(32, 1, 8, 2, 8, 4)
GPUE_PLAN(48, 0, 12, 2, 12, 4)
GPUE_PLAN(48, 1, 12, 3, 3, 4, 4)
GPUE_PLAN(64, 0, 16, 2, 16, 4)
GPUE_PLAN(64, 1, 8, 2, 8, 8)
GPUE_PLAN(96, 0, 24, 2, 12, 8)
GPUE_PLAN(96, 1, 12, 3, 12, 4, 2)
GPUE_PLAN(128, 0, 16, 2, 16, 8)
GPUE_PLAN(128, 1, 8, 3, 8, 8, 2)
GPUE_PLAN(192, 0, 12, 3, 12, 4, 4)
GPUE_PLAN(192, 1, 12, 4, 3, 4, 4, 4)
GPUE_PLAN(256, 0, 16, 2, 16, 16)
GPUE_PLAN(256, 1, 8, 3, 8, 8, 4)
GPUE_PLAN(384, 0, 24, 3, 12, 8, 4)
GPUE_PLAN(384, 1, 12, 4, 12, 4, 4, 2)
GPUE_PLAN(512, 0, 16, 3, 16, 16, 2)
GPUE_PLAN(512, 1, 8, 3, 8, 8, 8)
GPUE_PLAN(1024, 0, 16, 3, 16, 16, 4)
GPUE_PLAN(1024, 1, 8, 4, 8, 8, 8, 2)
GPUE_PLAN(2048, 0, 16, 3, 16, 16, 8)
GPUE_PLAN(2048, 1, 8, 4, 8, 8, 8, 4)
GPUE_PLAN(4096, 0, 16, 3, 16, 16, 16)
GPUE_PLAN(4096, 1, 8, 4, 8, 8, 8, 8)
#undef GPUE_PLAN

enum { ORD_FWD = 0, ORD_REV = 1 };

template <int N, int V, int ORD>
struct Seq {
  using P = Plan<N, V>;
  static constexpr int radix(int p) { return ORD == ORD_FWD ? P::R[p] : P::R[P::NP - 1 - p]; }
  static constexpr int first() { return radix(0); }
  static constexpr int last() { return radix(P::NP - 1); }
  // L before pass p
  static constexpr int Lbefore(int p) {
    int L = 1;
    for (int q = 0; q < p; ++q) L *= radix(q);
    return L;
  }
};

// Point index held in register slot r of thread tid in the boundary layout of radix R.
template <int N, int TPL, int R>
__device__ __forceinline__ constexpr int bpos(int tid, int r) {
  return tid + TPL * (r / R) + (r % R) * (N / R);
}

// Twiddle entries (tables laid out by TwTab): FP64 one complex double; FP32
// (hi.re, hi.im, lo.re, lo.im), lo = the rounded remainder.  SIGN > 0
// conjugates at use.
template <typename T>
struct TwE {
  using type = cx<T>;
};
template <>
struct TwE<float> {
  using type = float4;
};

template <typename T, int SIGN>
__device__ __forceinline__ cx<T> twmul_val(cx<T> a, typename TwE<T>::type w) {
  if constexpr (sizeof(T) == 4) {
    const float sy = SIGN > 0 ? -1.0f : 1.0f;
    return cmul_split(a, w.x, sy * w.y, w.z, sy * w.w);
  } else {
    if constexpr (SIGN > 0) w.y = -w.y;
    return cmul(a, w);
  }
}

// Pass twiddle table of a plan: for the forward order's passes, then the
// reverse order's, every pass P with L > 1 holds L*(R-1) entries, entry
// [t-1][k] = w^(t*k*STEP) with w = exp(-2*pi*i/N), STEP = N/(L*R) (the
// inverse transform conjugates at use).  A lane's twiddles of a pass are
// (t-1)*L entries apart from its own base: one address per pass, the loads
// take immediate offsets, and for each t the lanes (consecutive k) read
// consecutive entries.
template <int N, int V>
struct TwTab {
  template <int ORD>
  static constexpr int cnt(int P) {
    return Seq<N, V, ORD>::Lbefore(P) > 1 ? Seq<N, V, ORD>::Lbefore(P) * (Seq<N, V, ORD>::radix(P) - 1) : 0;
  }
  template <int ORD>
  static constexpr int off(int P) {
    int o = 0;
    if (ORD == ORD_REV)
      for (int q = 0; q < Plan<N, V>::NP; ++q) o += cnt<ORD_FWD>(q);
    for (int q = 0; q < P; ++q) o += cnt<ORD>(q);
    return o;
  }
  static constexpr int total() { return off<ORD_REV>(Plan<N, V>::NP); }
};

// Twiddles a lane needs for a whole transform (PRE): pass P's NB*(R-1)
// entries from off(P).  They depend only on the lane, so a transform can issue
// all of its loads before its first pass: the table reads (L2 hits) then
// overlap pass 0 and the LDS exchanges instead of stalling each pass.
template <int N, int V, int ORD>
struct TwPlan {
  using S = Seq<N, V, ORD>;
  static constexpr int NP = Plan<N, V>::NP, E = Plan<N, V>::E;
  static constexpr int cnt(int P) { return S::Lbefore(P) > 1 ? (E / S::radix(P)) * (S::radix(P) - 1) : 0; }
  static constexpr int off(int P) {
    int o = 0;
    for (int q = 0; q < P; ++q) o += cnt(q);
    return o;
  }
  static constexpr int total() { return off(NP) > 0 ? off(NP) : 1; }
};

template <typename T, int N, int V, int ORD, int P = 0>
__device__ __forceinline__ void tw_load(int tid, const typename TwE<T>::type* __restrict__ tw,
                                        typename TwE<T>::type* w) {
  using TP = TwPlan<N, V, ORD>;
  if constexpr (P < TP::NP) {
    constexpr int R = TP::S::radix(P), L = TP::S::Lbefore(P), E = TP::E, NB = E / R, TPL = N / E;
    if constexpr (L > 1) {
      const typename TwE<T>::type* tp = tw + TwTab<N, V>::template off<ORD>(P);
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int k = (tid + TPL * q) % L;
#pragma unroll
        for (int t = 1; t < R; ++t) w[TP::off(P) + q * (R - 1) + (t - 1)] = tp[(t - 1) * L + k];
      }
    }
    tw_load<T, N, V, ORD, P + 1>(tid, tw, w);
  }
}

// Compute pass p (radix R, L) on registers v[E] of thread tid.  PRE: the
// twiddles were loaded into w (from TwPlan::off(p)); otherwise they are read
// here from the pass's table tp (fewer registers live across the passes).
template <typename T, int N, int E, int R, int L, int SIGN, bool PRE>
__device__ __forceinline__ void pass_compute(cx<T>* v, int tid, const typename TwE<T>::type* __restrict__ tp,
                                             const typename TwE<T>::type* w) {
  constexpr int TPL = N / E, NB = E / R;
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    cx<T>* a = v + q * R;
    if constexpr (L > 1) {
      const int k = (tid + TPL * q) % L;
      const typename TwE<T>::type* tk = tp + k;
#pragma unroll
      for (int t = 1; t < R; ++t) a[t] = twmul_val<T, SIGN>(a[t], PRE ? w[q * (R - 1) + (t - 1)] : tk[(t - 1) * L]);
    }
    DFT<T, R, SIGN>::run(a);
  }
}

// Output position of register slot (q, s) after pass (R, L).
template <int N, int TPL, int R, int L>
__device__ __forceinline__ int pass_out(int tid, int q, int s) {
  const int j = tid + TPL * q;
  const int k = j % L;
  return (j / L) * L * R + k + s * L;
}

// LDS addressing: element e of a line.  Row lines: base + pad(e); strided
// lines: (e * LS + c) where LS = lines per tile (lanes vary in c fastest).
struct LdsRow {
  int base;
  __device__ __forceinline__ int at(int e) const { return base + e + (e >> 4); }
  static constexpr int footprint(int N) { return N + N / 16; }
};
struct LdsCol {
  int c, LS;
  __device__ __forceinline__ int at(int e) const { return e * LS + c; }
};

// Full transform on registers: input in boundary layout of first radix,
// output in boundary layout of last radix.  Inner passes exchange via LDS.
// `active` predicates LDS traffic for padding threads (they still hit barriers).
template <typename T, int N, int V, int ORD, int SIGN, bool PRE, int P, typename LA>
__device__ __forceinline__ void fft_from_pass(cx<T>* v, int tid, const typename TwE<T>::type* __restrict__ tw,
                                              const typename TwE<T>::type* w, cx<T>* lds, const LA& la,
                                              bool active) {
  using S = Seq<N, V, ORD>;
  constexpr int NP = Plan<N, V>::NP, E = Plan<N, V>::E, TPL = Plan<N, V>::TPL;
  if constexpr (P < NP) {
    if constexpr (P > 0) {
      constexpr int RP = S::radix(P - 1), LP = S::Lbefore(P - 1);
      __syncthreads();  // previous users of LDS are done
      if (active) {
#pragma unroll
        for (int r = 0; r < E; ++r) lds[la.at(pass_out<N, TPL, RP, LP>(tid, r / RP, r % RP))] = v[r];
      }
      __syncthreads();
      if (active) {
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = lds[la.at(bpos<N, TPL, S::radix(P)>(tid, r))];
      }
    }
    pass_compute<T, N, E, S::radix(P), S::Lbefore(P), SIGN, PRE>(v, tid, tw + TwTab<N, V>::template off<ORD>(P),
                                                                 w + TwPlan<N, V, ORD>::off(P));
    fft_from_pass<T, N, V, ORD, SIGN, PRE, P + 1>(v, tid, tw, w, lds, la, active);
  }
}

// PRE: load every twiddle of the transform before its first pass (for
// latency-bound launches with registers to spare); otherwise per pass.
template <typename T, int N, int V, int ORD, int SIGN, bool PRE, typename LA>
__device__ __forceinline__ void fft_regs(cx<T>* v, int tid, const typename TwE<T>::type* __restrict__ tw, cx<T>* lds,
                                         const LA& la,
                                         bool active) {
  typename TwE<T>::type w[PRE ? TwPlan<N, V, ORD>::total() : 1];
  if constexpr (PRE) tw_load<T, N, V, ORD>(tid, tw, w);
  fft_from_pass<T, N, V, ORD, SIGN, PRE, 0>(v, tid, tw, w, lds, la, active);
}

// The twiddles of both orders, loaded once at the start of a sweep (PRE
// variants): every transform of the sweep then runs on registers, so a second
// or third transform does not wait for its table reads.
template <typename T, int N, int V>
struct TwRegs {
  typename TwE<T>::type f[TwPlan<N, V, ORD_FWD>::total()];
  typename TwE<T>::type r[TwPlan<N, V, ORD_REV>::total()];
  __device__ __forceinline__ void load(int tid, const typename TwE<T>::type* __restrict__ tw) {
    tw_load<T, N, V, ORD_FWD>(tid, tw, f);
    tw_load<T, N, V, ORD_REV>(tid, tw, r);
  }
};

// fft_regs on preloaded twiddles (PRE) or from the table.
template <typename T, int N, int V, int ORD, int SIGN, bool PRE, typename LA>
__device__ __forceinline__ void fft_regs_tw(cx<T>* v, int tid, const typename TwE<T>::type* __restrict__ tw,
                                            const TwRegs<T, N, V>& wr, cx<T>* lds, const LA& la, bool active) {
  if constexpr (PRE)
    fft_from_pass<T, N, V, ORD, SIGN, true, 0>(v, tid, tw, ORD == ORD_FWD ? wr.f : wr.r, lds, la, active);
  else
    fft_from_pass<T, N, V, ORD, SIGN, false, 0>(v, tid, tw, wr.f, lds, la, active);
}

}  // namespace gpue
