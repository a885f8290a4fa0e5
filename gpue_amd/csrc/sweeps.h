// sweeps.h -- fused split-operator sweeps for gfx950.
//
// One "sweep" reads psi once from HBM, applies a chain of 1-D transforms
// along one axis with GPUE's pointwise operators fused between them, and
// writes psi once.  GPUE's step (src/evolution.cu:727-870) maps onto
//   2D: [V/2, FFTx] [FFTy, K, IFFTy] [IFFTx, V/2 (,x-gauge)] (+gauge sweeps)
//   3D: [V/2, FFTx] [FFTy] [FFTz, K, IFFTz] [IFFTy] [IFFTx, V/2 (,x-gauge)] (+gauge)
//   1D: [V/2, FFT, K, IFFT, V/2 (,gauge)]
// The imaginary-time renormalisation (parSum, src/split_op.cu:83-129) is
// fused: the last sweep of a step leaves per-workgroup partial sums of |psi|^2,
// the last workgroup to finish reduces them in a fixed order to
// norm = sqrt(sum * dV), and the next step's first sweep divides by it.
// Operator fields are regenerated in-kernel from 1-D coordinate arrays
// (harmonic V, rotation A, standard K) or read from tables (equation-string
// or file fields), so the only per-step HBM traffic is psi itself.
#pragma once
#include "fft_core.h"

namespace gpue {

// V: separable 1-D value tables (harmonic trap + 1/2 m |A|^2 of the rotation gauge)
// or a full table (file / equation fields, V re-evaluated per step when time dependent).
enum { V_SEP = 0, V_TABLE = 1 };
enum { A_ROTATION = 0, A_ZERO = 1, A_TABLE = 2 };
enum { K_SEP = 0, K_TABLE = 1 };

// Sweep modes: row (x) and strided (y/z) kernels.
enum { RM_FWD = 0, RM_INV = 1, RM_GAUGE = 2, RM_FULL = 3, RM_KSUM = 4 };
enum { CM_FWD = 0, CM_INV = 1, CM_K = 2, CM_GAUGE = 3, CM_KSUM = 4 };

// Flags of a sweep (runtime-uniform).
enum : unsigned {
  F_SCALE_IN = 1u,     // divide loaded psi by *norm_in (pending renormalisation)
  F_HALFV_PRE = 2u,    // V/2 (+ nonlinearity) before the first transform
  F_HALFV_POST = 4u,   // V/2 (+ nonlinearity) after the inverse transform
  F_GAUGE_POST = 8u,   // extra [FFT, A_axis, IFFT] after the post op
  F_NORM = 16u,        // accumulate |psi_out|^2 and produce norm_out
};

// Time-dependent equation program (postfix): CONST pushes val, X/Y/Z/T push a
// coordinate, OP+k pops (a, b) and pushes GPUE operator k (src/dynamic.cu:568-600).
enum : int { EQ_CONST = 0, EQ_X = 1, EQ_Y = 2, EQ_Z = 3, EQ_T = 4, EQ_OP = 16 };
constexpr int EQ_MAX = 48;
struct EqnProg {
  int n;
  int op[EQ_MAX];
  double val[EQ_MAX];
};

struct Phys {
  int xDim, yDim, zDim, dimnum;
  int gstate, gpe;
  int v_kind, a_kind, k_kind;
  double cV;      // Dt / (2 hbar)          (GV/EV exponent, src/operators.cu:1081,1100)
  double cH;      // (0.5*Dt) / hbar        (cMultDensity, src/kernels.cu:372)
  double g;       // interaction * gDenConst
  double cK;      // Dt / hbar              (GK/EK, src/operators.cu:1082,1102)
  double Dt;      // gauge exponent scale   (GpA/EpA, src/operators.cu:1093-1110)
  double hb2m;    // hbar^2 / (2 mass)      (simple_K, src/operators.cu:467)
  double mass, omega, omegaX, omegaY, omegaZe, gammaY, x0, y0, z0e;
  const double *xg, *yg, *zg, *pxg, *pyg, *pzg;
  const double *vs[3];   // V_SEP: V = vs[0][x] + vs[1][y] + vs[2][z]
  const double *ks[3];   // K_SEP: K = ks[0][kx] + ks[1][ky] + ks[2][kz]
  const double *kf[3];   // K_SEP: per-axis factors of the K operator for this gstate,
                         // interleaved (re, im): exp(-K_a Dt/hbar) or exp(-i K_a Dt/hbar)
  const double *Vt, *Kt, *At[3];
};

struct SweepArgs {
  int active;         // batched launches: 0 skips this instance
  const void* in;
  void* out;
  const void* tw;
  Phys ph;
  int lines;          // row sweeps: yDim*zDim
  int axis;           // 0 x (row), 1 y, 2 z
  int gauge_axis;     // axis of the A operator applied by F_GAUGE_POST / gauge modes
  unsigned flags;
  double invN;        // normalisation folded into K (1/N) or A (1/n_axis)
  const double* norm_in;   // pending renormalisation divisor (F_SCALE_IN)
  double* partials;        // per-workgroup partial sums (F_NORM, KSUM modes)
};

// ---------------------------------------------------------------------------
// Field evaluation (GPUE operators.cu kernels restated per point)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double A_val(const Phys& p, int axis, int xi, int yi, int zi, long gid) {
  if (p.a_kind == A_ROTATION) {           // krotation_Ax/Ay, kconstant_A (:630-647, :621-627)
    if (axis == 0) return -p.yg[yi] * p.omega * p.omegaX;
    if (axis == 1) return p.xg[xi] * p.omega * p.omegaY;
    return 0.0;
  }
  if (p.a_kind == A_ZERO) return 0.0;
  return p.At[axis][gid];
}

// kharmonic_V (src/operators.cu:994-1008) is a sum of per-axis terms once the
// rotation gauge is written out (Ax = Ax(y), Ay = Ay(x), Az = 0), so it is
// carried as three 1-D tables; other fields are full tables.
__device__ __forceinline__ double V_val(const Phys& p, int xi, int yi, int zi, long gid) {
  if (p.v_kind == V_SEP) return p.vs[0][xi] + p.vs[1][yi] + p.vs[2][zi];
  return p.Vt[gid];
}

// simple_K (src/operators.cu:460-469): hbar^2/(2m) (kx^2 + ky^2 + kz^2)
__device__ __forceinline__ double K_val(const Phys& p, int xi, int yi, int zi, long gid) {
  if (p.k_kind == K_SEP) return p.ks[0][xi] + p.ks[1][yi] + p.ks[2][zi];
  return p.Kt[gid];
}

__device__ __forceinline__ long gid_of(const Phys& p, int xi, int yi, int zi) {
  return (long)xi + (long)p.xDim * ((long)yi + (long)p.yDim * (long)zi);
}

// FP64 exp and sincos for the fused operators.  The library versions carry
// general special-case handling that dominates the cost of the V/2 sweeps; here
// the arguments are finite and moderate, so a short form suffices:
// Cody-Waite argument reduction (W. J. Cody and W. Waite, "Software Manual for
// the Elementary Functions", 1980) followed by the Taylor polynomial of the
// reduced argument in Horner form.  The split constants are ln 2 and pi/2
// rounded to 42 and 3x33 significant bits, so k*C_hi is exact for every k that
// occurs.  exp: |r| <= ln2/2, degree 13 (truncation < 5e-18 relative).
// sin/cos: |r| <= pi/4, degrees 17 and 18 (truncation < 1e-19).  Arguments
// outside |x| < 2^20 (sincos) go to the library; NaN propagates.
namespace fm {
constexpr double LN2_HI = 0.6931471805598903;
constexpr double LN2_LO = 5.497923018708371e-14;
constexpr double INV_LN2 = 1.4426950408889634;
constexpr double PIO2_1 = 1.5707963267341256;
constexpr double PIO2_2 = 6.077100506303966e-11;
constexpr double PIO2_3 = 2.0222662487959506e-21;
constexpr double TWO_OVER_PI = 0.6366197723675814;
// 1/k!
constexpr double F2 = 0.5, F3 = 0.16666666666666666, F4 = 0.041666666666666664, F5 = 0.008333333333333333,
                 F6 = 0.001388888888888889, F7 = 0.0001984126984126984, F8 = 2.48015873015873e-05,
                 F9 = 2.7557319223985893e-06, F10 = 2.755731922398589e-07, F11 = 2.505210838544172e-08,
                 F12 = 2.08767569878681e-09, F13 = 1.6059043836821613e-10, F14 = 1.1470745597729725e-11,
                 F15 = 7.647163731819816e-13, F16 = 4.779477332387385e-14, F17 = 2.8114572543455206e-15,
                 F18 = 1.5619206968586225e-16;
}  // namespace fm

__device__ __forceinline__ double exp_cw(double x) {
  using namespace fm;
  x = x < -750.0 ? -750.0 : x;   // below -745.14 the result rounds to 0; comparisons keep NaN
  x = x > 710.0 ? 710.0 : x;
  const double k = rint(x * INV_LN2);
  double r = fma(-k, LN2_HI, x);
  r = fma(-k, LN2_LO, r);
  double p = F13;
  p = fma(p, r, F12); p = fma(p, r, F11); p = fma(p, r, F10); p = fma(p, r, F9); p = fma(p, r, F8);
  p = fma(p, r, F7); p = fma(p, r, F6); p = fma(p, r, F5); p = fma(p, r, F4); p = fma(p, r, F3);
  p = fma(p, r, F2); p = fma(p, r, 1.0); p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

__device__ __forceinline__ void sincos_cw(double x, double* s, double* c) {
  using namespace fm;
  if (!(fabs(x) < 1048576.0)) {   // large, infinite or NaN: library
    sincos(x, s, c);
    return;
  }
  const double q = rint(x * TWO_OVER_PI);
  double r = fma(-q, PIO2_1, x);
  r = fma(-q, PIO2_2, r);
  r = fma(-q, PIO2_3, r);
  const double z = r * r;
  double ps = -F17;
  ps = fma(ps, z, F15); ps = fma(ps, z, -F13); ps = fma(ps, z, F11); ps = fma(ps, z, -F9);
  ps = fma(ps, z, F7); ps = fma(ps, z, -F5); ps = fma(ps, z, F3);
  const double sn = fma(-r * z, ps, r);   // r - r^3/3! + ...
  double pc = F18;
  pc = fma(pc, z, -F16); pc = fma(pc, z, F14); pc = fma(pc, z, -F12); pc = fma(pc, z, F10);
  pc = fma(pc, z, -F8); pc = fma(pc, z, F6); pc = fma(pc, z, -F4); pc = fma(pc, z, F2);
  const double cs = fma(-z, pc, 1.0);     // 1 - r^2/2! + ...
  const int n = ((int)q) & 3;
  const double s0 = (n & 1) ? cs : sn, c0 = (n & 1) ? sn : cs;
  *s = (n & 2) ? -s0 : s0;
  *c = ((n + 1) & 2) ? -c0 : c0;
}

// Transcendentals in the working precision: FP32 sweeps use the FP32 library
// functions, FP64 sweeps the reduced forms above.
template <typename T>
struct WMath;
template <>
struct WMath<double> {
  static __device__ __forceinline__ double ex(double a) { return exp_cw(a); }
  static __device__ __forceinline__ void sc(double a, double* s, double* c) { sincos_cw(a, s, c); }
};
template <>
struct WMath<float> {
  static __device__ __forceinline__ float ex(float a) { return expf(a); }
  static __device__ __forceinline__ void sc(float a, float* s, float* c) { sincosf(a, s, c); }
};

// psi * exp(-arg) (imaginary time) or psi * exp(-i arg) (real time), times s
template <typename T>
__device__ __forceinline__ cx<T> apply_phase(cx<T> v, double arg, int gstate, double s) {
  if (gstate == 0) {
    T f = WMath<T>::ex((T)-arg) * (T)s;
    return cscale(v, f);
  }
  T sn, cs;
  WMath<T>::sc((T)-arg, &sn, &cs);
  cx<T> w = mk((T)(cs * (T)s), (T)(sn * (T)s));
  return cmul(w, v);
}

// V/2 with the nonlinear density term: cMultDensity (src/kernels.cu:365-388) with
// GV/EV (src/operators.cu:1081,1100) merged into one exponential.
template <typename T>
__device__ __forceinline__ cx<T> op_halfV(const Phys& p, cx<T> v, int xi, int yi, int zi) {
  long gid = gid_of(p, xi, yi, zi);
  double V = V_val(p, xi, yi, zi, gid);
  double arg = V * p.cV;
  if (p.gpe) {
    double d = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    arg += p.g * d * p.cH;
  }
  return apply_phase(v, arg, p.gstate, 1.0);
}

// K step: GK/EK (src/operators.cu:1082,1102) with both 1/sqrt(N) of src/evolution.cu:761,777.
// For the separable K the operator is the product of per-axis factors (no
// transcendental per point).
template <typename T>
__device__ __forceinline__ cx<T> op_K(const Phys& p, cx<T> v, int xi, int yi, int zi, double invN) {
  if (p.k_kind == K_SEP) {
    const double* fx = p.kf[0] + 2 * xi;
    const double* fy = p.kf[1] + 2 * yi;
    const double* fz = p.kf[2] + 2 * zi;
    if (p.gstate == 0) return cscale(v, (T)(fx[0] * fy[0] * fz[0] * invN));
    const double ar = fx[0] * fy[0] - fx[1] * fy[1], ai = fx[0] * fy[1] + fx[1] * fy[0];
    const double br = (ar * fz[0] - ai * fz[1]) * invN, bi = (ar * fz[1] + ai * fz[0]) * invN;
    return cmul(mk((T)br, (T)bi), v);
  }
  long gid = gid_of(p, xi, yi, zi);
  return apply_phase(v, K_val(p, xi, yi, zi, gid) * p.cK, p.gstate, invN);
}

// gauge operator along `axis`: GpA*/EpA* (src/operators.cu:1089-1110) with both 1/sqrt(n)
template <typename T>
__device__ __forceinline__ cx<T> op_A(const Phys& p, cx<T> v, int axis, int xi, int yi, int zi, double invn) {
  long gid = gid_of(p, xi, yi, zi);
  double k = axis == 0 ? p.pxg[xi] : axis == 1 ? p.pyg[yi] : p.pzg[zi];
  double pA = A_val(p, axis, xi, yi, zi, gid) * k;
  return apply_phase(v, pA * p.Dt, p.gstate, invn);
}

// ---------------------------------------------------------------------------
// Deterministic two-stage reduction of NA doubles: every workgroup writes its
// partial sums (plain stores, no fences), a one-workgroup finalize kernel
// launched after it sums them in a fixed order.  The result is the same for
// every run and independent of scheduling.
// ---------------------------------------------------------------------------
template <int NA>
__device__ __forceinline__ void block_sum(double (&v)[NA], double* sh /* >= NA * 16 */) {
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    double x = v[a];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    v[a] = x;
  }
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int a = 0; a < NA; ++a) sh[a * 16 + w] = v[a];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      double s = 0.0;
      for (int i = 0; i < nw; ++i) s += sh[a * 16 + i];
      v[a] = s;
    }
  }
}

// Block b of nb (of one problem) stores its NA partial sums.
template <int NA>
__device__ __forceinline__ void block_partial(double (&v)[NA], double* partials, double* sh, unsigned b,
                                              unsigned nb) {
  block_sum<NA>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int a = 0; a < NA; ++a) partials[(size_t)a * nb + b] = v[a];
  }
}

}  // namespace gpue
