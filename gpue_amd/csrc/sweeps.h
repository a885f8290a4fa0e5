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
// or a full table (file / static equation fields).  Time-dependent equation
// fields are evaluated per point inside the step's sweeps (the EQ kernels):
// V_EQN, K_EQN and A_EQN; V_TRAP_A is the harmonic trap plus 1/2 m |A|^2 of a
// time-dependent A (GPUE builds V after A, src/operators.cu:726-829).
enum { V_SEP = 0, V_TABLE = 1, V_EQN = 2, V_TRAP_A = 3 };
enum { A_ROTATION = 0, A_ZERO = 1, A_TABLE = 2, A_EQN = 3 };
enum { K_SEP = 0, K_TABLE = 1, K_EQN = 2 };

// Sweep modes: row (x) and strided (y/z) kernels.
// RM_INV_FWD / RM_GAUGE_FWD / CM_GAUGE_FWD: the last sweep of a real-time step
// fused with the first of the next one on the same axis:
// ..., [V/2 of step i+1 (F_HALFV_PRE)], FFT.
// RM_K: FFT, K, IFFT along x (a step that leads with the y axis).
// CM_FWD_V: [scale] [V/2] FFT along y/z (a step that leads with that axis);
// CM_INV_VG: IFFT [V/2] [FFT, A, IFFT] along y/z (its closing sweep).
enum { RM_FWD = 0, RM_INV = 1, RM_GAUGE = 2, RM_FULL = 3, RM_KSUM = 4, RM_INV_FWD = 5, RM_GAUGE_FWD = 6, RM_K = 7 };
enum { CM_FWD = 0, CM_INV = 1, CM_K = 2, CM_GAUGE = 3, CM_KSUM = 4, CM_FWD_V = 5, CM_INV_VG = 6, CM_GAUGE_FWD = 7 };

// Strided-sweep modes that apply a phase operator (FP32 real time runs their
// bit-16 kernels).
__host__ __device__ inline bool col_phase_mode(int mode) {
  return mode == CM_K || mode == CM_GAUGE || mode == CM_FWD_V || mode == CM_INV_VG || mode == CM_GAUGE_FWD;
}

// Flags of a sweep (runtime-uniform).
enum : unsigned {
  F_SCALE_IN = 1u,     // divide loaded psi by *norm_in (pending renormalisation)
  F_HALFV_PRE = 2u,    // V/2 (+ nonlinearity) before the first transform
  F_HALFV_POST = 4u,   // V/2 (+ nonlinearity) after the inverse transform
  F_GAUGE_POST = 8u,   // extra [FFT, A_axis, IFFT] after the post op
  F_NORM = 16u,        // accumulate |psi_out|^2 and produce norm_out
  F_SCALE_PART = 32u,  // with F_SCALE_IN: the norm is reduced here from the previous
                       // sweep's partials (norm_nb of them), and block 0 stores it
                       // to norm_out, instead of a separate finalize
  F_DEFER = 64u,       // host only: the finalize of this F_NORM sweep is left to
                       // the next sweep (F_SCALE_PART)
  F_ASUM = 128u,       // KSUM modes: weight A_axis * k_axis (the gauge term of the
                       // evolved Hamiltonian) instead of K
  F_KAX = 256u,        // KSUM modes: weight ks[axis][k_axis], this axis's part of a
                       // separable K (Parseval along the other axes needs no transform)
};

// Time-dependent equation program (postfix): CONST pushes val, X/Y/Z/T push a
// coordinate, OP+k pops (a, b) and pushes GPUE operator k (src/dynamic.cu:568-600).
enum : int { EQ_CONST = 0, EQ_X = 1, EQ_Y = 2, EQ_Z = 3, EQ_T = 4, EQ_OP = 16 };
constexpr int EQ_MAX = 48;
struct EqnProg {
  int n;
  int depth;   // largest stack depth the program reaches
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
  const float *vsf[3];   // FP32 contexts: vs and kf rounded to FP32 once (the FP32
  const float *kff[3];   // sweeps read these; null for FP64 contexts)
  const double *Vt, *Kt, *At[3];
  // time-dependent fields (read by the EQ kernels only)
  const EqnProg* eq[5];   // programs of V, K, Ax, Ay, Az; null where the field is static
  const double* vst[3];   // V_TRAP_A: the trap's per-axis terms without 1/2 m |A|^2
  double t;               // time of the step, Dt * i (src/evolution.cu:408)
  double org[3];          // equation coordinates x = dx * xi - xMax (src/dynamic.cu:613-615)
  double dxyz[3];
  int eqk;                // 1: the step runs the EQ kernels
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
  int norm_nb;             // F_SCALE_PART: partials of the previous sweep
  double norm_scale;       // F_SCALE_PART: dV
  double* norm_out;        // F_SCALE_PART: where block 0 stores the norm
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

// |x| <= 1/32: Taylor polynomials without a reduction (the V/2 phases of
// real time and most imaginary-time exponents are this small); the first
// omitted terms, x^9/9! and r^9/9!, r^10/10!, are below 1e-18 relative.
constexpr double SMALL_ARG = 0.03125;
__device__ __forceinline__ double exp_small(double x) {
  using namespace fm;
  double p = F8;
  p = fma(p, x, F7); p = fma(p, x, F6); p = fma(p, x, F5); p = fma(p, x, F4); p = fma(p, x, F3);
  p = fma(p, x, F2); p = fma(p, x, 1.0);
  return fma(p, x, 1.0);
}
__device__ __forceinline__ void sincos_small(double r, double* sn_out, double* cs_out) {
  using namespace fm;
  const double z = r * r;
  double ps = -F7;
  ps = fma(ps, z, F5); ps = fma(ps, z, -F3);
  *sn_out = fma(r * z, ps, r);            // r - r^3/3! + r^5/5! - r^7/7!
  double pc = F8;
  pc = fma(pc, z, -F6); pc = fma(pc, z, F4); pc = fma(pc, z, -F2);
  *cs_out = fma(z, pc, 1.0);              // 1 - r^2/2! + ... + r^8/8!
}

// sin and cos of |r| <= pi/4 (no reduction)
__device__ __forceinline__ void sincos_poly(double r, double* sn_out, double* cs_out) {
  using namespace fm;
  const double z = r * r;
  double ps = -F17;
  ps = fma(ps, z, F15); ps = fma(ps, z, -F13); ps = fma(ps, z, F11); ps = fma(ps, z, -F9);
  ps = fma(ps, z, F7); ps = fma(ps, z, -F5); ps = fma(ps, z, F3);
  const double sn = fma(-r * z, ps, r);   // r - r^3/3! + ...
  double pc = F18;
  pc = fma(pc, z, -F16); pc = fma(pc, z, F14); pc = fma(pc, z, -F12); pc = fma(pc, z, F10);
  pc = fma(pc, z, -F8); pc = fma(pc, z, F6); pc = fma(pc, z, -F4); pc = fma(pc, z, F2);
  const double cs = fma(-z, pc, 1.0);     // 1 - r^2/2! + ...
  *sn_out = sn;
  *cs_out = cs;
}

// core for |x| < 2^20 (no branch, so the polynomial chains of several points
// can be interleaved); sincos_cw adds the library path for other arguments
__device__ __forceinline__ void sincos_core(double x, double* s, double* c) {
  using namespace fm;
  const double q = rint(x * TWO_OVER_PI);
  double r = fma(-q, PIO2_1, x);
  r = fma(-q, PIO2_2, r);
  r = fma(-q, PIO2_3, r);
  double sn, cs;
  sincos_poly(r, &sn, &cs);
  const int n = ((int)q) & 3;
  const double s0 = (n & 1) ? cs : sn, c0 = (n & 1) ? sn : cs;
  *s = (n & 2) ? -s0 : s0;
  *c = ((n + 1) & 2) ? -c0 : c0;
}

__device__ __forceinline__ bool sincos_small_ok(double x) { return fabs(x) <= 0.78539816339744828; }

__device__ __forceinline__ bool sincos_fast_ok(double x) { return fabs(x) < 1048576.0; }

__device__ __forceinline__ void sincos_cw(double x, double* s, double* c) {
  if (!sincos_fast_ok(x)) {   // large, infinite or NaN: library
    sincos(x, s, c);
    return;
  }
  sincos_core(x, s, c);
}

// FP32 sincos in the same form: pi/2 split into 12+12 significant bits and a
// rounded tail (q*P1, q*P2 exact for |q| < 2^12), Taylor degrees 9 and 10 on
// |r| <= pi/4 (truncation < 2e-9, below FP32 rounding); |x| >= 6000, infinities
// and NaN go to the library.
__device__ __forceinline__ void sincos_poly(float r, float* sn_out, float* cs_out) {
  const float z = r * r;
  float ps = 2.7557318844628753e-06f;
  ps = fmaf(ps, z, -0.00019841270113829523f);
  ps = fmaf(ps, z, 0.008333333767950535f);
  ps = fmaf(ps, z, -0.1666666716337204f);
  const float sn = fmaf(r * z, ps, r);
  float pc = -2.755731998149713e-07f;
  pc = fmaf(pc, z, 2.4801587642286904e-05f);
  pc = fmaf(pc, z, -0.0013888889225199819f);
  pc = fmaf(pc, z, 0.0416666679084301f);
  pc = fmaf(pc, z, -0.5f);
  const float cs = fmaf(z, pc, 1.0f);
  *sn_out = sn;
  *cs_out = cs;
}

__device__ __forceinline__ void sincos_core(float x, float* s, float* c) {
  const float q = rintf(x * 0.6366197466850281f);
  float r = fmaf(-q, 1.57080078125f, x);
  r = fmaf(-q, -4.453584551811218e-06f, r);
  r = fmaf(-q, -8.705515752716053e-10f, r);
  float sn, cs;
  sincos_poly(r, &sn, &cs);
  const int n = ((int)q) & 3;
  const float s0 = (n & 1) ? cs : sn, c0 = (n & 1) ? sn : cs;
  *s = (n & 2) ? -s0 : s0;
  *c = ((n + 1) & 2) ? -c0 : c0;
}

__device__ __forceinline__ bool sincos_small_ok(float x) { return fabsf(x) <= 0.78539816f; }

__device__ __forceinline__ bool sincos_fast_ok(float x) { return fabsf(x) < 6000.0f; }

__device__ __forceinline__ void sincos_cw(float x, float* s, float* c) {
  if (!sincos_fast_ok(x)) {
    sincosf(x, s, c);
    return;
  }
  sincos_core(x, s, c);
}

// Transcendentals in the working precision: FP64 sweeps use the reduced forms
// above; FP32 sweeps the FP32 library exp and the FP32 reduced sincos.
template <typename T>
struct WMath;
template <>
struct WMath<double> {
  static __device__ __forceinline__ double ex(double a) { return exp_cw(a); }
  static __device__ __forceinline__ void sc(double a, double* s, double* c) { sincos_cw(a, s, c); }
  static __device__ __forceinline__ void sc_core(double a, double* s, double* c) { sincos_core(a, s, c); }
  static __device__ __forceinline__ bool sc_ok(double a) { return sincos_fast_ok(a); }
  static __device__ __forceinline__ bool sc_small_ok(double a) { return sincos_small_ok(a); }
  static __device__ __forceinline__ void sc_small(double a, double* s, double* c) { sincos_poly(a, s, c); }
};
template <>
struct WMath<float> {
  // hardware exp2 of a*log2(e) (__expf): the FP32 step rounds its operands to
  // FP32 anyway, and |a| stays small (V dt/2, K dt), so the product's rounding
  // adds |a|*2^-24 relative
  static __device__ __forceinline__ float ex(float a) { return __expf(a); }
  static __device__ __forceinline__ void sc(float a, float* s, float* c) { sincos_cw(a, s, c); }
  static __device__ __forceinline__ void sc_core(float a, float* s, float* c) { sincos_core(a, s, c); }
  static __device__ __forceinline__ bool sc_ok(float a) { return sincos_fast_ok(a); }
  static __device__ __forceinline__ bool sc_small_ok(float a) { return sincos_small_ok(a); }
  static __device__ __forceinline__ void sc_small(float a, float* s, float* c) { sincos_poly(a, s, c); }
};

// ---------------------------------------------------------------------------
// Time-dependent equation fields, evaluated per point in the EQ kernels.
// GPUE's operators (src/dynamic.cu:568-600): 1 +, 2 -, 3 *, 4 /, 5 pow with an
// integer exponent, 6 cos, 7 sin, 8 tan, 9 sqrt, 10 exp, 11 jn, 12 yn.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double eqn_apply(int k, double a, double b) {
  switch (k) {
    case 1: return a + b;
    case 2: return a - b;
    case 3: return a * b;
    case 4: return a / b;
    case 5: return pow(a, (int)b);
    case 6: return cos(a);
    case 7: return sin(a);
    case 8: return tan(a);
    case 9: return sqrt(a);
    case 10: return exp(a);
    case 11: { const int n = (int)a; return n == 0 ? j0(b) : n == 1 ? j1(b) : jn(n, b); }
    case 12: { const int n = (int)a; return n == 0 ? y0(b) : n == 1 ? y1(b) : yn(n, b); }
    default: return 0.0;
  }
}

// A postfix program with its stack in registers: EQ_REG_DEPTH named slots,
// shifted on every push and pop so that each index is a compile-time constant
// (a dynamically indexed array would live in scratch memory).  The host rejects
// programs that reach a deeper stack.  The op sequence is the same for every
// lane, so the dispatch branches are uniform.
// One out-of-line copy per translation unit: the EQ kernels call it per point
// (inlined E times over every kernel it would multiply their code and compile
// time for no gain; these kernels are not on the path of static fields).
constexpr int EQ_REG_DEPTH = 8;
__device__ __noinline__ double eqn_run(const EqnProg* __restrict__ p, double x, double y, double z, double t) {
  double s[EQ_REG_DEPTH];
#pragma unroll
  for (int k = 0; k < EQ_REG_DEPTH; ++k) s[k] = 0.0;
  const int n = p->n;
  for (int i = 0; i < n; ++i) {
    const int op = p->op[i];
    if (op < EQ_OP) {
      const double v = op == EQ_CONST ? p->val[i] : op == EQ_X ? x : op == EQ_Y ? y : op == EQ_Z ? z : t;
#pragma unroll
      for (int k = EQ_REG_DEPTH - 1; k > 0; --k) s[k] = s[k - 1];
      s[0] = v;
    } else {
      const double r = eqn_apply(op - EQ_OP, s[1], s[0]);
#pragma unroll
      for (int k = 1; k + 1 < EQ_REG_DEPTH; ++k) s[k] = s[k + 1];
      s[0] = r;
    }
  }
  return s[0];
}

struct Pt {
  int x, y, z;
};

__device__ __forceinline__ double eq_field(const Phys& p, const EqnProg* prog, Pt q) {
  return eqn_run(prog, p.dxyz[0] * q.x + p.org[0], p.dxyz[1] * q.y + p.org[1], p.dxyz[2] * q.z + p.org[2], p.t);
}

// A_axis of an A_EQN context: the program where the component depends on t,
// else its table (components are tables once any of them is time dependent)
__device__ __forceinline__ double A_eq(const Phys& p, int axis, Pt q, long gid) {
  if (p.eq[2 + axis]) return eq_field(p, p.eq[2 + axis], q);
  return p.At[axis] ? p.At[axis][gid] : 0.0;
}

// V of a V_EQN or V_TRAP_A context
__device__ __forceinline__ double V_eq(const Phys& p, Pt q, long gid) {
  if (p.v_kind == V_EQN) return eq_field(p, p.eq[0], q);
  const double A0 = A_eq(p, 0, q, gid), A1 = A_eq(p, 1, q, gid), A2 = A_eq(p, 2, q, gid);
  return (p.vst[0][q.x] + p.vst[1][q.y] + p.vst[2][q.z]) + 0.5 * p.mass * (A0 * A0 + A1 * A1 + A2 * A2);
}

// ---------------------------------------------------------------------------
// Batched forms of the pointwise operators for the E points a lane holds.  All
// table loads of the E points are issued before any arithmetic and every
// branch on a field kind or gstate is taken once, outside the point loop:
// per-point branches around loads serialise one memory round trip per point.
// pos(r) gives the (x, y, z) grid indices of register slot r.
// ---------------------------------------------------------------------------
// The operator arithmetic runs in the working precision T: FP64 sweeps in
// double; FP32 sweeps in float, with the table values (held in double) rounded
// once at load.  Only the reductions stay in FP64 in the FP32 form.

// psi * exp(-arg) (imaginary time) or psi * exp(-i arg) (real time), times s
// X (FP32 real-time kernels only): the phase factor is formed in FP32, then its
// modulus is corrected in FP64 by one Newton step towards 1/sqrt(c^2 + s^2),
// and the product is formed in FP64 and rounded to FP32 once.  A factor
// rounded to FP32 carries a modulus error of ~1e-7 that is the same at every
// step for static V, A and K, so the norm and energy would drift linearly with
// the steps; after the Newton step the modulus is 1 within ~1e-14, and what is
// left (the FP32 phase, ~1e-7 relative) acts like a rounded V and does not grow.
// Separate kernels, so the imaginary-time ones keep their registers.
template <typename T, int E, bool X = false>
__device__ __forceinline__ void apply_phase_all(cx<T>* v, const T (&arg)[E], int gstate, double s_) {
  const T s = (T)s_;
  // FP64: a lane whose arguments are all small takes the short polynomials
  bool tiny = false;
  if constexpr (sizeof(T) == 8) {
    tiny = true;
#pragma unroll
    for (int r = 0; r < E; ++r) tiny = tiny && fabs(arg[r]) <= SMALL_ARG;
  }
  if (gstate == 0) {
    if constexpr (sizeof(T) == 8) {
      if (tiny) {
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = cscale(v[r], exp_small(-arg[r]) * s);
        return;
      }
    }
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const T f = WMath<T>::ex(-arg[r]) * s;
      v[r] = cscale(v[r], f);
    }
  } else if constexpr (X && sizeof(T) == 4) {
    bool small = true, fast = true;
#pragma unroll
    for (int r = 0; r < E; ++r) {
      small = small && sincos_small_ok(-arg[r]);
      fast = fast && sincos_fast_ok(-arg[r]);
    }
    float sn[E], cs[E];
    if (small) {
#pragma unroll
      for (int r = 0; r < E; ++r) sincos_poly(-arg[r], &sn[r], &cs[r]);
    } else if (fast) {
#pragma unroll
      for (int r = 0; r < E; ++r) sincos_core(-arg[r], &sn[r], &cs[r]);
    } else {
#pragma unroll
      for (int r = 0; r < E; ++r) sincos_cw(-arg[r], &sn[r], &cs[r]);
    }
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const double c0 = cs[r], d0 = sn[r];
      const double f = fma(-0.5, fma(c0, c0, d0 * d0), 1.5) * s_;
      const double c = c0 * f, d = d0 * f, x = v[r].x, y = v[r].y;
      v[r] = mk((float)(c * x - d * y), (float)(c * y + d * x));
    }
  } else {
    // one range check for the lane's E arguments, then branch-free cores; a
    // real-time phase is usually within pi/4 (no reduction needed)
    bool small = true, fast = true;
#pragma unroll
    for (int r = 0; r < E; ++r) {
      small = small && WMath<T>::sc_small_ok(-arg[r]);
      fast = fast && WMath<T>::sc_ok(-arg[r]);
    }
    T sn[E], cs[E];
    if constexpr (sizeof(T) == 8) {
      if (tiny) {
#pragma unroll
        for (int r = 0; r < E; ++r) sincos_small(-arg[r], &sn[r], &cs[r]);
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = cmul(mk(cs[r] * s, sn[r] * s), v[r]);
        return;
      }
    }
    if (small) {
#pragma unroll
      for (int r = 0; r < E; ++r) WMath<T>::sc_small(-arg[r], &sn[r], &cs[r]);
    } else if (fast) {
#pragma unroll
      for (int r = 0; r < E; ++r) WMath<T>::sc_core(-arg[r], &sn[r], &cs[r]);
    } else {
#pragma unroll
      for (int r = 0; r < E; ++r) WMath<T>::sc(-arg[r], &sn[r], &cs[r]);
    }
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = cmul(mk(cs[r] * s, sn[r] * s), v[r]);
  }
}

// V/2 with the nonlinear density term: cMultDensity (src/kernels.cu:365-388)
// with GV/EV (src/operators.cu:1081,1100) merged into one exponential.
// In two parts: halfV_vargs gathers V*cV of the lane's points, halfV_apply
// adds the density term and applies the phase (halfV_all does both in a row).
template <typename T, int E, bool EQ = false, typename F>
__device__ __forceinline__ void halfV_vargs(const Phys& p, F pos, T (&arg)[E]) {
  if (EQ && (p.v_kind == V_EQN || p.v_kind == V_TRAP_A)) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const Pt q = pos(r);
      arg[r] = (T)V_eq(p, q, gid_of(p, q.x, q.y, q.z));
    }
  } else if (p.v_kind == V_SEP) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const Pt q = pos(r);
      if constexpr (sizeof(T) == 4)
        arg[r] = p.vsf[0][q.x] + p.vsf[1][q.y] + p.vsf[2][q.z];
      else
        arg[r] = p.vs[0][q.x] + p.vs[1][q.y] + p.vs[2][q.z];
    }
  } else {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const Pt q = pos(r);
      arg[r] = (T)p.Vt[gid_of(p, q.x, q.y, q.z)];
    }
  }
  const T cV = (T)p.cV;
#pragma unroll
  for (int r = 0; r < E; ++r) arg[r] = arg[r] * cV;
}

template <typename T, int E, bool X = false>
__device__ __forceinline__ void halfV_apply(const Phys& p, cx<T>* v, T (&arg)[E]) {
  if (p.gpe) {
    const T g = (T)p.g, cH = (T)p.cH;
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const T d = v[r].x * v[r].x + v[r].y * v[r].y;
      arg[r] += g * d * cH;
    }
  }
  apply_phase_all<T, E, X>(v, arg, p.gstate, 1.0);
}

template <typename T, int E, bool X = false, bool EQ = false, typename F>
__device__ __forceinline__ void halfV_all(const Phys& p, cx<T>* v, F pos) {
  T arg[E];
  halfV_vargs<T, E, EQ>(p, pos, arg);
  halfV_apply<T, E, X>(p, v, arg);
}

// gauge operator along AXIS: GpA*/EpA* (src/operators.cu:1089-1110) with both
// 1/sqrt(n) of the axis transforms folded into invn.  gauge_args gathers the
// exponents A*k*Dt of the lane's points; apply_phase_all applies the phase.
template <typename T, int E, int AXIS, bool EQ = false, typename F>
__device__ __forceinline__ void gauge_args(const Phys& p, F pos, T (&arg)[E]) {
  if (EQ && p.a_kind == A_EQN) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const Pt q = pos(r);
      arg[r] = (T)A_eq(p, AXIS, q, gid_of(p, q.x, q.y, q.z));
    }
  } else if (p.a_kind == A_TABLE) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const Pt q = pos(r);
      arg[r] = (T)p.At[AXIS][gid_of(p, q.x, q.y, q.z)];
    }
  } else {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const Pt q = pos(r);
      double A = 0.0;   // A_ZERO, and A_z of the rotation
      if (p.a_kind == A_ROTATION) {
        if constexpr (AXIS == 0) A = -p.yg[q.y] * p.omega * p.omegaX;
        if constexpr (AXIS == 1) A = p.xg[q.x] * p.omega * p.omegaY;
      }
      arg[r] = (T)A;
    }
  }
  const T Dt = (T)p.Dt;
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const Pt q = pos(r);
    const T k = (T)(AXIS == 0 ? p.pxg[q.x] : AXIS == 1 ? p.pyg[q.y] : p.pzg[q.z]);
    arg[r] = arg[r] * k * Dt;
  }
}

template <typename T, int E, int AXIS, bool X = false, bool EQ = false, typename F>
__device__ __forceinline__ void gauge_all(const Phys& p, cx<T>* v, F pos, double invn) {
  T arg[E];
  gauge_args<T, E, AXIS, EQ>(p, pos, arg);
  apply_phase_all<T, E, X>(v, arg, p.gstate, invn);
}

// The rotation gauge on a line along AXIS (x: A_x = -y omega omegaX; y:
// A_y = x omega omegaY; both constant along the line) multiplies the point of
// signed wave-number index m by exp(-a m) (imaginary time) or exp(-i a m) (real
// time), a = A dk Dt, dk = k[1].  In the boundary layout of radix R a lane holds
// the points tid + TPL q + (N/R) s (r = q R + s), less N from the Nyquist index
// on, so the E factors are products of four per-lane factors and their powers:
// exp(-a tid), exp(-a N/R), exp(-a TPL) and exp(a N), i.e. three or four
// exponentials per lane instead of E (each product adds a rounding: a few ulp).
// FP64 kernels only; false (nothing applied) for other gauges, for FP32, or when
// |a| N is large enough for exp(a N) to leave the double range.
template <typename T, int N, int TPL, int R, int E, int AXIS, bool X, bool EQ>
__device__ __forceinline__ bool gauge_rot_lin(const Phys& p, cx<T>* v, int tid, int other, double invn) {
  if constexpr (sizeof(T) != 8 || X || AXIS == 2) {
    return false;
  } else {
    if (p.a_kind != A_ROTATION) return false;
    const double A = AXIS == 0 ? -p.yg[other] * p.omega * p.omegaX : p.xg[other] * p.omega * p.omegaY;
    const double dk = AXIS == 0 ? p.pxg[1] : p.pyg[1];
    const double a = A * dk * p.Dt;
    if (!(fabs(a) * N <= 600.0)) return false;
    constexpr int Q = E / R, NR = N / R;
    if (p.gstate == 0) {
      double w[R];
      w[0] = WMath<double>::ex(-a * tid) * invn;
      const double W = WMath<double>::ex(-a * NR), Z = WMath<double>::ex(a * N);
      const double U = Q > 1 ? WMath<double>::ex(-a * TPL) : 1.0;
#pragma unroll
      for (int k = 1; k < R; ++k) w[k] = w[k - 1] * W;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const int r = q * R + k;
          const double f = (tid + TPL * q + NR * k >= N / 2) ? w[k] * Z : w[k];
          v[r] = cscale(v[r], f);
        }
        if (q + 1 < Q) {
#pragma unroll
          for (int k = 0; k < R; ++k) w[k] *= U;
        }
      }
    } else {
      // exp(-i x) = (cos x, -sin x)
      auto ph = [](double x) {
        double sn, cs;
        WMath<double>::sc(x, &sn, &cs);
        return mk(cs, -sn);
      };
      cx<double> w[R];
      const cx<double> b = ph(a * tid);
      w[0] = mk(b.x * invn, b.y * invn);
      const cx<double> W = ph(a * NR), Zc = ph(a * N);
      const cx<double> Z = mk(Zc.x, -Zc.y);   // exp(+i a N)
      const cx<double> U = Q > 1 ? ph(a * TPL) : mk(1.0, 0.0);
#pragma unroll
      for (int k = 1; k < R; ++k) w[k] = cmul(w[k - 1], W);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const int r = q * R + k;
          const cx<double> f = (tid + TPL * q + NR * k >= N / 2) ? cmul(w[k], Z) : w[k];
          v[r] = cmul(f, v[r]);
        }
        if (q + 1 < Q) {
#pragma unroll
          for (int k = 0; k < R; ++k) w[k] = cmul(w[k], U);
        }
      }
    }
    return true;
  }
}

// Per-axis K factor table in the working precision.
template <typename T>
struct KTab {
  static __device__ __forceinline__ const double* get(const Phys& p, int a) { return p.kf[a]; }
};
template <>
struct KTab<float> {
  static __device__ __forceinline__ const float* get(const Phys& p, int a) { return p.kff[a]; }
};

// K step: GK/EK (src/operators.cu:1082,1102) with both 1/sqrt(N) of
// src/evolution.cu:761,777 folded into invN.  For the separable K the operator
// is the product of per-axis factors (no transcendental per point).
template <typename T, int E, bool X = false, bool EQ = false, typename F>
__device__ __forceinline__ void kstep_all(const Phys& p, cx<T>* v, F pos, double invN_) {
  const T invN = (T)invN_;
  if constexpr (EQ) {
    if (p.k_kind == K_EQN) {   // as a K table, with the value of the program at time t
      T arg[E];
      const T cK = (T)p.cK;
#pragma unroll
      for (int r = 0; r < E; ++r) arg[r] = (T)eq_field(p, p.eq[1], pos(r)) * cK;
      apply_phase_all<T, E, X>(v, arg, p.gstate, invN_);
      return;
    }
  }
  if constexpr (X && sizeof(T) == 4) {
    if (p.k_kind == K_SEP && p.gstate != 0) {
      // factor from the FP64 tables, product in FP64 (see apply_phase_all)
#pragma unroll
      for (int r = 0; r < E; ++r) {
        const Pt q = pos(r);
        const double* gx = p.kf[0] + 2 * q.x;
        const double* gy = p.kf[1] + 2 * q.y;
        const double* gz = p.kf[2] + 2 * q.z;
        const double ar = gx[0] * gy[0] - gx[1] * gy[1], ai = gx[0] * gy[1] + gx[1] * gy[0];
        const double br = (ar * gz[0] - ai * gz[1]) * invN_, bi = (ar * gz[1] + ai * gz[0]) * invN_;
        const double x = v[r].x, y = v[r].y;
        v[r] = mk((T)(br * x - bi * y), (T)(br * y + bi * x));
      }
      return;
    }
  }
  if (p.k_kind == K_SEP) {
    T f[E][6];
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const Pt q = pos(r);
      const auto* fx = KTab<T>::get(p, 0) + 2 * q.x;
      const auto* fy = KTab<T>::get(p, 1) + 2 * q.y;
      const auto* fz = KTab<T>::get(p, 2) + 2 * q.z;
      f[r][0] = (T)fx[0]; f[r][1] = (T)fx[1]; f[r][2] = (T)fy[0];
      f[r][3] = (T)fy[1]; f[r][4] = (T)fz[0]; f[r][5] = (T)fz[1];
    }
    if (p.gstate == 0) {
#pragma unroll
      for (int r = 0; r < E; ++r) v[r] = cscale(v[r], f[r][0] * f[r][2] * f[r][4] * invN);
    } else {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        const T ar = f[r][0] * f[r][2] - f[r][1] * f[r][3], ai = f[r][0] * f[r][3] + f[r][1] * f[r][2];
        const T br = (ar * f[r][4] - ai * f[r][5]) * invN, bi = (ar * f[r][5] + ai * f[r][4]) * invN;
        v[r] = cmul(mk(br, bi), v[r]);
      }
    }
    return;
  }
  T arg[E];
  const T cK = (T)p.cK;
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const Pt q = pos(r);
    arg[r] = (T)p.Kt[gid_of(p, q.x, q.y, q.z)] * cK;
  }
  apply_phase_all<T, E, X>(v, arg, p.gstate, invN_);
}

// K step on a strided line (axis AXIS = 1 or 2 varies along the line; x = xi
// and the other coordinate o are fixed), separable K: the fixed axes' factors
// are loaded once, only the varying axis' factor per point.  Same products as
// kstep_all.
template <typename T, int E, int AXIS>
struct KCol {
  T fx0, fx1, fo0, fo1, fe[E][2];
  template <typename F>
  __device__ __forceinline__ void load(const Phys& p, int xi, int o, F epos) {
    const auto* fxp = KTab<T>::get(p, 0) + 2 * xi;
    fx0 = (T)fxp[0];
    fx1 = (T)fxp[1];
    const auto* fop = KTab<T>::get(p, AXIS == 1 ? 2 : 1) + 2 * o;
    fo0 = (T)fop[0];
    fo1 = (T)fop[1];
    const auto* fvp = KTab<T>::get(p, AXIS);
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const int e = epos(r);
      fe[r][0] = (T)fvp[2 * e];
      fe[r][1] = (T)fvp[2 * e + 1];
    }
  }
  __device__ __forceinline__ void apply(const Phys& p, cx<T>* v, double invN_) const {
    const T invN = (T)invN_;
    if (p.gstate == 0) {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        const T fy0 = AXIS == 1 ? fe[r][0] : fo0, fz0 = AXIS == 1 ? fo0 : fe[r][0];
        v[r] = cscale(v[r], fx0 * fy0 * fz0 * invN);
      }
    } else {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        const T fy0 = AXIS == 1 ? fe[r][0] : fo0, fy1 = AXIS == 1 ? fe[r][1] : fo1;
        const T fz0 = AXIS == 1 ? fo0 : fe[r][0], fz1 = AXIS == 1 ? fo1 : fe[r][1];
        const T ar = fx0 * fy0 - fx1 * fy1, ai = fx0 * fy1 + fx1 * fy0;
        const T br = (ar * fz0 - ai * fz1) * invN, bi = (ar * fz1 + ai * fz0) * invN;
        v[r] = cmul(mk(br, bi), v[r]);
      }
    }
  }
};

template <typename T, int E, int AXIS, bool X = false, bool EQ = false, typename F>
__device__ __forceinline__ void kstep_col(const Phys& p, cx<T>* v, int xi, int o, F epos, double invN_) {
  if (p.k_kind != K_SEP || (X && sizeof(T) == 4 && p.gstate != 0)) {
    kstep_all<T, E, X, EQ>(p, v, [&](int r) { const int e = epos(r); return AXIS == 1 ? Pt{xi, e, o} : Pt{xi, o, e}; },
                           invN_);
    return;
  }
  KCol<T, E, AXIS> kc;
  kc.load(p, xi, o, epos);
  kc.apply(p, v, invN_);
}

// ---------------------------------------------------------------------------
// Deterministic two-stage reduction of NA doubles: every workgroup writes its
// partial sums (plain stores, no fences), a one-workgroup finalize kernel
// launched after it sums them in a fixed order.  The result is the same for
// every run and independent of scheduling.
// ---------------------------------------------------------------------------
template <int NA>
__device__ __forceinline__ void block_sum(double (&v)[NA], double* sh /* >= NA * 16 */) {
  // a workgroup smaller than a wave (strided sweeps of 16- and 32-point lines
  // with narrow tiles) has lanes past its end: the butterfly stays within the
  // workgroup's lanes (the offsets not taken are >= blockDim.x)
  const int wl = (int)blockDim.x;
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    double x = v[a];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      if (o < wl) x += __shfl_xor(x, o, 64);
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

// Sum of nb partials in a fixed order, as lane t of a B-lane finalize block
// adds them: elements t, t+B, t+2B, ... in four interleaved accumulators (so
// four loads are in flight per lane).  The order depends only on nb and B.
__device__ __forceinline__ double fin_sum_at(const double* __restrict__ p, int nb, int t, int B) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = t;
  for (; i + 3 * B < nb; i += 4 * B) {
    const double a = p[i], b = p[i + B], c = p[i + 2 * B], d = p[i + 3 * B];
    s0 += a;
    s1 += b;
    s2 += c;
    s3 += d;
  }
  for (; i < nb; i += B) s0 += p[i];
  return (s0 + s1) + (s2 + s3);
}

// The 1024-lane form (nb >= 4096): eight accumulators, so eight loads are in
// flight per lane (a 32768-partial sum takes four rounds instead of eight).
__device__ __forceinline__ double fin_sum8_at(const double* __restrict__ p, int nb, int t, int B) {
  double s[8] = {};
  int i = t;
  for (; i + 7 * B < nb; i += 8 * B) {
    double x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = p[i + k * B];
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] += x[k];
  }
  for (; i < nb; i += B) s[0] += p[i];
  return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// The finalize block size for nb partials (launch_finalize).
__host__ __device__ constexpr int fin_block(int nb) { return nb >= 4096 ? 1024 : 256; }

// fin_sum_at(p, nb, t, 256) for the four lanes t = l + 64j (j < 4) of one
// finalize block: in chunks of four consecutive elements per lane (k = 4c ..
// 4c+3, element t + k*256; zero past nb), all sixteen loads of a chunk first,
// then the same additions in the same order: a chunk is one iteration of the
// first loop while t + (4c+3)*256 < nb (monotone in c), else its elements go
// one by one into s0 as the second loop takes them (adding an exact zero
// changes nothing).
__device__ __forceinline__ void fin_sum4(const double* __restrict__ p, int nb, int l, double (&out)[4]) {
  constexpr int B = 256;
  double s[4][4] = {};
  const int nch = (nb + 4 * B - 1) / (4 * B);
  for (int c = 0; c < nch; ++c) {
    double v[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = l + 64 * j + (4 * c + q) * B;
        v[j][q] = i < nb ? p[i] : 0.0;
      }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (l + 64 * j + (4 * c + 3) * B < nb) {
#pragma unroll
        for (int q = 0; q < 4; ++q) s[j][q] += v[j][q];
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) s[j][0] += v[j][q];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) out[j] = (s[j][0] + s[j][1]) + (s[j][2] + s[j][3]);
}

// The parSum norm sqrt(sum * scale) of nb < 4096 partials, computed by one full
// wave (every lane gets it; the host defers a norm only to sweeps whose
// workgroups are whole waves) with the additions of k_finalize in its 256-lane
// form, bit for bit: lane l stands for finalize lanes l + 64j (j < 4); each j
// is one finalize wave, reduced by the same xor butterfly as block_sum (the
// four run interleaved), and the four wave totals are added in order.
__device__ __forceinline__ double part_norm(const double* __restrict__ p, int nb, double scale) {
  const int l = threadIdx.x & 63;
  double x[4];
  fin_sum4(p, nb, l, x);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int j = 0; j < 4; ++j) x[j] += __shfl_xor(x[j], o, 64);
  double tot = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) tot += x[j];
  return sqrt(tot * scale);
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
