// misc_kernels.hip -- pointwise kernels around the sweeps: renormalisation,
// the position-space part of GPUE's energy, the vortex plaquette detector and
// the per-point equation evaluator.
#include "misc_kernels.h"

#include <math.h>

namespace gpue {

// ---------------------------------------------------------------------------
// Equation programs: postfix form of GPUE's AST (src/dynamic.cu:527-603).
// ---------------------------------------------------------------------------
__device__ static double eval_eqn(const EqnProg* p, double x, double y, double z, double t) {
  double st[EQ_MAX];
  int sp = 0;
  const int n = p->n;
  for (int i = 0; i < n; ++i) {
    const int op = p->op[i];
    if (op == EQ_CONST) { st[sp++] = p->val[i]; continue; }
    if (op == EQ_X) { st[sp++] = x; continue; }
    if (op == EQ_Y) { st[sp++] = y; continue; }
    if (op == EQ_Z) { st[sp++] = z; continue; }
    if (op == EQ_T) { st[sp++] = t; continue; }
    const double b = st[--sp], a = st[sp - 1];
    double r = 0.0;
    switch (op - EQ_OP) {
      case 1: r = a + b; break;
      case 2: r = a - b; break;
      case 3: r = a * b; break;
      case 4: r = a / b; break;
      case 5: r = pow(a, (int)b); break;          // pow_gpu (:88-90)
      case 6: r = cos(a); break;
      case 7: r = sin(a); break;
      case 8: r = tan(a); break;
      case 9: r = sqrt(a); break;
      case 10: r = exp(a); break;
      case 11: { int k = (int)a; r = k == 0 ? j0(b) : k == 1 ? j1(b) : jn(k, b); break; }   // :138-148
      case 12: { int k = (int)a; r = k == 0 ? y0(b) : k == 1 ? y1(b) : yn(k, b); break; }   // :150-160
      default: r = 0.0; break;
    }
    st[sp - 1] = r;
  }
  return sp > 0 ? st[sp - 1] : 0.0;
}

// find_field (src/dynamic.cu:605-616): static field at dx*xid - xMax, ...
__global__ void k_eqn_field(EqnProg prog, double* out, int xDim, int yDim, int zDim, double dx, double dy,
                            double dz, double xMax, double yMax, double zMax, double t) {
  const long N = (long)xDim * yDim * zDim;
  for (long gid = blockIdx.x * (long)blockDim.x + threadIdx.x; gid < N; gid += (long)gridDim.x * blockDim.x) {
    const int xi = (int)(gid % xDim);
    const int yi = (int)((gid / xDim) % yDim);
    const int zi = (int)(gid / ((long)xDim * yDim));
    out[gid] = eval_eqn(&prog, dx * xi - xMax, dy * yi - yMax, dz * zi - zMax, t);
  }
}

// Time-dependent fields at time p.t into their tables (for the energy): the same
// per-point values the EQ sweeps evaluate.
__global__ void k_fields_at(Phys p, double* V, double* K, double* A0, double* A1, double* A2, long N) {
  for (long gid = blockIdx.x * (long)blockDim.x + threadIdx.x; gid < N; gid += (long)gridDim.x * blockDim.x) {
    const Pt q{(int)(gid % p.xDim), (int)((gid / p.xDim) % p.yDim), (int)(gid / ((long)p.xDim * p.yDim))};
    if (V) V[gid] = V_eq(p, q, gid);
    if (K) K[gid] = eq_field(p, p.eq[1], q);
    if (A0) A0[gid] = eq_field(p, p.eq[2], q);
    if (A1) A1[gid] = eq_field(p, p.eq[3], q);
    if (A2) A2[gid] = eq_field(p, p.eq[4], q);
  }
}

// ---------------------------------------------------------------------------
// psi /= norm  (pending renormalisation made explicit)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_apply_norm(cx<T>* psi, long N, const double* norm) {
  const double nrm = *norm;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    cx<T> v = psi[i];
    psi[i] = mk((T)((double)v.x / nrm), (T)((double)v.y / nrm));
  }
}

// parSum on its own (src/split_op.cu:83-129): sum |psi|^2 then divide
template <typename T>
__global__ void k_norm_sum(const cx<T>* psi, long N, double* partials) {
  __shared__ double sh[64];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    cx<T> v = psi[i];
    acc += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
  double pv[1] = {acc};
  block_partial<1>(pv, partials, sh, blockIdx.x, gridDim.x);
}

// Second stage of every reduction: one workgroup sums nb partials of each of
// na accumulators in a fixed order (fin_sum_at, then block_sum).  mode 0:
// out[a] = tot[a] * scale; mode 1: out[0] = sqrt(tot[0] * scale) (the parSum
// norm, src/kernels.cu:482).
__device__ __forceinline__ double fin_sum(const double* __restrict__ p, int nb) {
  return blockDim.x >= 1024 ? fin_sum8_at(p, nb, threadIdx.x, blockDim.x) : fin_sum_at(p, nb, threadIdx.x, blockDim.x);
}

__global__ void __launch_bounds__(1024) k_finalize(const double* partials, int nb, int na, int mode, double scale,
                                                   double* out) {
  __shared__ double sh[64];
  for (int a = 0; a < na; ++a) {
    double acc[1] = {fin_sum(partials + (size_t)a * nb, nb)};
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) out[a] = mode == 1 ? sqrt(acc[0] * scale) : acc[0] * scale;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Position-space terms of energy_calc (src/split_op.cu:275-360):
//   [0] sum V |psi|^2      [1] sum |psi|^4
//   [2] sum Re conj(psi) (i hbar) sum_a A_a D_a psi   (forward differences on the
//       flat array, derive src/kernels.cu:58-71; included iff corotating && dimnum>1)
//   [3] sum |psi|^2
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_energy_pos(const cx<T>* psi, Phys p, int gauge, double hbar, double dx, double dy,
                             double dz, double* partials) {
  __shared__ double sh[64];
  const long N = (long)p.xDim * p.yDim * p.zDim;
  const long sy = p.xDim, sz = (long)p.xDim * p.yDim;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long gid = blockIdx.x * (long)blockDim.x + threadIdx.x; gid < N; gid += (long)gridDim.x * blockDim.x) {
    const int xi = (int)(gid % p.xDim);
    const int yi = (int)((gid / p.xDim) % p.yDim);
    const int zi = (int)(gid / sz);
    const cx<T> w = psi[gid];
    const double wx = w.x, wy = w.y;
    const double dens = wx * wx + wy * wy;
    acc[0] += V_val(p, xi, yi, zi, gid) * dens;
    acc[1] += dens * dens;
    acc[3] += dens;
    if (gauge) {
      double sx = 0.0, syy = 0.0;
      const long strides[3] = {1, sy, sz};
      const double h[3] = {dx, dy, dz};
      const int naxes = p.dimnum;
      for (int ax = 0; ax < naxes; ++ax) {
        double ddx, ddy;
        if (gid + strides[ax] < N) {
          const cx<T> q = psi[gid + strides[ax]];
          ddx = ((double)q.x - wx) / h[ax];
          ddy = ((double)q.y - wy) / h[ax];
        } else {
          ddx = wx / h[ax];
          ddy = wy / h[ax];
        }
        const double A = A_val(p, ax, xi, yi, zi, gid);
        sx += ddx * A;
        syy += ddy * A;
      }
      // Re[conj(psi) * (i hbar) * S] = hbar * (psi.y S.x - psi.x S.y)
      acc[2] += wx * (-syy * hbar) - (-wy) * (sx * hbar);
    }
  }
  block_partial<4>(acc, partials, sh, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Vortex plaquettes (src/tracker.cu:102-161 with src/minions.cu:62-93).
// Pass 1: every in-mask plaquette computes its winding candidate (0 = none).
// Pass 2: a candidate is accepted iff the number of consecutive candidates
// immediately before it in the raster stream of in-mask plaquettes is even --
// the closed form of GPUE's sequential cond_x/cond_y skip rule.
// Arithmetic is kept unfused so that it rounds as GPUE's host code does.
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
__device__ __forceinline__ double2 h_scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double h_mag(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }
__device__ __forceinline__ double2 h_div(double2 num, double2 den) {
  double2 c = make_double2(den.x, -den.y);
  double2 m = make_double2(num.x * c.x - num.y * c.y, num.x * c.y + num.y * c.x);
  return h_scale(m, 1.0 / (den.x * den.x + den.y * den.y));
}
__device__ __forceinline__ double h_phase(double2 a, double2 b) {
  const double PI_ = 3.141592653589793;
  double2 g = h_scale(h_div(a, b), h_mag(b) / h_mag(a));
  double ph = atan2(g.y, g.x);
  if (ph <= -PI_) ph += 2 * PI_;
  return ph;
}

template <typename T>
__global__ void k_vortex_cand(const cx<T>* psi, const unsigned char* mask, int xDim, int* cand) {
  const int P = xDim - 1;
  const long np = (long)P * P;
  const double PI_ = 3.141592653589793;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < np; idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx / P), j = (int)(idx % P);
    int c = 0;
    if (mask[idx]) {
      auto ld = [&](int a, int b) {
        cx<T> v = psi[(long)a * xDim + b];
        return make_double2((double)v.x, (double)v.y);
      };
      double2 a = ld(i, j), b = ld(i + 1, j), cc = ld(i + 1, j + 1), d = ld(i, j + 1);
      double sum = h_phase(a, b) + h_phase(b, cc);
      sum = sum + h_phase(cc, d);
      sum = sum + h_phase(d, a);
      long rnd = lround(sum / (2 * PI_));
      // |sum| >= 1.9 pi rounds to a winding of magnitude >= 1, so hits are nonzero
      if (sum >= 1.9 * PI_) c = (int)rnd;
      else if (sum <= -1.9 * PI_) c = (int)-rnd;
    } else {
      c = INT_MAX;   // not in mask: not part of the stream
    }
    cand[idx] = c;
  }
}
#pragma clang fp contract(on)

__global__ void k_vortex_resolve(const int* cand, int xDim, int* marker, int* found) {
  const int P = xDim - 1;
  const long np = (long)P * P;
  int local = 0;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < np; idx += (long)gridDim.x * blockDim.x) {
    const int c = cand[idx];
    if (c == 0 || c == INT_MAX) continue;
    // count consecutive candidates before idx in the in-mask stream
    int run = 0;
    for (long q = idx - 1; q >= 0; --q) {
      const int cq = cand[q];
      if (cq == INT_MAX) continue;     // outside the mask
      if (cq == 0) break;              // in-mask, no candidate: run ends
      ++run;
    }
    if ((run & 1) == 0) {
      const int i = (int)(idx / P), j = (int)(idx % P);
      marker[(long)i * xDim + j] = c;
      ++local;
    }
  }
  if (local) atomicAdd(found, local);
}

// 16-byte copy (state snapshots for sweep timing); grid-stride
__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ in, uint4* __restrict__ out, long n16) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) out[i] = in[i];
}

hipError_t launch_copy16(const void* in, void* out, size_t bytes, hipStream_t s) {
  if (bytes % 16 != 0) return hipErrorInvalidValue;
  const long n16 = (long)(bytes / 16);
  const long want = (n16 + 255) / 256;
  const int blocks = (int)(want < 8192 ? (want > 0 ? want : 1) : 8192);
  hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, s, (const uint4*)in, (uint4*)out, n16);
  return hipGetLastError();
}

__global__ void k_set_double(double* p, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

hipError_t launch_set_double(double* p, double v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_double, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

// Energy terms of psi / norm from sums over the stored psi: results[0] is the
// pending norm (1.0 when none), results[2..5] the position sums (V|psi|^2,
// |psi|^4, gauge term, |psi|^2); with naxes > 0 the kinetic sum is the total of
// the per-axis sums results[9..9+naxes-1], else results[1] already holds it.
__global__ void k_energy_fold(double* r, int naxes) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double inv2 = 1.0 / (r[0] * r[0]);
  double kin = r[1];
  if (naxes > 0) {
    kin = r[9];
    for (int a = 1; a < naxes; ++a) kin += r[9 + a];
  }
  r[1] = kin * inv2;
  r[2] *= inv2;
  r[3] *= inv2 * inv2;
  r[4] *= inv2;
  r[5] *= inv2;
}

hipError_t launch_energy_fold(double* results, int naxes, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_fold, dim3(1), dim3(64), 0, s, results, naxes);
  return hipGetLastError();
}

// the FP64 transcendentals of the fused operators, pointwise (verification)
__global__ void k_math_eval(int fn, const double* __restrict__ x, double* __restrict__ y, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double s, c;
    float sf, cf;
    if (fn == 0) {
      y[i] = WMath<double>::ex(x[i]);
    } else if (fn <= 2) {
      WMath<double>::sc(x[i], &s, &c);
      y[i] = fn == 1 ? s : c;
    } else if (fn <= 4) {
      WMath<float>::sc((float)x[i], &sf, &cf);
      y[i] = fn == 3 ? sf : cf;
    } else if (fn == 5) {
      y[i] = exp_small(x[i]);   // the short forms, for |x| <= SMALL_ARG
    } else {
      sincos_small(x[i], &s, &c);
      y[i] = fn == 6 ? s : c;
    }
  }
}

hipError_t launch_math_eval(int fn, const double* x, double* y, long n, hipStream_t s) {
  const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_math_eval, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, s, fn, x, y, n);
  return hipGetLastError();
}

__global__ void k_fill_int(int* p, long n, int v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline int gs_blocks(long n, int bs, int cap) {
  long b = (n + bs - 1) / bs;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (int)b;
}

hipError_t launch_fields_at(const Phys& p, double* V, double* K, double* A0, double* A1, double* A2, long N,
                            hipStream_t s) {
  const int blocks = (int)std::min<long>((N + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fields_at, dim3(blocks), dim3(256), 0, s, p, V, K, A0, A1, A2, N);
  return hipGetLastError();
}

hipError_t launch_eqn_field(const EqnProg& prog, double* out, int xDim, int yDim, int zDim, double dx, double dy,
                            double dz, double xMax, double yMax, double zMax, double t, hipStream_t s) {
  long N = (long)xDim * yDim * zDim;
  hipLaunchKernelGGL(k_eqn_field, dim3(gs_blocks(N, 256, 8192)), dim3(256), 0, s, prog, out, xDim, yDim, zDim, dx, dy,
                     dz, xMax, yMax, zMax, t);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_apply_norm(void* psi, long N, const double* norm, hipStream_t s) {
  hipLaunchKernelGGL(k_apply_norm<T>, dim3(gs_blocks(N, 256, 16384)), dim3(256), 0, s, (cx<T>*)psi, N, norm);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_norm_sum(const void* psi, long N, double* partials, int* nb, hipStream_t s) {
  *nb = gs_blocks(N, 256, REDUCE_BLOCKS);
  hipLaunchKernelGGL(k_norm_sum<T>, dim3(*nb), dim3(256), 0, s, (const cx<T>*)psi, N, partials);
  return hipGetLastError();
}

// Batched finalize: one 1024-lane block per instance, each summing in the
// order k_finalize uses for its nb (lanes past fin_block(nb) add nothing, and
// the zero wave totals they leave are exact), so an instance's results are
// the same batched or alone.
__global__ void __launch_bounds__(1024) k_finalize_b(const FinArgs* __restrict__ args) {
  const FinArgs& f = args[blockIdx.x];
  if (!f.active) return;
  __shared__ double sh[64];
  const int B = fin_block(f.nb);
  for (int a = 0; a < f.na; ++a) {
    const double* pa = f.partials + (size_t)a * f.nb;
    double acc[1] = {(int)threadIdx.x >= B ? 0.0
                     : B >= 1024 ? fin_sum8_at(pa, f.nb, threadIdx.x, B) : fin_sum_at(pa, f.nb, threadIdx.x, B)};
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) f.out[a] = f.mode == 1 ? sqrt(acc[0] * f.scale) : acc[0] * f.scale;
    __syncthreads();
  }
}

hipError_t launch_finalize_batched(const FinArgs* d, int ninst, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_b, dim3(ninst), dim3(1024), 0, s, d);
  return hipGetLastError();
}

hipError_t launch_finalize(const double* partials, int nb, int na, int mode, double scale, double* out, hipStream_t s) {
  const int bs = fin_block(nb);   // block_sum takes up to 16 waves
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(bs), 0, s, partials, nb, na, mode, scale, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_energy_pos(const void* psi, const Phys& p, int gauge, double hbar, double dx, double dy, double dz,
                             double* partials, int* nb, hipStream_t s) {
  long N = (long)p.xDim * p.yDim * p.zDim;
  *nb = gs_blocks(N, 256, REDUCE_BLOCKS);
  hipLaunchKernelGGL(k_energy_pos<T>, dim3(*nb), dim3(256), 0, s, (const cx<T>*)psi, p, gauge, hbar, dx, dy, dz,
                     partials);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_vortex(const void* psi, const unsigned char* mask, int xDim, int* cand, int* marker, int* found,
                         hipStream_t s) {
  long np = (long)(xDim - 1) * (xDim - 1);
  long nm = (long)xDim * xDim;
  hipLaunchKernelGGL(k_fill_int, dim3(gs_blocks(nm, 256, 4096)), dim3(256), 0, s, marker, nm, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, found, 1L, 0);
  hipLaunchKernelGGL(k_vortex_cand<T>, dim3(gs_blocks(np, 256, 8192)), dim3(256), 0, s, (const cx<T>*)psi, mask, xDim,
                     cand);
  hipLaunchKernelGGL(k_vortex_resolve, dim3(gs_blocks(np, 256, 8192)), dim3(256), 0, s, cand, xDim, marker, found);
  return hipGetLastError();
}

template hipError_t launch_apply_norm<double>(void*, long, const double*, hipStream_t);
template hipError_t launch_apply_norm<float>(void*, long, const double*, hipStream_t);
template hipError_t launch_norm_sum<double>(const void*, long, double*, int*, hipStream_t);
template hipError_t launch_norm_sum<float>(const void*, long, double*, int*, hipStream_t);
template hipError_t launch_energy_pos<double>(const void*, const Phys&, int, double, double, double, double, double*,
                                              int*, hipStream_t);
template hipError_t launch_energy_pos<float>(const void*, const Phys&, int, double, double, double, double, double*,
                                             int*, hipStream_t);
template hipError_t launch_vortex<double>(const void*, const unsigned char*, int, int*, int*, int*, hipStream_t);
template hipError_t launch_vortex<float>(const void*, const unsigned char*, int, int*, int*, int*, hipStream_t);

}  // namespace gpue
