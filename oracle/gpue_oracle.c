/*
 * gpue_oracle.c -- CPU FP64 restatement of GPUE's split-operator step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Nothing in the product
 * (gpue_amd/, libgpue_amd.so) links or calls this file.  Built by
 * oracle/Makefile into oracle/libgpue_oracle.so and loaded from tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Citations are GPUE source paths (relative to the GPUE repository root).
 * The arithmetic order of every pointwise operation follows the cited kernel
 * so that FP64 results differ from GPUE only through the FFT rounding and the
 * summation order of reductions (GPUE's own tree order is unspecified:
 * src/split_op.cu:64-78 runs multipass in place).
 */
#include "gpue_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.141592653589793   /* include/constants.h:19 */

/* ------------------------------------------------------------------------ */
/* FFT: own mixed-radix Stockham autosort, FP64, unnormalised.              */
/* Pass with radix R after L points already combined, m = n/R:              */
/*   a_t = x[j + t m] * w^(t k),  k = j mod L,  w = exp(sign 2 pi i/(L R))  */
/*   y[(j/L) L R + k + s L] = sum_t a_t exp(sign 2 pi i s t / R)            */
/* Replaces cuFFT Z2Z (src/init.cu:309-319, src/ds.cu:11-79).               */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n, nf, fac[64];
    or_c *tw;          /* tw[m] = exp(-2 pi i m / n) */
} or_plan;

static void plan_init(or_plan *p, int n) {
    p->n = n; p->nf = 0;
    int r = n;
    while (r % 4 == 0) { p->fac[p->nf++] = 4; r /= 4; }
    while (r % 2 == 0) { p->fac[p->nf++] = 2; r /= 2; }
    for (int q = 3; r > 1; q += 2) {
        while (r % q == 0) { p->fac[p->nf++] = q; r /= q; }
        if ((long)q * q > r && r > 1) { p->fac[p->nf++] = r; r = 1; }
    }
    p->tw = (or_c *)malloc(sizeof(or_c) * (size_t)n);
    for (int m = 0; m < n; ++m) {
        double a = 2.0 * OR_PI * (double)m / (double)n;
        p->tw[m].x = cos(a);
        p->tw[m].y = -sin(a);
    }
}

static void plan_free(or_plan *p) { free(p->tw); }

static inline or_c cmul(or_c a, or_c b) {
    or_c r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
static inline or_c twid(const or_plan *p, long idx, int sign) {
    or_c w = p->tw[idx % p->n];
    if (sign > 0) w.y = -w.y;
    return w;
}

/* transform one contiguous line x (length n) using scratch y; result in x */
static void fft_line(const or_plan *p, or_c *x, or_c *y, int sign) {
    const int n = p->n;
    or_c *src = x, *dst = y;
    long L = 1;
    or_c a[64], b[64];
    for (int f = 0; f < p->nf; ++f) {
        const int R = p->fac[f];
        const long m = n / R;
        const long step = n / (L * R);
        for (long j = 0; j < m; ++j) {
            const long k = j % L;
            if (R <= 64) {
                for (int t = 0; t < R; ++t) {
                    or_c v = src[j + t * m];
                    a[t] = (k == 0 || t == 0) ? v : cmul(v, twid(p, (long)t * k * step, sign));
                }
                if (R == 2) {
                    b[0].x = a[0].x + a[1].x; b[0].y = a[0].y + a[1].y;
                    b[1].x = a[0].x - a[1].x; b[1].y = a[0].y - a[1].y;
                } else if (R == 4) {
                    or_c s0 = {a[0].x + a[2].x, a[0].y + a[2].y};
                    or_c s1 = {a[0].x - a[2].x, a[0].y - a[2].y};
                    or_c s2 = {a[1].x + a[3].x, a[1].y + a[3].y};
                    or_c s3 = {a[1].x - a[3].x, a[1].y - a[3].y};
                    /* multiply s3 by sign*i */
                    or_c js3 = {-sign * s3.y, sign * s3.x};
                    b[0].x = s0.x + s2.x; b[0].y = s0.y + s2.y;
                    b[2].x = s0.x - s2.x; b[2].y = s0.y - s2.y;
                    b[1].x = s1.x + js3.x; b[1].y = s1.y + js3.y;
                    b[3].x = s1.x - js3.x; b[3].y = s1.y - js3.y;
                } else {
                    const long rs = n / R;
                    for (int s = 0; s < R; ++s) {
                        or_c acc = a[0];
                        for (int t = 1; t < R; ++t) {
                            or_c w = twid(p, (long)((s * t) % R) * rs, sign);
                            or_c v = cmul(a[t], w);
                            acc.x += v.x; acc.y += v.y;
                        }
                        b[s] = acc;
                    }
                }
                for (int s = 0; s < R; ++s) dst[(j / L) * L * R + k + s * L] = b[s];
            } else {
                /* large prime factor: direct sum */
                const long rs = n / R;
                for (int s = 0; s < R; ++s) {
                    or_c acc = {0.0, 0.0};
                    for (int t = 0; t < R; ++t) {
                        or_c v = src[j + t * m];
                        if (k != 0 && t != 0) v = cmul(v, twid(p, (long)t * k * step, sign));
                        if (t != 0) v = cmul(v, twid(p, (long)((s * (long)t) % R) * rs, sign));
                        acc.x += v.x; acc.y += v.y;
                    }
                    dst[(j / L) * L * R + k + s * L] = acc;
                }
            }
        }
        or_c *tmp = src; src = dst; dst = tmp;
        L *= R;
    }
    if (src != x) memcpy(x, src, sizeof(or_c) * (size_t)n);
}

void or_fft_lines(or_c *data, int n, int howmany, long stride, long dist, int sign) {
    if (n <= 1) return;
    or_plan p; plan_init(&p, n);
    #pragma omp parallel
    {
        or_c *buf = (or_c *)malloc(sizeof(or_c) * (size_t)n * 2);
        #pragma omp for schedule(static)
        for (int b = 0; b < howmany; ++b) {
            or_c *base = data + (long)b * dist;
            for (int i = 0; i < n; ++i) buf[i] = base[(long)i * stride];
            fft_line(&p, buf, buf + n, sign);
            for (int i = 0; i < n; ++i) base[(long)i * stride] = buf[i];
        }
        free(buf);
    }
    plan_free(&p);
}

/* transform along one axis of an x-fastest (xDim, yDim, zDim) array */
static void fft_axis(or_c *data, int xDim, int yDim, int zDim, int axis, int sign) {
    if (axis == 0) {
        or_fft_lines(data, xDim, yDim * zDim, 1, xDim, sign);
    } else if (axis == 1) {
        for (int z = 0; z < zDim; ++z)
            or_fft_lines(data + (long)z * xDim * yDim, yDim, xDim, xDim, 1, sign);
    } else {
        or_fft_lines(data, zDim, xDim * yDim, (long)xDim * yDim, 1, sign);
    }
}

void or_fft_nd(or_c *data, int xDim, int yDim, int zDim, int sign) {
    if (xDim > 1) fft_axis(data, xDim, yDim, zDim, 0, sign);
    if (yDim > 1) fft_axis(data, xDim, yDim, zDim, 1, sign);
    if (zDim > 1) fft_axis(data, xDim, yDim, zDim, 2, sign);
}

void or_dft_bruteforce(const or_c *in, or_c *out, int n, int sign) {
    for (int k = 0; k < n; ++k) {
        long double ax = 0.0L, ay = 0.0L;
        for (int j = 0; j < n; ++j) {
            long double a = 2.0L * 3.14159265358979323846264338327950288L *
                            (long double)(((long)j * k) % n) / (long double)n;
            long double c = cosl(a), s = sinl(a) * (long double)sign;
            ax += (long double)in[j].x * c - (long double)in[j].y * s;
            ay += (long double)in[j].x * s + (long double)in[j].y * c;
        }
        out[k].x = (double)ax; out[k].y = (double)ay;
    }
}

/* ------------------------------------------------------------------------ */
/* Context: grids, fields and GPUE's operator arrays.                        */
/* ------------------------------------------------------------------------ */

struct or_ctx {
    or_params p;
    long N;
    double dx, dy, dz, dpx, dpy, dpz, pxMax, pyMax, pzMax;
    double *x, *y, *z, *px, *py, *pz;
    double *V, *K, *Ax, *Ay, *Az;
    int v_user;                    /* V table supplied by the caller */
    or_c *GV, *EV, *GK, *EK, *GpAx, *GpAy, *GpAz, *EpAx, *EpAy, *EpAz;
    or_c *wfc;
    int ramp_type, ramp_steps;     /* -R: 0 off, 1 or 2 (src/evolution.cu:409-428) */
    or_c *RpA[3];                  /* ramped gauge operators of the current step */
};

double or_dV(const or_ctx *c) { return c->dx * c->dy * c->dz; }

/* src/operators.cu:277-432 (generate_p_space) with the box of src/init.cu:124-169 */
static void build_grids(or_ctx *c) {
    const or_params *p = &c->p;
    int xD = p->xDim, yD = p->yDim, zD = p->zDim;
    c->pxMax = (OR_PI / p->xMax) * (xD >> 1);
    c->pyMax = (OR_PI / p->yMax) * (yD >> 1);
    c->pzMax = (OR_PI / p->zMax) * (zD >> 1);
    c->dx = p->xMax / (xD >> 1);
    c->dy = p->yMax / (yD >> 1);
    c->dz = p->zMax / (zD >> 1);
    if (p->dimnum < 3) c->dz = 1;
    if (p->dimnum < 2) c->dy = 1;
    c->dpx = OR_PI / p->xMax;
    c->dpy = OR_PI / p->yMax;
    c->dpz = OR_PI / p->zMax;
    c->x = calloc((size_t)xD, sizeof(double)); c->px = calloc((size_t)xD, sizeof(double));
    c->y = calloc((size_t)yD, sizeof(double)); c->py = calloc((size_t)yD, sizeof(double));
    c->z = calloc((size_t)zD, sizeof(double)); c->pz = calloc((size_t)zD, sizeof(double));
    for (int i = 0; i < xD / 2; ++i) {
        c->x[i] = -p->xMax + i * c->dx;
        c->x[i + xD / 2] = i * c->dx;
        c->px[i] = i * c->dpx;
        c->px[i + xD / 2] = -c->pxMax + i * c->dpx;
    }
    if (p->dimnum >= 2) {
        for (int i = 0; i < yD / 2; ++i) {
            c->y[i] = -p->yMax + i * c->dy;
            c->y[i + yD / 2] = i * c->dy;
            c->py[i] = i * c->dpy;
            c->py[i + yD / 2] = -c->pyMax + i * c->dpy;
        }
    }
    if (p->dimnum == 3) {
        for (int i = 0; i < zD / 2; ++i) {
            c->z[i] = -p->zMax + i * c->dz;
            c->z[i + zD / 2] = i * c->dz;
            c->pz[i] = i * c->dpz;
            c->pz[i + zD / 2] = -c->pzMax + i * c->dpz;
        }
    }
}

#define GID(xi, yi, zi) ((long)(xi) + (long)c->p.xDim * ((long)(yi) + (long)c->p.yDim * (long)(zi)))

/* K: src/operators.cu:460-469 (simple_K); A: :621-647 (krotation_Ax/Ay, kconstant_A);
 * V: :994-1008 (kharmonic_V, items[] of :757-793) */
static void build_fields(or_ctx *c) {
    const or_params *p = &c->p;
    const double hbar = p->hbar, mass = p->mass;
    const double oZ = (p->dimnum == 3) ? p->omegaZ : 0.0;      /* items[5] */
    const double z0 = (p->dimnum == 3) ? p->z0_shift : 0.0;    /* items[8] */
    #pragma omp parallel for collapse(2) schedule(static)
    for (int zi = 0; zi < p->zDim; ++zi)
        for (int yi = 0; yi < p->yDim; ++yi)
            for (int xi = 0; xi < p->xDim; ++xi) {
                long gid = GID(xi, yi, zi);
                if (p->k_kind == OR_K_STANDARD)
                    c->K[gid] = (hbar * hbar / (2 * mass)) *
                                (c->px[xi] * c->px[xi] + c->py[yi] * c->py[yi] + c->pz[zi] * c->pz[zi]);
                if (p->a_kind == OR_A_ROTATION) {
                    c->Ax[gid] = -c->y[yi] * p->omega * p->omegaX;
                    c->Ay[gid] = c->x[xi] * p->omega * p->omegaY;
                    c->Az[gid] = 0;
                } else if (p->a_kind == OR_A_ZERO) {
                    c->Ax[gid] = 0; c->Ay[gid] = 0; c->Az[gid] = 0;
                }
                if (p->v_kind == OR_V_HARMONIC) {
                    double V_x = p->omegaX * (c->x[xi] + p->x0_shift);
                    double V_y = p->gammaY * p->omegaY * (c->y[yi] + p->y0_shift);
                    double V_z = 1.0 * oZ * (c->z[zi] + z0);
                    c->V[gid] = 0.5 * mass * ((V_x * V_x + V_y * V_y + V_z * V_z) +
                                              (c->Ax[gid] * c->Ax[gid] + c->Ay[gid] * c->Ay[gid] +
                                               c->Az[gid] * c->Az[gid]));
                }
            }
}

/* src/operators.cu:1069-1111 (aux_fields) */
static void build_ops(or_ctx *c) {
    const or_params *p = &c->p;
    const double hbar = p->hbar, gdt = p->gdt, dt = p->dt;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int zi = 0; zi < p->zDim; ++zi)
        for (int yi = 0; yi < p->yDim; ++yi)
            for (int xi = 0; xi < p->xDim; ++xi) {
                long gid = GID(xi, yi, zi);
                double V = c->V[gid], K = c->K[gid];
                c->GV[gid].x = exp(-V * (gdt / (2 * hbar))); c->GV[gid].y = 0.0;
                c->GK[gid].x = exp(-K * (gdt / hbar));       c->GK[gid].y = 0.0;
                double pAx = c->Ax[gid] * c->px[xi];
                double pAy = c->Ay[gid] * c->py[yi];
                double pAz = c->Az[gid] * c->pz[zi];
                c->GpAx[gid].x = exp(-pAx * gdt); c->GpAx[gid].y = 0;
                c->GpAy[gid].x = exp(-pAy * gdt); c->GpAy[gid].y = 0;
                c->GpAz[gid].x = exp(-pAz * gdt); c->GpAz[gid].y = 0;
                c->EV[gid].x = cos(-V * (dt / (2 * hbar))); c->EV[gid].y = sin(-V * (dt / (2 * hbar)));
                c->EK[gid].x = cos(-K * (dt / hbar));       c->EK[gid].y = sin(-K * (dt / hbar));
                c->EpAz[gid].x = cos(-pAz * dt); c->EpAz[gid].y = sin(-pAz * dt);
                c->EpAy[gid].x = cos(-pAy * dt); c->EpAy[gid].y = sin(-pAy * dt);
                c->EpAx[gid].x = cos(-pAx * dt); c->EpAx[gid].y = sin(-pAx * dt);
            }
}

or_ctx *or_create(const or_params *pin) {
    or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
    c->p = *pin;
    if (c->p.dimnum < 3) c->p.zDim = 1;
    if (c->p.dimnum < 2) c->p.yDim = 1;
    c->N = (long)c->p.xDim * c->p.yDim * c->p.zDim;
    size_t N = (size_t)c->N;
    build_grids(c);
    c->V = calloc(N, sizeof(double)); c->K = calloc(N, sizeof(double));
    c->Ax = calloc(N, sizeof(double)); c->Ay = calloc(N, sizeof(double)); c->Az = calloc(N, sizeof(double));
    or_c **ops[] = {&c->GV, &c->EV, &c->GK, &c->EK, &c->GpAx, &c->GpAy, &c->GpAz,
                    &c->EpAx, &c->EpAy, &c->EpAz, &c->wfc};
    for (unsigned i = 0; i < sizeof(ops) / sizeof(ops[0]); ++i) *ops[i] = calloc(N, sizeof(or_c));
    build_fields(c);
    build_ops(c);
    return c;
}

void or_destroy(or_ctx *c) {
    if (!c) return;
    for (int a = 0; a < 3; ++a) free(c->RpA[a]);
    free(c->x); free(c->y); free(c->z); free(c->px); free(c->py); free(c->pz);
    free(c->V); free(c->K); free(c->Ax); free(c->Ay); free(c->Az);
    free(c->GV); free(c->EV); free(c->GK); free(c->EK); free(c->GpAx); free(c->GpAy);
    free(c->GpAz); free(c->EpAx); free(c->EpAy); free(c->EpAz); free(c->wfc);
    free(c);
}

void or_set_field(or_ctx *c, int which, const double *vals) {
    size_t bytes = sizeof(double) * (size_t)c->N;
    switch (which) {
        case OR_FIELD_V:  memcpy(c->V, vals, bytes); c->p.v_kind = OR_V_TABLE; break;
        case OR_FIELD_K:  memcpy(c->K, vals, bytes); c->p.k_kind = OR_K_TABLE; break;
        case OR_FIELD_AX: memcpy(c->Ax, vals, bytes); c->p.a_kind = OR_A_TABLE; break;
        case OR_FIELD_AY: memcpy(c->Ay, vals, bytes); c->p.a_kind = OR_A_TABLE; break;
        case OR_FIELD_AZ: memcpy(c->Az, vals, bytes); c->p.a_kind = OR_A_TABLE; break;
    }
}

void or_rebuild(or_ctx *c) { build_fields(c); build_ops(c); }

void or_get_field(const or_ctx *c, int which, double *out) {
    const double *src = which == OR_FIELD_V ? c->V : which == OR_FIELD_K ? c->K :
                        which == OR_FIELD_AX ? c->Ax : which == OR_FIELD_AY ? c->Ay : c->Az;
    memcpy(out, src, sizeof(double) * (size_t)c->N);
}

void or_get_grid(const or_ctx *c, int axis, double *out) {
    const double *g[6] = {c->x, c->y, c->z, c->px, c->py, c->pz};
    int n[6] = {c->p.xDim, c->p.yDim, c->p.zDim, c->p.xDim, c->p.yDim, c->p.zDim};
    memcpy(out, g[axis], sizeof(double) * (size_t)n[axis]);
}

void or_set_wfc(or_ctx *c, const double *v) { memcpy(c->wfc, v, sizeof(or_c) * (size_t)c->N); }
void or_get_wfc(const or_ctx *c, double *v) { memcpy(v, c->wfc, sizeof(or_c) * (size_t)c->N); }
void or_set_g(or_ctx *c, double g) { c->p.g = g; }
/* OpenMP threads of the following calls (the CPU baseline runs at all cores and at 1) */
void or_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int or_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------ */
/* Pointwise kernels restated                                                */
/* ------------------------------------------------------------------------ */

/* src/kernels.cu:365-388 (cMultDensity); dt argument is 0.5*Dt (src/evolution.cu:738) */
static void cmult_density(or_ctx *c, const or_c *op, double dt, int gstate, double g) {
    const double hbar = c->p.hbar;
    #pragma omp parallel for schedule(static)
    for (long gid = 0; gid < c->N; ++gid) {
        or_c t1 = op[gid], t2 = c->wfc[gid], r;
        double gDensity = g * (t2.x * t2.x + t2.y * t2.y) * (dt / hbar);
        if (gstate == 0) {
            double tmp = t1.x * exp(-gDensity);
            r.x = tmp * t2.x - t1.y * t2.y;
            r.y = tmp * t2.y + t1.y * t2.x;
        } else {
            or_c tmp;
            tmp.x = t1.x * cos(-gDensity) - t1.y * sin(-gDensity);
            tmp.y = t1.y * cos(-gDensity) + t1.x * sin(-gDensity);
            r.x = tmp.x * t2.x - tmp.y * t2.y;
            r.y = tmp.x * t2.y + tmp.y * t2.x;
        }
        c->wfc[gid] = r;
    }
}

/* src/kernels.cu:282-290 (cMult(in1, in2)).  Both argument orders appear in GPUE:
 * cMult(op, wfc) (src/evolution.cu:771) and cMult(wfc, op) (src/evolution.cu:29). */
static void cmult_op_wfc(or_ctx *c, const or_c *op) {
    #pragma omp parallel for schedule(static)
    for (long gid = 0; gid < c->N; ++gid) {
        or_c a = op[gid], b = c->wfc[gid], r;
        r.x = a.x * b.x - a.y * b.y;
        r.y = a.x * b.y + a.y * b.x;
        c->wfc[gid] = r;
    }
}

static void cmult_wfc_op(or_ctx *c, const or_c *op) {
    #pragma omp parallel for schedule(static)
    for (long gid = 0; gid < c->N; ++gid) {
        or_c a = c->wfc[gid], b = op[gid], r;
        r.x = a.x * b.x - a.y * b.y;
        r.y = a.x * b.y + a.y * b.x;
        c->wfc[gid] = r;
    }
}

/* src/kernels.cu:452-459 (scalarMult double2 x double) */
static void scalar_mult(or_c *v, long N, double f) {
    #pragma omp parallel for schedule(static)
    for (long gid = 0; gid < N; ++gid) { v[gid].x = v[gid].x * f; v[gid].y = v[gid].y * f; }
}

/* One gauge axis: src/evolution.cu:174-190 etc.  FFT along `axis`, x 1/sqrt(n_a), x op,
 * inverse FFT, x 1/sqrt(n_a).  GPUE's 2D code uses swapped factor names
 * (src/evolution.cu:175,195,237); for the square grids GPUE supports they coincide. */
static void gauge_axis(or_ctx *c, int axis, const or_c *op) {
    const or_params *p = &c->p;
    int n = axis == 0 ? p->xDim : axis == 1 ? p->yDim : p->zDim;
    double rf = 1.0 / pow(n, 0.5);
    fft_axis(c->wfc, p->xDim, p->yDim, p->zDim, axis, -1);
    scalar_mult(c->wfc, c->N, rf);
    cmult_wfc_op(c, op);
    fft_axis(c->wfc, p->xDim, p->yDim, p->zDim, axis, +1);
    scalar_mult(c->wfc, c->N, rf);
}

/* src/split_op.cu:83-129 (parSum) with deterministic row-partial summation */
static double sum_abs2(const or_c *w, long N, long rowlen) {
    long rows = N / rowlen;
    double *part = (double *)malloc(sizeof(double) * (size_t)rows);
    #pragma omp parallel for schedule(static)
    for (long r = 0; r < rows; ++r) {
        double s = 0.0;
        const or_c *q = w + r * rowlen;
        for (long i = 0; i < rowlen; ++i) s += q[i].x * q[i].x + q[i].y * q[i].y;
        part[r] = s;
    }
    double s = 0.0;
    for (long r = 0; r < rows; ++r) s += part[r];
    free(part);
    return s;
}

void or_parsum_raw(or_c *wfc, int dimnum, int xDim, int yDim, int zDim,
                   double dx, double dy, double dz) {
    long gsize = xDim;
    double dg = dx;
    if (dimnum > 1) { gsize *= yDim; dg *= dy; }
    if (dimnum > 2) { gsize *= zDim; dg *= dz; }
    double sum = sum_abs2(wfc, gsize, xDim);
    double norm = sqrt(sum * dg);           /* src/kernels.cu:479-486 (scalarDiv_wfcNorm) */
    #pragma omp parallel for schedule(static)
    for (long gid = 0; gid < gsize; ++gid) { wfc[gid].x = wfc[gid].x / norm; wfc[gid].y = wfc[gid].y / norm; }
}

void or_parsum(or_ctx *c) {
    or_parsum_raw(c->wfc, c->p.dimnum, c->p.xDim, c->p.yDim, c->p.zDim, c->dx, c->dy, c->dz);
}

/* ------------------------------------------------------------------------ */
/* -R ramp: src/evolution.cu:409-428 sets omega_0 per step, :811-832 raises  */
/* the gauge operators to it in place, cumulatively (src/kernels.cu:491-497  */
/* scalarPow acts on re and im separately).  Imaginary time: GpA is real and */
/* positive, so after step i the operator is GpA^(f_0 f_1 ... f_i), kept    */
/* here literally by repeated pow.  Real time: the element-wise pow of a    */
/* complex phase is not a phase; the defined semantics is the same exponent */
/* scaling, exp(-i pA dt F_i) with F_i = f_0 ... f_i.                       */
/* ------------------------------------------------------------------------ */

void or_set_ramp(or_ctx *c, int type, int num_steps) {
    c->ramp_type = type;
    c->ramp_steps = num_steps;
    for (int a = 0; a < 3; ++a) {
        if (!c->RpA[a]) c->RpA[a] = (or_c *)calloc((size_t)c->N, sizeof(or_c));
    }
}

/* omega_0 of step i (src/evolution.cu:412-427) */
double or_ramp_factor(int type, double omega, int i, int num_steps) {
    if (type == 1) return i == 0 ? omega : (double)i / (double)(i + 1);
    return i == 0 ? omega / (double)num_steps : (double)(i + 1) / (double)i;
}

static void ramp_ops(or_ctx *c, int i, int gstate) {
    const or_params *p = &c->p;
    const double w0 = or_ramp_factor(c->ramp_type, p->omega, i, c->ramp_steps);
    const or_c *G[3] = {c->GpAx, c->GpAy, c->GpAz};
    if (gstate == 0) {
        for (int a = 0; a < 3; ++a) {
            or_c *R = c->RpA[a];
            if (i == 0) memcpy(R, G[a], sizeof(or_c) * (size_t)c->N);
            #pragma omp parallel for schedule(static)
            for (long gid = 0; gid < c->N; ++gid) { R[gid].x = pow(R[gid].x, w0); R[gid].y = pow(R[gid].y, w0); }
        }
        return;
    }
    double F = 1.0;
    for (int j = 0; j <= i; ++j) F *= or_ramp_factor(c->ramp_type, p->omega, j, c->ramp_steps);
    #pragma omp parallel for collapse(2) schedule(static)
    for (int zi = 0; zi < p->zDim; ++zi)
        for (int yi = 0; yi < p->yDim; ++yi)
            for (int xi = 0; xi < p->xDim; ++xi) {
                long gid = GID(xi, yi, zi);
                double pA[3] = {c->Ax[gid] * c->px[xi], c->Ay[gid] * c->py[yi], c->Az[gid] * c->pz[zi]};
                for (int a = 0; a < 3; ++a) {
                    c->RpA[a][gid].x = cos(-pA[a] * (p->dt * F));
                    c->RpA[a][gid].y = sin(-pA[a] * (p->dt * F));
                }
            }
}

/* ------------------------------------------------------------------------ */
/* One step: src/evolution.cu:727-870                                        */
/* ------------------------------------------------------------------------ */

void or_step(or_ctx *c, int i, int gstate) {
    const or_params *p = &c->p;
    const double Dt = gstate == 0 ? p->gdt : p->dt;                 /* :344-351 */
    const or_c *Vop = gstate == 0 ? c->GV : c->EV;                   /* src/init.cu:403-495 */
    const or_c *Kop = gstate == 0 ? c->GK : c->EK;
    const or_c *Aop[3] = {gstate == 0 ? c->GpAx : c->EpAx, gstate == 0 ? c->GpAy : c->EpAy,
                          gstate == 0 ? c->GpAz : c->EpAz};
    const double rnd = 1.0 / pow((double)c->N, 0.5);                 /* :336 */

    /* U_r(dt/2) (:728-755) */
    if (p->gpe) cmult_density(c, Vop, 0.5 * Dt, gstate, p->g);
    else        cmult_op_wfc(c, Vop);
    /* U_p(dt) (:758-777) */
    or_fft_nd(c->wfc, p->xDim, p->yDim, p->zDim, -1);
    scalar_mult(c->wfc, c->N, rnd);
    cmult_op_wfc(c, Kop);
    or_fft_nd(c->wfc, p->xDim, p->yDim, p->zDim, +1);
    scalar_mult(c->wfc, c->N, rnd);
    /* U_r(dt/2) (:780-809) */
    if (p->gpe) cmult_density(c, Vop, 0.5 * Dt, gstate, p->g);
    else        cmult_op_wfc(c, Vop);
    /* gauge (:811-866); flip = i%2 (:837,842) */
    if (p->corotating) {
        int flip = i % 2;
        if (c->ramp_type) {
            ramp_ops(c, i, gstate);
            for (int a = 0; a < 3; ++a) Aop[a] = c->RpA[a];
        }
        if (p->dimnum == 1) {
            gauge_axis(c, 0, Aop[0]);
        } else if (p->dimnum == 2) {
            if (flip) { gauge_axis(c, 0, Aop[0]); gauge_axis(c, 1, Aop[1]); }
            else      { gauge_axis(c, 1, Aop[1]); gauge_axis(c, 0, Aop[0]); }
        } else {
            if (flip) { gauge_axis(c, 0, Aop[0]); gauge_axis(c, 1, Aop[1]); gauge_axis(c, 2, Aop[2]); }
            else      { gauge_axis(c, 2, Aop[2]); gauge_axis(c, 1, Aop[1]); gauge_axis(c, 0, Aop[0]); }
        }
    }
    if (gstate == 0) or_parsum(c);                                    /* :868-870 */
}

/* ------------------------------------------------------------------------ */
/* Energy: src/split_op.cu:208-375 (energy_calc) with derive src/kernels.cu:58-71.
 * mu (not in GPUE) is the same sum with the interaction weight g instead of g/2. */
/* ------------------------------------------------------------------------ */

static or_c derive_at(const or_c *d, long gid, long stride, long gsize, double h) {
    or_c o;
    if (gid + stride < gsize) { o.x = (d[gid + stride].x - d[gid].x) / h; o.y = (d[gid + stride].y - d[gid].y) / h; }
    else                      { o.x = d[gid].x / h; o.y = d[gid].y / h; }
    return o;
}

double or_energy(or_ctx *c, double *mu_out) {
    const or_params *p = &c->p;
    const long N = c->N;
    const double rf = 1.0 / pow((double)N, 0.5);
    const double dg = c->dx * c->dy * c->dz;
    or_c *ek = (or_c *)malloc(sizeof(or_c) * (size_t)N);
    memcpy(ek, c->wfc, sizeof(or_c) * (size_t)N);
    or_fft_nd(ek, p->xDim, p->yDim, p->zDim, -1);
    #pragma omp parallel for schedule(static)
    for (long gid = 0; gid < N; ++gid) {
        or_c v = ek[gid];
        v.x = v.x * rf; v.y = v.y * rf;                  /* scalarMult */
        v.x = v.x * c->K[gid]; v.y = v.y * c->K[gid];    /* vecMult(wfc_k, K) */
        ek[gid] = v;
    }
    or_fft_nd(ek, p->xDim, p->yDim, p->zDim, +1);
    const int gauge = p->corotating && p->dimnum > 1;
    const double half_g = 0.5 * p->g;
    long rows = N / p->xDim;
    double *pe = (double *)malloc(sizeof(double) * (size_t)rows);
    double *pm = (double *)malloc(sizeof(double) * (size_t)rows);
    #pragma omp parallel for schedule(static)
    for (long r = 0; r < rows; ++r) {
        double se = 0.0, sm = 0.0;
        for (long gid = r * p->xDim; gid < (r + 1) * p->xDim; ++gid) {
            or_c w = c->wfc[gid];
            or_c wc = {w.x, -w.y};                                   /* vecConjugate */
            or_c k = ek[gid]; k.x = k.x * rf; k.y = k.y * rf;
            double kin = wc.x * k.x - wc.y * k.y;                    /* cMult(wfc_c, energy_k).x */
            double dens = w.x * w.x + w.y * w.y;
            double rc_e = p->gpe ? dens * half_g + c->V[gid] : c->V[gid];
            double rc_m = p->gpe ? dens * p->g + c->V[gid] : c->V[gid];
            or_c er = {w.x * rc_e, w.y * rc_e};                      /* vecMult(wfc, real_comp) */
            or_c em = {w.x * rc_m, w.y * rc_m};
            double pot_e = wc.x * er.x - wc.y * er.y;
            double pot_m = wc.x * em.x - wc.y * em.y;
            double e = pot_e + kin, m = pot_m + kin;                 /* energy_sum */
            if (gauge) {
                or_c dx = derive_at(c->wfc, gid, 1, N, c->dx);
                or_c el = {dx.x * c->Ax[gid], dx.y * c->Ax[gid]};
                or_c dy = derive_at(c->wfc, gid, p->xDim, N, c->dy);
                or_c dw = {dy.x * c->Ay[gid], dy.y * c->Ay[gid]};
                el.x = dw.x + el.x; el.y = dw.y + el.y;              /* sum(dwfc, energy_l) */
                if (p->dimnum == 3) {
                    or_c dz = derive_at(c->wfc, gid, (long)p->xDim * p->yDim, N, c->dz);
                    or_c dw3 = {dz.x * c->Az[gid], dz.y * c->Az[gid]};
                    el.x = dw3.x + el.x; el.y = dw3.y + el.y;
                }
                or_c s = {el.x * 0 - el.y * p->hbar, el.x * p->hbar + el.y * 0};  /* x (0, hbar) */
                double lz = wc.x * s.x - wc.y * s.y;                 /* cMult(wfc_c, energy_l).x */
                e = e + lz; m = m + lz;                              /* energy_lsum */
            }
            se += e; sm += m;
        }
        pe[r] = se; pm[r] = sm;
    }
    double E = 0.0, M = 0.0;
    for (long r = 0; r < rows; ++r) { E += pe[r]; M += pm[r]; }
    free(pe); free(pm); free(ek);
    if (mu_out) *mu_out = M * dg;
    return E * dg;
}

/* The gauge term of the Hamiltonian GPUE evolves.  apply_gauge multiplies by
 * exp(-pA_a gdt) (imaginary) or exp(-i pA_a dt) (real) between 1-D transforms
 * along axis a, pA_a = A_a k_a in the mixed representation (k_a along a,
 * positions across; src/operators.cu:1089-1110, src/evolution.cu:163-257), so
 * the evolved term is hbar A_a k_a.  energy_calc adds +i hbar A.grad instead
 * (src/split_op.cu:312-360), i.e. minus that.  Here, by Parseval along each axis:
 *   hbar * sum_a sum_mixed A_a[gid] k_a |FFT_a psi|^2 / n_a * dV. */
double or_gauge_spectral(or_ctx *c) {
    const or_params *p = &c->p;
    const long N = c->N;
    const int dims[3] = {p->xDim, p->yDim, p->zDim};
    const double *A[3] = {c->Ax, c->Ay, c->Az};
    const double *k[3] = {c->px, c->py, c->pz};
    or_c *w = (or_c *)malloc(sizeof(or_c) * (size_t)N);
    double total = 0.0;
    for (int a = 0; a < p->dimnum; ++a) {
        memcpy(w, c->wfc, sizeof(or_c) * (size_t)N);
        fft_axis(w, p->xDim, p->yDim, p->zDim, a, -1);
        double s = 0.0;
        #pragma omp parallel for reduction(+ : s) schedule(static)
        for (long gid = 0; gid < N; ++gid) {
            const long idx = a == 0 ? gid % p->xDim : a == 1 ? (gid / p->xDim) % p->yDim
                                                            : gid / ((long)p->xDim * p->yDim);
            s += A[a][gid] * k[a][idx] * (w[gid].x * w[gid].x + w[gid].y * w[gid].y);
        }
        total += s / dims[a];
    }
    free(w);
    return p->hbar * total * c->dx * c->dy * c->dz;
}

/* energy_calc with the gauge term replaced by or_gauge_spectral (when GPUE
 * includes one: corotating and dimnum > 1) */
double or_energy_consistent(or_ctx *c, double *mu_out) {
    const int gauge = c->p.corotating && c->p.dimnum > 1;
    const int corot = c->p.corotating;
    c->p.corotating = 0;
    double mu = 0.0;
    double e = or_energy(c, &mu);
    c->p.corotating = corot;
    if (gauge) {
        const double t = or_gauge_spectral(c);
        e += t;
        mu += t;
    }
    if (mu_out) *mu_out = mu;
    return e;
}

/* src/evolution.cu:407-904 without host I/O: the step loop with energy cadence and
 * the imaginary-time early stop (:872-903).  Returns the number of steps taken. */
int or_evolve(or_ctx *c, int nsteps, int gstate, int energy_every, double thr,
              double *energies, int max_energies, int *n_energies) {
    int ne = 0;
    double old = 0.0;
    int i;
    for (i = 0; i < nsteps; ++i) {
        or_step(c, i, gstate);
        if (energy_every > 0 && i % energy_every == 0) {
            double e = or_energy(c, NULL);
            if (energies && ne < max_energies) energies[ne] = e;
            ++ne;
            double olde = (i != 0) ? old : 0.0;
            old = e;
            if (i != 0 && fabs(olde - e) < thr * olde && gstate == 0) { ++i; break; }
        }
    }
    if (n_energies) *n_energies = ne;
    return i;
}

/* ------------------------------------------------------------------------ */
/* Vortex detection: src/tracker.cu:102-161 with src/minions.cu:62-93        */
/* ------------------------------------------------------------------------ */

static double cmag(or_c a) { return sqrt(a.x * a.x + a.y * a.y); }
static or_c cscale(or_c a, double s) { or_c r = {a.x * s, a.y * s}; return r; }
static or_c cdiv(or_c num, or_c den) {
    or_c cj = {den.x, -den.y};
    or_c m = {num.x * cj.x - num.y * cj.y, num.x * cj.y + num.y * cj.x};
    return cscale(m, 1.0 / (den.x * den.x + den.y * den.y));
}

int or_find_vortex(int *marker, const double *wfc_il, double radius, int xDim, const double *x) {
    const or_c *wfc = (const or_c *)wfc_il;
    int cond_x = 0, cond_y = 0, found = 0;
    for (int i = 0; i < xDim - 1; ++i) {
        for (int j = 0; j < xDim - 1; ++j) {
            if (sqrt(x[i] * x[i] + x[j] * x[j]) < radius) {
                or_c a = wfc[i * xDim + j], b = wfc[(i + 1) * xDim + j];
                or_c cc = wfc[(i + 1) * xDim + (j + 1)], d = wfc[i * xDim + (j + 1)];
                or_c g[4];
                g[0] = cscale(cdiv(a, b), cmag(b) / cmag(a));
                g[1] = cscale(cdiv(b, cc), cmag(cc) / cmag(b));
                g[2] = cscale(cdiv(cc, d), cmag(d) / cmag(cc));
                g[3] = cscale(cdiv(d, a), cmag(a) / cmag(d));
                double ph[4];
                for (int k = 0; k < 4; ++k) {
                    ph[k] = atan2(g[k].y, g[k].x);
                    if (ph[k] <= -OR_PI) ph[k] += 2 * OR_PI;
                }
                double sum = ph[0] + ph[1] + ph[2] + ph[3];
                long rnd = lround(sum / (2 * OR_PI));
                if (sum >= 1.9 * OR_PI && cond_x <= 0 && cond_y <= 0) {
                    marker[i * xDim + j] = (int)rnd; ++found; cond_x = 2; cond_y = 2;
                } else if (sum <= -1.9 * OR_PI && cond_x <= 0 && cond_y <= 0) {
                    marker[i * xDim + j] = (int)-rnd; ++found; cond_x = 2; cond_y = 2;
                }
                --cond_x; --cond_y;
            }
        }
    }
    return found;
}

/* src/tracker.cu:212-229 (vortPos): raster-ordered (i, j, winding) triples */
int or_vort_pos(const int *marker, int xDim, int *out, int max_out) {
    int n = 0;
    for (int i = 0; i < xDim; ++i)
        for (int j = 0; j < xDim; ++j)
            if (abs(marker[i * xDim + j]) >= 1) {
                if (n < max_out) { out[3 * n] = i; out[3 * n + 1] = j; out[3 * n + 2] = marker[i * xDim + j]; }
                ++n;
            }
    return n;
}

/* src/tracker.cu:317-345 (lsFit) with the matrix of include/tracker.h:46-48:
 * least-squares plane psi ~ a i + b j + c through the plaquette corners, the
 * sub-grid core at (i, j) - R with a R_i + b R_j = -c solved as GPUE does.
 * ijw: n triples (i, j, winding); out: n pairs (xd, yd). */
void or_ls_fit(const int *ijw, int n, const double *wfc_il, int xDim, double *out) {
    static const double lsq[3][4] = {{-0.5, 0.5, -0.5, 0.5}, {-0.5, -0.5, 0.5, 0.5}, {0.75, 0.25, 0.25, -0.25}};
    const or_c *wfc = (const or_c *)wfc_il;
    for (int v = 0; v < n; ++v) {
        int x = ijw[3 * v], y = ijw[3 * v + 1];
        or_c g[4] = {wfc[x * xDim + y], wfc[(x + 1) * xDim + y], wfc[x * xDim + (y + 1)], wfc[(x + 1) * xDim + (y + 1)]};
        or_c res[3];
        for (int jj = 0; jj < 3; ++jj) {
            res[jj].x = lsq[jj][0] * g[0].x + lsq[jj][1] * g[1].x + lsq[jj][2] * g[2].x + lsq[jj][3] * g[3].x;
            res[jj].y = lsq[jj][0] * g[0].y + lsq[jj][1] * g[1].y + lsq[jj][2] * g[2].y + lsq[jj][3] * g[3].y;
        }
        double det = 1.0 / (res[0].x * res[1].y - res[0].y * res[1].x);
        double Rx = det * (res[1].y * res[2].x - res[0].y * res[2].y);
        double Ry = det * (-res[1].x * res[2].x + res[0].x * res[2].y);
        out[2 * v] = x - Rx;
        out[2 * v + 1] = y - Ry;
    }
}
