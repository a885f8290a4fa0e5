/*
 * gpue_oracle.h -- CPU FP64 restatement of GPUE's split-operator time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  No product
 * code path (gpue_amd/, libgpue_amd.so) links, loads or calls it.
 *
 * Every function cites the GPUE source it restates (paths relative to the
 * GPUE repository root).  The FFT is this project's own mixed-radix FP64
 * Stockham transform (FFTW is not available offline); it is pinned against a
 * brute-force DFT and numpy.fft golden vectors in tests/.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   or_fft*            brute-force DFT (n<=64), numpy.fft fixtures, Parseval
 *   or_parsum          GPUE KAT src/unit_test.cu:841-947 (2+2i on 64^2 / 16^3)
 *   or_energy          closed forms: HO ground state E = d/2 (src/unit_test.cu:1162-1278),
 *                      plane-wave kinetic energy, forward-difference gauge term on
 *                      analytic fields, Thomas-Fermi bound
 *   or_gauge_spectral  plane waves (closed form hbar A k0 <1>), and the forward-difference
 *                      term of or_energy on smooth fields (equal and opposite to O(dx))
 *   or_step/or_evolve  matlab/GPE_2d.m ordering transcribed independently in numpy
 *                      (tests/golden/make_golden.py), unitarity in real time,
 *                      linear HO ground state
 *   or_find_vortex     imprinted phase windings at known plaquettes incl. the
 *                      skip rule, sign loss and mask (src/tracker.cu:102-161)
 */
#ifndef GPUE_ORACLE_H
#define GPUE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y; } or_c;   /* mirrors CUDA double2 */

enum { OR_V_HARMONIC = 0, OR_V_TABLE = 1 };
enum { OR_A_ROTATION = 0, OR_A_ZERO = 1, OR_A_TABLE = 2 };
enum { OR_K_STANDARD = 0, OR_K_TABLE = 1 };
enum { OR_FIELD_V = 0, OR_FIELD_K = 1, OR_FIELD_AX = 2, OR_FIELD_AY = 3, OR_FIELD_AZ = 4 };

typedef struct {
    int dimnum, xDim, yDim, zDim;
    double xMax, yMax, zMax;          /* box half-widths (box_size or 6*Rxy*a0) */
    double hbar, mass, g;             /* g = interaction * gDenConst */
    double omega, omegaX, omegaY, omegaZ, gammaY;
    double x0_shift, y0_shift, z0_shift;
    double gdt, dt;
    int gpe, corotating;
    int v_kind, a_kind, k_kind;
} or_params;

typedef struct or_ctx or_ctx;

/* --- FFT (own FP64 mixed-radix Stockham; unnormalised; sign -1 = forward) --- */
void or_fft_lines(or_c *data, int n, int howmany, long stride, long dist, int sign);
void or_fft_nd(or_c *data, int xDim, int yDim, int zDim, int sign);
void or_dft_bruteforce(const or_c *in, or_c *out, int n, int sign);

/* --- context: grids, fields, operators, psi --- */
or_ctx *or_create(const or_params *p);
void    or_destroy(or_ctx *c);
void    or_set_field(or_ctx *c, int which, const double *vals);   /* switches kind to TABLE */
void    or_rebuild(or_ctx *c);                                   /* re-derive V (incl. 1/2 m A^2) and ops */
void    or_get_field(const or_ctx *c, int which, double *out);
void    or_get_grid(const or_ctx *c, int axis, double *out);      /* 0..2 x,y,z ; 3..5 px,py,pz */
void    or_set_wfc(or_ctx *c, const double *interleaved);
void    or_get_wfc(const or_ctx *c, double *interleaved);
void    or_set_g(or_ctx *c, double g);
void    or_set_threads(int n);     /* OpenMP threads of the following calls */
int     or_max_threads(void);
double  or_dV(const or_ctx *c);

/* --- hot path --- */
void    or_step(or_ctx *c, int i, int gstate);
void    or_parsum(or_ctx *c);
double  or_energy(or_ctx *c, double *mu_out);
/* the gauge term hbar <A.k> of the evolved Hamiltonian (spectral, per axis), and
 * energy_calc with it in place of GPUE's +i hbar A.grad forward difference */
double  or_gauge_spectral(or_ctx *c);
double  or_energy_consistent(or_ctx *c, double *mu_out);
int     or_evolve(or_ctx *c, int nsteps, int gstate, int energy_every, double thr,
                  double *energies, int max_energies, int *n_energies);

/* -R ramp of the gauge operators (0 off, 1 or 2; num_steps = steps of the
 * evolve call, step indices count from 0 within it) */
void    or_set_ramp(or_ctx *c, int type, int num_steps);
double  or_ramp_factor(int type, double omega, int i, int num_steps);

/* --- standalone kernels restated (for KATs) --- */
void    or_parsum_raw(or_c *wfc, int dimnum, int xDim, int yDim, int zDim,
                      double dx, double dy, double dz);
int     or_find_vortex(int *marker, const double *wfc_interleaved, double radius,
                       int xDim, const double *x);
int     or_vort_pos(const int *marker, int xDim, int *out_ij_wind, int max_out);
void    or_ls_fit(const int *ijw, int n, const double *wfc_interleaved, int xDim, double *out_xy);

#ifdef __cplusplus
}
#endif
#endif
