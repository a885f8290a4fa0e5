/*
 * gpue_amd.h -- C ABI of libgpue_amd.so, the MI355X (gfx950) split-operator
 * step that stands in for GPUE's evolution/split_op path.
 *
 * Conventions
 *  - Every call returns a gpue_status; the library never exits the process
 *    (GPUE's cudaHandleError/cudaCheckError exit(1), src/split_op.cu:40-53).
 *    gpue_last_error() gives the message of the last failing call on a ctx.
 *  - Host buffers are caller-owned and only touched during the call.
 *  - psi is complex, interleaved (re, im) doubles, index x + xDim*(y + yDim*z)
 *    (x fastest), exactly GPUE's wfc layout (src/kernels.cu:28-35).
 *  - One gpue_ctx is driven by one host thread at a time; contexts are
 *    independent (each owns its device memory and one HIP stream).
 *  - Device memory is allocated only by gpue_create / gpue_set_field* and freed
 *    by gpue_destroy; the step loop allocates nothing.
 */
#ifndef GPUE_AMD_H
#define GPUE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define GPUE_ABI_VERSION 1

typedef struct gpue_ctx gpue_ctx;

typedef enum {
    GPUE_OK = 0,
    GPUE_EINVAL = 1,        /* bad argument / parameter */
    GPUE_ENOMEM = 2,        /* device allocation failed */
    GPUE_EUNSUPPORTED = 3,  /* grid size or feature not implemented */
    GPUE_EHIP = 4,          /* HIP runtime error (message in gpue_last_error) */
    GPUE_ENONFINITE = 5,    /* norm or energy became NaN/Inf */
    GPUE_EBUDGET = 6,       /* serving: request exceeds the HBM budget */
    GPUE_EPARSE = 7,        /* equation string rejected */
    GPUE_ETIMEOUT = 8       /* serving: gpue_server_drain ran out of time */
} gpue_status;

enum { GPUE_FIELD_V = 0, GPUE_FIELD_K = 1, GPUE_FIELD_AX = 2, GPUE_FIELD_AY = 3, GPUE_FIELD_AZ = 4 };
enum { GPUE_A_ROTATION = 0, GPUE_A_CONSTANT = 1 };   /* GPUE -A rotation | constant (src/ds.cu:300-331) */

/* GPUE's parameter set for the step (keys of Grid, include/ds.h:86-197, with the
 * defaults of src/parser.cu:14-67).  hbar, mass and g are explicit so that
 * harmonic-oscillator units can be used; gpue_params_default() fills GPUE's SI
 * values (HBAR include/constants.h:20, Rb-87 mass src/init.cu:98). */
typedef struct {
    int dimnum, xDim, yDim, zDim;
    double xMax, yMax, zMax;        /* box half-widths: box_size, or 6*Rxy*a0 (src/init.cu:124-137) */
    double hbar, mass;
    double g;                       /* interaction * gDenConst (src/evolution.cu:739) */
    double omega, omegaX, omegaY, omegaZ, gammaY;
    double x0_shift, y0_shift, z0_shift;
    double gdt, dt;                 /* imaginary / real time step (src/parser.cu:21-22) */
    int gpe;                        /* -s: nonlinear term on */
    int corotating;                 /* -l: gauge (A.p) operators on */
    int a_kind;                     /* GPUE_A_ROTATION | GPUE_A_CONSTANT (tables via gpue_set_field) */
    int precision;                  /* 64: FP64 throughout; 32: FP32 bulk, FP64 reductions */
    int device;                     /* HIP device ordinal (-C, src/parser.cu:168) */
} gpue_params;

typedef struct {
    int steps_done;     /* steps executed (early stop included) */
    int early_stop;     /* 1 if the imaginary-time energy criterion stopped the loop */
    int n_energies;     /* energies evaluated (cadence hits) */
    double energy;      /* last evaluated energy (GPUE formula), 0 if none */
    double mu;          /* chemical potential at the same point */
} gpue_evolve_result;

/* GPUE Grid defaults (src/parser.cu:14-67) with dimnum/xDim/... as given there. */
gpue_status gpue_params_default(gpue_params *p);

/* Replaces init()'s field/operator construction and plan creation for one problem
 * (src/init.cu:40-356, src/operators.cu:722-992).  Fields default to GPUE's
 * kharmonic_V (incl. 1/2 m |A|^2), krotation_A / kconstant_A and simple_K. */
gpue_status gpue_create(const gpue_params *p, gpue_ctx **out);
void        gpue_destroy(gpue_ctx *c);
const char *gpue_last_error(const gpue_ctx *c);

/* Static field from a host table of N doubles (GPUE -A file / file_A, and the
 * V/K/A arrays GPUE keeps in Grid: src/operators.cu:505-520). */
gpue_status gpue_set_field(gpue_ctx *c, int which, const double *host_N);
/* Field from a GPUE equation string (src/dynamic.cu:187-470).  Expressions
 * without 't' are evaluated once on the device at find_field's coordinates
 * (src/dynamic.cu:605-616); a V with 't' is compiled into the fused V/2 pass.
 * `names`/`values` supply the Grid doubles the expression may reference. */
gpue_status gpue_set_field_eqn(gpue_ctx *c, int which, const char *expr, const char *const *names,
                               const double *values, int n_names);

/* set_variables' H2D of psi (src/init.cu:422,470) / D2H (src/evolution.cu:430). */
gpue_status gpue_upload_wfc(gpue_ctx *c, const double *interleaved);
gpue_status gpue_download_wfc(gpue_ctx *c, double *interleaved);

/* evolve() (src/evolution.cu:260-908) without host I/O: steps i = first_step ..
 * first_step+nsteps-1 of GPUE's loop, gstate 0 = imaginary, 1 = real time.
 * energy_every > 0 evaluates energy_calc after steps with i % energy_every == 0,
 * storing up to max_energies values, and stops early in imaginary time when
 * |E_old - E| < energy_thr * E_old (src/evolution.cu:872-903). */
gpue_status gpue_evolve(gpue_ctx *c, int first_step, int nsteps, int gstate, int energy_every,
                        double energy_thr, double *energies, int max_energies, gpue_evolve_result *res);

/* parSum (src/split_op.cu:83-129): psi <- psi / sqrt(sum |psi|^2 dV). */
gpue_status gpue_parsum(gpue_ctx *c);

/* energy_calc (src/split_op.cu:208-375), plus the chemical potential (same sum
 * with the interaction weighted g instead of g/2). */
gpue_status gpue_energy(gpue_ctx *c, double *energy, double *mu);

/* energy_calc with its gauge term replaced by the one of the Hamiltonian the
 * step evolves.  GPUE's energy adds +i hbar A.grad (forward differences,
 * src/split_op.cu:312-360), while apply_gauge evolves exp(-A_a k_a Dt)
 * (src/operators.cu:1089-1110), i.e. +hbar A.k: the two gauge terms have
 * opposite signs.  Here hbar sum_a <A_a k_a> is taken spectrally (FFT along
 * each gauge axis, Parseval), so E is the quantity real-time steps conserve and
 * imaginary-time steps lower.  Equal to gpue_energy without a gauge term
 * (corotating off, or 1D).  The -R ramp factor is not applied. */
gpue_status gpue_energy_consistent(gpue_ctx *c, double *energy, double *mu);

/* Tracker::findVortex (src/tracker.cu:102-161) on the current 2D psi:
 * marker_out (xDim*xDim ints, overwritten) and the count of hits. */
gpue_status gpue_vortex_marker_2d(gpue_ctx *c, double mask_radius, int *marker_out, int *found);
/* No counterpart in the reference: gpue_vortex_marker_2d for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same markers and count; its scratch
 * (plaquette mask, candidates, marker) is a fixed device array of the library,
 * one per GPU, held until the process ends, for square grids up to 1024^2
 * (larger: GPUE_EINVAL).  Calls on one GPU share that array, so calls that
 * overlap on one GPU race; each call returns after its results are copied. */
gpue_status gpue_vortex_marker_2d_noalloc(gpue_ctx *c, double mask_radius, int *marker_out, int *found);
/* findVortex + vortPos (src/tracker.cu:212-229): raster-ordered (i, j, winding). */
gpue_status gpue_vortex_list_2d(gpue_ctx *c, double mask_radius, int *ijw, int max_vortices, int *n_vortices);

/* x, y, z (axis 0..2) and px, py, pz (3..5) grids of generate_p_space
 * (src/operators.cu:277-432). */
gpue_status gpue_get_grid(gpue_ctx *c, int axis, double *out);

/* Interaction change between phases (quench).  GPUE has one `interaction`
 * (src/evolution.cu:739); this is the ABI extension the quench workload needs. */
gpue_status gpue_set_g(gpue_ctx *c, double g);

/* GPUE's -R ramp (src/evolution.cu:409-428, 811-832): 0 off; 1 or 2 scale the
 * gauge exponent of step i (counted from 0 within the ramped evolve call of
 * num_steps steps) by omega/(i+1) or omega*(i+1)/num_steps, GPUE's cumulative
 * omega_0.  Steps of a ramped context are launched one by one (no graphs). */
gpue_status gpue_set_ramp(gpue_ctx *c, int type, int num_steps);

/* Wait for all work on the ctx stream. */
gpue_status gpue_sync(gpue_ctx *c);

/* FFT kernel shape per axis (no reference counterpart: cuFFT plans choose their
 * own).  gpue_autotune times every variant of every sweep of the current
 * configuration (graph replays of back-to-back launches in place on copies
 * of psi, as gpue_sweep_profile) and keeps the fastest per axis, then refines
 * the choice over whole graph-replayed steps on a copy of the state;
 * variants[3] receives the choice (0..5: bit 0 the plan, 0 = wide lines,
 * 1 = narrow lines with fewer points per lane; variant >> 1 = 0, 1, 2 narrows
 * the tiles of the y/z sweeps to 1, 1/2 or 1/4 of their default width, to
 * spread small grids over more CUs, and when > 0 each transform loads its
 * twiddles before its first pass; the x axis takes 0..3, where 4, 5 act as
 * 2, 3).  gpue_set_variants forces a choice. */
gpue_status gpue_autotune(gpue_ctx *c, int reps, int *variants);
gpue_status gpue_set_variants(gpue_ctx *c, const int *variants);
/* The kernel variants in use per axis (what gpue_set_variants or an autotuner set). */
gpue_status gpue_get_variants(const gpue_ctx *c, int *variants);
/* Steady-state steps are captured as HIP graphs of up to `steps` consecutive
 * steps (even, default 8) and replayed; 0 launches every sweep individually.
 * No reference counterpart (GPUE launches each kernel from the host loop). */
gpue_status gpue_set_graph_steps(gpue_ctx *c, int steps);
/* Capture, without launching, the step blocks that gpue_evolve with the same
 * first_step, nsteps, gstate and energy_every would replay (blocks of up to
 * `steps` steps; shorter ones end the call and precede each energy
 * evaluation), so that call spends no time on captures.  No reference
 * counterpart. */
gpue_status gpue_prepare_evolve(gpue_ctx *c, int first_step, int nsteps, int gstate, int energy_every);

/* --- measurement hooks (no reference counterpart: GPUE prints clock() deltas,
 * src/evolution.cu:436-441) --- */
/* When on, gpue_evolve records a HIP event pair around every sweep launch. */
gpue_status gpue_set_timing(gpue_ctx *c, int on);
/* Per-sweep-kind accumulated kernel time: kind index, name, launches, ms. */
gpue_status gpue_timing_get(gpue_ctx *c, int kind, const char **name, long *launches, double *ms);
/* HBM bytes of psi one steady-state step moves in the current configuration:
 * 2 x (bytes of psi) per sweep, averaged over the step parities and, when graphs
 * are on, over a captured block (real time fuses the sweeps at step boundaries);
 * *sweeps_per_step receives the sweep count. */
gpue_status gpue_step_bytes(gpue_ctx *c, int gstate, double *bytes_per_step, double *sweeps_per_step);
/* Per-kernel cost of the step (no reference counterpart): each distinct sweep
 * of the gstate schedule is launched `reps` times back to back from a graph,
 * in place in the scratch buffer (psi is untouched): each launch on a fresh
 * copy of psi, the copies timed alone and subtracted, except for states of
 * 256 MiB and more in imaginary time, which run on one copy; HIP events on the context
 * stream.  Fills up to max_kinds entries: kernel name (owned by the
 * ctx, valid until the next call), microseconds per launch, launches per step
 * (averaged over the two gauge orders); *n_kinds = number of distinct sweeps. */
gpue_status gpue_sweep_profile(gpue_ctx *c, int gstate, int reps, int max_kinds, int *n_kinds,
                               const char **names, double *us_per_launch, double *launches_per_step);
/* The exp (fn 0), sin (1) and cos (2) the FP64 fused operators evaluate on the
 * device, the sin (3) / cos (4) of the FP32 ones (input rounded to float), and
 * the short small-argument forms exp (5), sin (6), cos (7) used for |x| <= 1/32,
 * applied to n host values (verification of the device math; no reference
 * counterpart). */
gpue_status gpue_math_eval(int device, int fn, const double *x, double *y, long n);

/* --- equation strings (host-side, src/dynamic.cu:187-522) --- */
typedef struct {          /* EqnNode_gpu (include/ds.h:71-80) */
    double val;
    int is_dynamic;
    char var;
    int left, right;
    int op_num;
} gpue_eqn_node;

gpue_status gpue_eqn_parse(const char *expr, const char *const *names, const double *values, int n_names,
                           gpue_eqn_node *nodes, int max_nodes, int *n_nodes, int *time_dependent);
double      gpue_eqn_eval(const gpue_eqn_node *nodes, int n_nodes, double x, double y, double z, double t);

/* --- batched instances (no reference counterpart: GPUE solves one problem per
 * process; its bin/run.sh starts several processes).  Instances of one shape
 * share every sweep launch; the shape groups run on at most four streams (the
 * device's hardware queues per process: the three groups with the most work
 * get one each, the rest share the fourth), two steps of a group captured as
 * a HIP graph and replayed, so the groups overlap on the device.
 * The kernel variants are the contexts' own (equal within a group; a batch
 * may prefer other variants than a lone instance: the groups share the
 * device).  Contexts stay owned by the caller and must outlive the
 * batch; results are read through the contexts (gpue_energy, gpue_download_wfc).
 * Contexts with a time-dependent V or on different devices are rejected. --- */
typedef struct gpue_batch gpue_batch;
gpue_status gpue_batch_create(gpue_ctx *const *ctxs, int n, gpue_batch **out);
/* Exclude (0) or include (1) instance `index` from the following steps. */
gpue_status gpue_batch_set_active(gpue_batch *b, int index, int active);
/* Advance every active instance by nsteps steps (gstate 0 imaginary, 1 real);
 * each instance continues from its own step index (GPUE's gauge-order parity). */
gpue_status gpue_batch_evolve(gpue_batch *b, int nsteps, int gstate);
/* Energy and chemical potential of every instance (arrays of n, in the order
 * of gpue_batch_create), evaluated on all instances' streams at once and
 * waited for once: the values are those gpue_energy returns per context.
 * An instance with a separable K and static fields on its own stream has its
 * evaluation captured as one HIP graph at the first call and replayed after;
 * any change of its parameters, fields or variants drops the graph. */
gpue_status gpue_batch_energy(gpue_batch *b, double *energy, double *mu);
/* How pairs of steps are replayed: 0 (default) one graph per shape group on the
 * group's stream, launched group by group, most work first; 1 one graph for the whole batch, a
 * branch per group forked onto the group streams, so a pair of steps of every
 * instance is one launch; 2 no graphs (every sweep launched on its own, for
 * kernel tracers). */
gpue_status gpue_batch_set_graph_mode(gpue_batch *b, int mode);
/* No counterpart in the reference (GPUE runs one problem per process): choose
 * the kernel variants of every shape group, and the graph mode, for the batch
 * as a whole (a lone instance's choice serves latency; in a batch the groups
 * share the device).  Coordinate descent over the groups, largest first, and
 * their axes; each candidate is timed over `steps` steps of gstate of the whole
 * batch.  Every instance's psi, step counter and pending norm are restored
 * afterwards.  us_per_step, graph_mode (either may be NULL): the best time per
 * batch step and the graph mode chosen. */
gpue_status gpue_batch_autotune(gpue_batch *b, int gstate, int steps, double *us_per_step, int *graph_mode);
const char *gpue_batch_last_error(const gpue_batch *b);
void        gpue_batch_destroy(gpue_batch *b);

/* --- solve service (no reference counterpart: GPUE runs one problem per
 * process, bin/run.sh starts several).  A worker thread admits queued requests
 * in arrival order while a slot and the HBM budget allow, keeps every state
 * resident on the device and advances each admitted request on its own stream
 * in chunks of check_every steps, each followed by an energy evaluation.  A
 * request's result does not depend on the other requests in flight. --- */
typedef struct {
    gpue_params params;      /* problem; params.device is overridden by the server's */
    unsigned long long seed; /* psi0 = exp(-r^2/2)(1 + 0.01 u), u uniform from the seed */
    int random_phase;        /* 1: psi0 also gets a uniformly random phase per point */
    int check_every;         /* steps between energy checks (<=0: 100) */
    int max_imag_steps;      /* imaginary-time step limit */
    double energy_tol;       /* stop imaginary time when |E_k - E_k-1| < tol |E_k-1| */
    int real_steps;          /* real-time steps after the ground state (0: none) */
    double quench_factor;    /* g is multiplied by this before real time */
    double *psi_out;         /* optional: 2*N doubles receiving the final psi
                                (must stay valid until the result is polled) */
} gpue_request;

typedef struct {
    unsigned long long id;
    int status;              /* gpue_status of the request */
    int imag_steps, real_steps;
    double energy, mu;       /* at the last check */
    double t_submit, t_start, t_done;   /* seconds on the server clock */
    double t_dev_start, t_dev_done;     /* seconds on the device clock since the server started:
                                           the request's first GPU work and the end of its last chunk */
} gpue_result;

typedef struct gpue_server gpue_server;
/* budget_bytes: device memory the admitted requests may hold together;
 * max_slots: requests in flight at once. */
gpue_status gpue_server_create(int device, double budget_bytes, int max_slots, gpue_server **out);
/* Queue a request; GPUE_EBUDGET if it can never fit the budget. Thread-safe. */
gpue_status gpue_server_submit(gpue_server *s, const gpue_request *r, unsigned long long id);
/* Move up to max finished results to out; returns how many. Thread-safe. */
int         gpue_server_poll(gpue_server *s, gpue_result *out, int max);
/* Wait until every submitted request has finished (GPUE_ETIMEOUT on timeout). */
gpue_status gpue_server_drain(gpue_server *s, double timeout_s);
double      gpue_server_clock(const gpue_server *s);
/* Finishes the requests in flight and queued, then stops the worker. */
void        gpue_server_destroy(gpue_server *s);

#ifdef __cplusplus
}
#endif
#endif /* GPUE_AMD_H */
