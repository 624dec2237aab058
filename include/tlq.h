/*
 * tlq.h -- C-ABI of the MI355X (gfx950) HydroCore time step.
 *
 * Drop-in for the hot path of tlaloque's main loop (source/mainHD.f90:61-100).
 * The reference calls implicit-interface Fortran subroutines on module arrays
 * (source/globals.f90:5-11); a Fortran caller binds these functions with
 * ISO_C_BINDING (tlaloque_amd/fortran/tlq_c.f90).  Every function
 * returns 0 on success or a negative TLQ_ERR_* code; the last error message
 * is available from tlq_last_error().
 *
 * State lives on the GPU as structure-of-arrays (one contiguous array per
 * conserved variable, x fastest, two ghost layers).  Host arrays crossing
 * this boundary use the reference's Fortran layout: real(8) U(5, -1:nx+2,
 * -1:ny+2, -1:nz+2), component fastest (init.f90:82-88).
 *
 * Threading: one host thread per context at a time.  All work of a context
 * is ordered on one HIP stream; functions that return data synchronise it.
 */
#ifndef TLQ_H
#define TLQ_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define TLQ_OK 0
#define TLQ_ERR_ARG -1          /* invalid argument (bad id, size, pointer) */
#define TLQ_ERR_HIP -2          /* HIP runtime failure */
#define TLQ_ERR_NOMEM -3        /* device allocation failed */
#define TLQ_ERR_UNSUPPORTED -4  /* valid in the reference but not supported here */
#define TLQ_ERR_STATE -5        /* call out of order (e.g. advance before upload) */
#define TLQ_ERR_NONFINITE -6    /* non-finite or non-positive density found */

/* Boundary ids, source/constants.f90:20-23. */
#define TLQ_BC_OUTFLOW 1
#define TLQ_BC_REFLECTIVE 2
#define TLQ_BC_PERIODIC 3
#define TLQ_BC_INFLOW 4   /* a no-op stub in the reference; rejected here */

/* Limiter ids, source/constants.f90:26-31. */
#define TLQ_LIM_NONE 1
#define TLQ_LIM_VANLEER 2
#define TLQ_LIM_MINMOD 3
#define TLQ_LIM_UMIST 4
#define TLQ_LIM_WOODWARD 5
#define TLQ_LIM_SUPERBEE 6

/* Riemann solver: source/HLLC.f90 (default) or source/HLL.f90. */
#define TLQ_SOLVER_HLLC 0
#define TLQ_SOLVER_HLL 1

/* Precision: all-FP64, or FP32 storage and sweeps with the pressure recovery
 * E - rho v^2/2 and the conserved update accumulated in FP64. */
#define TLQ_FP64 0
#define TLQ_FP32 1

/* Cooling: off, exact (cooling.f90:47-119), or the learned surrogate with a
 * per-cell fallback to the exact integration (TLQ_COOL_INTEGRATED's). */
#define TLQ_COOL_OFF 0
#define TLQ_COOL_EXACT 1
#define TLQ_COOL_SURROGATE 2
/* The reference formula integrated to convergence: substeps of at most
 * 0.01 t_cool with Lambda re-read every substep (the surrogate's target and
 * fallback; one substep is exactly cooling.f90:67-112). */
#define TLQ_COOL_SUBCYCLED 3
/* The cooling equation integrated exactly for the tabulated Lambda (a power
 * law on each table segment; Townsend 2009, ApJS 181, 391): the limit of
 * TLQ_COOL_SUBCYCLED as its substeps shrink, in the same dense pass.  Cells
 * with dt <= 0.01 t_cool take cooling.f90:67-112's one step, as in
 * TLQ_COOL_SUBCYCLED.  The surrogate's integrator cells and fallbacks use the
 * same integration.  (An addition: the reference has only the one-step
 * formula.) */
#define TLQ_COOL_INTEGRATED 4

/* Self-gravity (the -DGRAV build, sources.f90:99-897): off; on with the
 * potential's ghost faces kept as uploaded (the serial build, whose neighbour
 * ranks are never negative, sources.f90:114-136); on with multipole values on
 * the faces every 10 steps (the MPI build at the domain edge). */
#define TLQ_SG_OFF 0
#define TLQ_SG_FIXED 1
#define TLQ_SG_MULTIPOLE 2

/* Device arrays of the reference's globals (source/globals.f90:5-11). */
#define TLQ_BUF_U 0
#define TLQ_BUF_UP 1
#define TLQ_BUF_UPP 2
#define TLQ_BUF_UPRIM 3
#define TLQ_BUF_F 4
#define TLQ_BUF_G 5
#define TLQ_BUF_H 6
#define TLQ_BUF_S 7

/* Problem parameters: user/parameters.f90:10-49 as run-time values. */
typedef struct tlq_params {
    int nx, ny, nz;                 /* interior cells (parameters.f90:63-65) */
    double xl, xr, yl, yr, zl, zr;  /* domain (parameters.f90:37-42) */
    double gamma;                   /* parameters.f90:43 */
    double mu0;                     /* parameters.f90:44 */
    double cfl;                     /* parameters.f90:46 */
    double eta;                     /* artificial viscosity, parameters.f90:47 */
    double cooling_limit;           /* parameters.f90:49 */
    int limiter;                    /* TLQ_LIM_*, parameters.f90:18 */
    int bc[6];                      /* LB, RB, DB, TB, FB, BB (x-, x+, y-, y+, z-, z+) */
    int solver;                     /* TLQ_SOLVER_* */
    int precision;                  /* TLQ_FP64 / TLQ_FP32 */
    int cooling;                    /* TLQ_COOL_* */
    double softening;               /* Plummer length of the point masses (0: 1/r) */
    int self_gravity;               /* TLQ_SG_*: MultiGrid inside Godunov (HLLC.f90:15-24) */
    int mg_levels;                  /* lvlm (parameters.f90:94); 0: as many as the mesh halves */
    double gconst;                  /* Gconst (parameters.f90:45); GR for the reference's */
    double mg_tol;                  /* V-cycles while Dif > mg_tol; 0: 10 (sources.f90:140) */
} tlq_params;

/* spherical_wind_type, source/winds.f90:5-13. */
typedef struct tlq_wind {
    double xc, yc, zc, vx, vy, vz, radius, mdot, vinf, temp, mu;
} tlq_wind;

typedef struct tlq_ctx tlq_ctx;

/* Lifecycle: replaces initmain's allocation (source/init.f90:82-88). */
int tlq_create(const tlq_params *params, tlq_ctx **out);
int tlq_destroy(tlq_ctx *ctx);
const char *tlq_last_error(void);

/* Sources.  Winds: the spherical_wind array driven by user_boundaries
 * (user/user.f90:57-70, winds.f90:32-90).  Point masses: x, y, z, mass per
 * mass (pointsource_gravity.f90:12-39; source term sources.f90:872-897); in
 * TLQ_FP32 mode the call also stores the masses' acceleration at every cell
 * (12 bytes per cell of device memory, allocated by the first call with
 * n > 0), which the fused step reads.
 * Cooling table: loaddata's table (cooling.f90:21-45), T and log10 Lambda. */
int tlq_set_winds(tlq_ctx *ctx, const tlq_wind *winds, int n);
int tlq_set_point_masses(tlq_ctx *ctx, const double *xyzm, int n);
int tlq_set_cooling_table(tlq_ctx *ctx, const double *temp, const double *log_lambda, int n);

/* Host <-> device copies of one global array in the reference's Fortran
 * layout (5 x (nx+4) x (ny+4) x (nz+4) doubles). */
int tlq_upload(tlq_ctx *ctx, int buf, const double *host_aos);
int tlq_download(tlq_ctx *ctx, int buf, double *host_aos);

/* Per-routine drop-ins, each acting on the device globals like its
 * reference counterpart.  dst/src name TLQ_BUF_* arrays. */
int tlq_calc_prims(tlq_ctx *ctx, int src);                        /* HydroCore.f90:1-33 */
int tlq_time_step(tlq_ctx *ctx, double *dt);                      /* HydroCore.f90:113-155 */
int tlq_fluxes(tlq_ctx *ctx, int src, int choice);               /* HLLC.f90:64-496 / HLL.f90:71-297 */
int tlq_godunov(tlq_ctx *ctx, int dst, double dt);                /* HLLC.f90:1-62 (+ S_Solver) */
int tlq_borders(tlq_ctx *ctx, int buf, int order);                /* boundaries.f90:1-26 */
int tlq_step(tlq_ctx *ctx, double *t, int *it, double dt);        /* HydroCore.f90:170-211 */
int tlq_cooling(tlq_ctx *ctx, double dt);                         /* cooling.f90:9-19 */

/* Fused time step: one iteration of mainHD.f90:74-99 in two kernels.
 * tlq_init_state performs mainHD.f90:38-44 (borders(U,1), calc_prims) on the
 * uploaded TLQ_BUF_U.  tlq_advance runs one step and, when dt/t/it are not
 * NULL, synchronises and returns them; tlq_advance_n enqueues n steps
 * without synchronising. */
int tlq_init_state(tlq_ctx *ctx);
int tlq_advance(tlq_ctx *ctx, double *dt, double *t, int *it);
int tlq_advance_n(tlq_ctx *ctx, int n);
int tlq_get_time(tlq_ctx *ctx, double *dt, double *t, int *it);
/* Restart: continue a run from a state uploaded with tlq_upload(U), at time
 * t after `it` steps; in place of tlq_init_state.  fill_ghosts = 0 keeps the
 * uploaded ghosts (a BIN dump holds them, and U's ghosts are images of UP,
 * not of U's interior, HydroCore.f90:195-206), so the run continues to
 * rounding; fill_ghosts = 1 fills them from the interior first (a VTK dump
 * holds the interior only).  Replaces the unfinished warmstart
 * (warmstart.f90:1-10; its autosave #warmstart.f90#:11-13 reads a VTK). */
int tlq_resume(tlq_ctx *ctx, double t, int it, int fill_ghosts);
int tlq_sync(tlq_ctx *ctx);

/* Self-gravity.  PHI(0:nx+1, 0:ny+1, 0:nz+1) (globals.f90:19, i fastest) in
 * FP64: upload its initial values (the faces are the boundary values) before
 * tlq_init_state, as user_initconds sets PHI (user.f90:139-141), and read it
 * back.  tlq_init_state then solves once (mainHD.f90:50-52); every fused step
 * solves once from U^n and both stages add S_Solver's source with
 * PHIT = PHIP + PHI (HLLC.f90:15-24; the reference solves twice from the same
 * U^n, mainHD.f90:80,88); tlq_godunov solves at every call, as the
 * reference.  tlq_gravity_solve runs one MultiGrid call on the current U at
 * the current step count; *cycles and *dif receive its V-cycles and final
 * Dif.  tlq_gravity_info: levels (-1 when off), the last solve's cycles and
 * Dif, and the step count the mod(it, 10) test reads. */
int tlq_upload_phi(tlq_ctx *ctx, const double *phi);
int tlq_download_phi(tlq_ctx *ctx, double *phi);
int tlq_gravity_solve(tlq_ctx *ctx, int *cycles, double *dif);
int tlq_gravity_info(tlq_ctx *ctx, int *levels, int *cycles, double *dif, int *it);

/* Diagnostics: device-side check that every interior density is finite and
 * positive and every energy finite; *bad receives the count of bad cells. */
int tlq_check_state(tlq_ctx *ctx, long long *bad);
/* Of the interior cells handed to the network in the last surrogate-cooling
 * step, the fraction whose answer failed the check and took the converged
 * integrator instead. */
int tlq_fallback_fraction(tlq_ctx *ctx, double *frac);
/* The same step's counts: cells handed to the network (mode 2: T >= T_min and
 * dt >= the threshold below, in units of t_cool) or to the integrator (modes
 * 3 and 4: T >= T_min and dt > 0.01 t_cool; below 0.01 t_cool one substep of
 * cooling.f90:67-112 is the converged answer), and, of the network's cells,
 * those that took the exact integration instead (mode 2). */
int tlq_cooling_counts(tlq_ctx *ctx, long long *queued, long long *fallbacks);

/* Learned cooling surrogate (MLP 3 -> 64 -> 64 -> 1, GELU, bf16 weights,
 * FP32 accumulation; TLQ_MLP_NPARAMS parameters laid out as W1[64][3], b1[64],
 * W2[64][64], b2[64], W3[64], b3).  No counterpart in the reference, which
 * has only the exact formula; the training targets come from that formula
 * integrated to convergence (TLQ_COOL_SUBCYCLED).  Inputs per sample:
 * x = (log10 n [cm^-3], log10 T [K], log10(dt / t_cool)) in the training box
 * [-3,4] x [4,9] x [-2,2]; output g = r / r1 with r = log10(T_new / T) and
 * r1 = -log10(e) d / (1 + d), d = dt / t_cool. */
#define TLQ_MLP_NPARAMS 4481
/* Random initial parameters (seeded) and fresh Adam moments. */
int tlq_surrogate_init(tlq_ctx *ctx, unsigned long long seed);
/* In TLQ_COOL_SURROGATE mode, cells with 0.01 < dt / t_cool < dtc_min take
 * the exact integration (TLQ_COOL_INTEGRATED's) instead of the network:
 * there the integrators cost less than the network's 128 activations.
 * Default TLQ_SURROGATE_MIN_DTC; 0.01 or less hands every such cell to the
 * network. */
#define TLQ_SURROGATE_MIN_DTC 0.3
int tlq_surrogate_set_threshold(tlq_ctx *ctx, double dtc_min);
/* FP32 master parameters in / out (n = TLQ_MLP_NPARAMS); setting them also
 * readies the context for TLQ_COOL_SURROGATE. */
int tlq_surrogate_set_params(tlq_ctx *ctx, const float *theta, int n);
int tlq_surrogate_get_params(tlq_ctx *ctx, float *theta, int n);
/* nsamples uniform draws of the training box with their targets g, made on
 * the GPU by the FP64 converged integrator (kept on the device). */
int tlq_surrogate_generate(tlq_ctx *ctx, long long nsamples, unsigned long long seed);
/* Copies samples [first, first + n) out: x as n triples, y as n targets. */
int tlq_surrogate_samples(tlq_ctx *ctx, long long first, long long n, float *x, float *y);
/* steps of Adam on the mean squared error, each over the next contiguous batch
 * of the (independently drawn) generated samples, cycling through the set;
 * learning rate on a cosine from lr_max to lr_min. */
int tlq_surrogate_train(tlq_ctx *ctx, int steps, int batch, double lr_max, double lr_min,
                        double *final_loss);
/* Network output g for n input triples. */
int tlq_surrogate_eval(tlq_ctx *ctx, const float *x, long long n, float *out);
/* Gradient (TLQ_MLP_NPARAMS values) and value of the mean squared error of
 * the network on the n samples (x, y) given. */
int tlq_surrogate_grad(tlq_ctx *ctx, const float *x, const float *y, int n, double *grad,
                       double *loss);
/* Allocation-free variant of tlq_surrogate_grad.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight training loops).  x (n triples), y (n
 * targets), grad (TLQ_MLP_NPARAMS floats) and loss (1 float) are DEVICE
 * pointers; the call only enqueues two kernels on the context's stream
 * (read the results after tlq_sync) and allocates nothing: the partial
 * gradients go to a static scratch array of 256 blocks x (TLQ_MLP_NPARAMS + 1)
 * floats (4.6 MB) that each GPU holds until the process ends, so n is at
 * most TLQ_NOALLOC_MAX_SAMPLES (TLQ_ERR_ARG otherwise).  Every call on one
 * GPU shares that array (each call rewrites all it reads): calls from
 * contexts whose streams overlap on one GPU race.  Results equal
 * tlq_surrogate_grad's for the same samples, bit for bit. */
#define TLQ_NOALLOC_MAX_SAMPLES 16384
int tlq_surrogate_grad_noalloc(tlq_ctx *ctx, const float *x, const float *y, int n, float *grad,
                               float *loss);

/* Timing: HIP events on the context's stream.  tlq_timer_* brackets any
 * enqueued work; with kernel timing on, every kernel launch of the fused step
 * is bracketed and its durations accumulate per kernel (names as in
 * tlq_kernel_stats: "stage1", "stage2", "init", "cooling", "multigrid" (whole
 * self-gravity solves), "mg_relax0" (the finest level's relaxation launches)). */
int tlq_timer_start(tlq_ctx *ctx);
int tlq_timer_stop(tlq_ctx *ctx, double *ms);
int tlq_kernel_timing(tlq_ctx *ctx, int enable);
int tlq_kernel_stats(tlq_ctx *ctx, int index, char *name, int name_len, double *total_ms,
                     long long *launches);

/* Streaming-copy probe on the context's stream: copies `bytes` bytes between
 * two device buffers `iters` times and returns the mean time of one copy
 * (calibrates HBM counters and bandwidth; no counterpart in the reference). */
int tlq_copy_probe(tlq_ctx *ctx, long long bytes, int iters, double *ms_per_copy);

/* The same copy in one configuration: `unroll` (1, 2, 4 or 8) independent
 * 16-byte loads in flight per thread, `blocks` blocks of 256 threads, and
 * nontemporal loads and stores when `nontemporal` is nonzero (a sweep over
 * these finds the practical HBM ceiling; no counterpart in the reference). */
int tlq_copy_probe_config(tlq_ctx *ctx, long long bytes, int iters, int unroll, int blocks,
                          int nontemporal, double *ms_per_copy);

/* Sizes: bytes of one device field and the number of cells per field. */
int tlq_layout(tlq_ctx *ctx, long long *cells_per_field, int *pitch_x, int *origin_x);

#ifdef __cplusplus
}
#endif
#endif /* TLQ_H */
