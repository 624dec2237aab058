/*
 * tlq_oracle.h -- FP64 CPU restatement of the tlaloque HydroCore time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (tlaloque_amd, libtlq.so) never calls it.
 *
 * Arrays use the reference's Fortran layout: U(e, i, j, k) with
 * e = 1..5 fastest, then i = -1..nx+2, j = -1..ny+2, k = -1..nz+2
 * (user/parameters.f90:81-87, init.f90:82-88).  In C the component index e is
 * 0-based and the cell indices keep the Fortran values (ghosts -1, 0, n+1, n+2).
 */
#ifndef TLQ_ORACLE_H
#define TLQ_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary ids, source/constants.f90:20-23. */
enum { ORC_BC_OUTFLOW = 1, ORC_BC_REFLECTIVE = 2, ORC_BC_PERIODIC = 3, ORC_BC_INFLOW = 4 };
/* Limiter ids, source/constants.f90:26-31. */
enum { ORC_LIM_NONE = 1, ORC_LIM_VANLEER = 2, ORC_LIM_MINMOD = 3, ORC_LIM_UMIST = 4,
       ORC_LIM_WOODWARD = 5, ORC_LIM_SUPERBEE = 6 };
/* Riemann solvers: HLLC.f90 (linked by the reference) and HLL.f90. */
enum { ORC_SOLVER_HLLC = 0, ORC_SOLVER_HLL = 1 };

typedef struct orc_wind {          /* spherical_wind_type, source/winds.f90:5-13 */
    double xc, yc, zc, vx, vy, vz, radius, mdot, vinf, temp, mu;
} orc_wind;

typedef struct orc_params {
    int nx, ny, nz;                /* interior cells, user/parameters.f90:63-65 */
    double xl, xr, yl, yr, zl, zr; /* user/parameters.f90:37-42 */
    double gamma, mu0, cfl, eta;   /* user/parameters.f90:43-47 */
    double cooling_limit;          /* user/parameters.f90:49 */
    int limiter;                   /* user/parameters.f90:18 */
    int bc[6];                     /* LB, RB, DB, TB, FB, BB: user/parameters.f90:10-15 */
    int solver;                    /* ORC_SOLVER_* */
    int cooling;                   /* 0 off, 1 apply_cooling (cooling.f90:47-119) */
    int n_winds;                   /* user_boundaries loop, user/user.f90:64-68 */
    const orc_wind *winds;
    int n_masses;                  /* point masses: x, y, z, mass (4 doubles each) */
    const double *masses;
    double softening;              /* Plummer length; 0 gives pointsource_gravity.f90:32-33 */
    int n_table;                   /* cooling table rows (cooling.f90:32-39) */
    const double *table_T;         /* T_i [K] */
    const double *table_L;         /* log10 Lambda_i (= -column 2 of Z1.0.dat) */
} orc_params;

/* Number of doubles in one 5-field array with two ghost layers. */
size_t orc_array_len(const orc_params *p);

/* Individual routines (each mirrors one reference routine). */
void orc_calc_prims(const orc_params *p, double *prim, const double *uu);
double orc_time_step(const orc_params *p, const double *prim);
void orc_fluxes(const orc_params *p, const double *uu, const double *prim,
                double *ff, double *gg, double *hh, int choice);
void orc_godunov(const orc_params *p, double *uout, const double *uu, const double *ff,
                 const double *gg, const double *hh, const double *src, double dt);
void orc_boundaries1(const orc_params *p, double *u);
void orc_boundaries2(const orc_params *p, double *u);
void orc_winds(const orc_params *p, double *u);
void orc_step(const orc_params *p, double *u, double *up, double *upp, double dt);
void orc_cooling(const orc_params *p, double *prim, double *u, double dt);
double orc_find_coef(const orc_params *p, double temp);
void orc_gravity_source(const orc_params *p, const double *prim, double *src);
void orc_limiter_avg(int lim, const double *a, const double *b, double *out, int n);
void orc_hllc_face(const orc_params *p, const double *pl, const double *pr,
                   const double *ul, const double *ur, int dim, double *f);
void orc_hll_face(const orc_params *p, const double *pl, const double *pr,
                  const double *ul, const double *ur, int dim, double *f);

/* One full iteration of mainHD.f90:74-99 on state (u, up, upp, prim) with
 * scratch fluxes (ff, gg, hh) and, when point masses are set, src.  Returns dt. */
double orc_advance(const orc_params *p, double *u, double *up, double *upp, double *prim,
                   double *ff, double *gg, double *hh, double *src);

/* Self-gravity (source/sources.f90:99-897, HLLC.f90:15-24). */
typedef struct orc_sg {
    int levels;      /* lvlm, user/parameters.f90:94 */
    int boundary;    /* 1: phi's ghost faces kept (serial build); 2: multipole
                        values every 10 steps (MPI build, domain edge) */
    double gconst;   /* Gconst, user/parameters.f90:45 */
    double *phi;     /* PHI(0:nx+1, 0:ny+1, 0:nz+1), i fastest */
    double *work;    /* orc_sg_work_len doubles: Error, Residue, ret per level */
    int it;          /* step counter (globals it) */
    int cycles;      /* out: V-cycles of the last MultiGrid */
    double dif;      /* out: its final Dif */
    double tol;      /* V-cycles while Dif > tol; 0: the reference's 10 */
} orc_sg;
size_t orc_sg_work_len(const orc_params *p, int levels);
int orc_multigrid(const orc_params *p, orc_sg *sg, const double *dens);
void orc_moments(const orc_params *p, const double *rho, double *c0, double *cm, double *sm);
void orc_phi_boundaries(const orc_params *p, double gconst, const double *c0, const double *cm,
                        const double *sm, double *phi);
void orc_density(const orc_params *p, const double *u, double *dens);
void orc_gravity_source_phi(const orc_params *p, const double *prim, const double *phi,
                            double *src);
double orc_advance_sg(const orc_params *p, double *u, double *up, double *upp, double *prim,
                      double *ff, double *gg, double *hh, double *src, orc_sg *sg, double *dens);

/* Cooling integrated to convergence (the reference formula sub-cycled). */
double orc_cool_subcycled(const orc_params *p, double n, double T, double dt, double eps,
                          int max_sub, int *nsub);
/* Cooling integrated exactly (closed form for the power-law table segments);
 * gate_frac T_min when the gas reaches the gate within dt. */
double orc_cool_exact_int(const orc_params *p, double n, double T, double dt, double gate_frac);
/* Cooling surrogate MLP: parameter count, forward, loss and gradient. */
int orc_mlp_nparams(void);
double orc_mlp(const float *theta, const double *x, const double *y, int B, double *out,
               double *grad);
/* 1 (default): bf16 roundings as the GPU kernels; 0: plain FP64 network. */
void orc_mlp_set_quant(int q);

/* OpenMP thread count for the oracle loops (1 = serial). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
