/*
 * sanitize_main.c -- drives every routine of the C oracle once on small
 * problems, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (make sanitize).  TEST INFRASTRUCTURE ONLY: run by the CPU test suite; it
 * exits non-zero when a result is not finite or a sanitizer reports.
 *
 * Usage: sanitize_oracle <cooling table: first line n, then rows T -log10(Lambda)>
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tlq_oracle.h"

#define PC 3.085677588e18
#define AMU 1.660538782e-24
#define KB 1.380650400e-16
#define MSUN 1.988920000e33
#define YEAR 3.155673600e7

static unsigned long long g_rng = 88172645463325252ull;
static double urand(void) {  /* xorshift64 in [0, 1) */
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return (double)(g_rng >> 11) * (1.0 / 9007199254740992.0);
}

static int g_bad = 0;
static void check_finite(const char *what, const double *a, size_t n) {
    for (size_t q = 0; q < n; ++q)
        if (!isfinite(a[q])) {
            fprintf(stderr, "%s: non-finite value at %zu\n", what, q);
            ++g_bad;
            return;
        }
}

/* Interior and ghost cells of a random cgs state: n within a factor
 * 10^(spread/2) of 1 cm^-3, T within 10^(2 spread) of 1e5 K, |v| < 20 km/s. */
static void random_state(const orc_params *p, double *u, double spread) {
    const size_t ncell = orc_array_len(p) / 5;
    for (size_t c = 0; c < ncell; ++c) {
        const double nd = pow(10.0, spread * (urand() - 0.5)),
                     T = pow(10.0, 5.0 + 4.0 * spread * (urand() - 0.5));
        const double rho = nd * p->mu0 * AMU, pr = nd * KB * T;
        double v[3];
        for (int d = 0; d < 3; ++d) v[d] = 2e6 * (2.0 * urand() - 1.0);
        u[5 * c + 0] = rho;
        for (int d = 0; d < 3; ++d) u[5 * c + 1 + d] = rho * v[d];
        u[5 * c + 4] = 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + pr / (p->gamma - 1.0);
    }
}

typedef struct {
    double *u, *up, *upp, *prim, *ff, *gg, *hh, *src;
    size_t n;
} state;

static state alloc_state(const orc_params *p) {
    state s;
    s.n = orc_array_len(p);
    double **f[] = {&s.u, &s.up, &s.upp, &s.prim, &s.ff, &s.gg, &s.hh, &s.src};
    for (size_t q = 0; q < sizeof f / sizeof f[0]; ++q) *f[q] = calloc(s.n, sizeof(double));
    return s;
}

static void free_state(state *s) {
    free(s->u); free(s->up); free(s->upp); free(s->prim);
    free(s->ff); free(s->gg); free(s->hh); free(s->src);
}

static void start(const orc_params *p, state *s) {  /* mainHD.f90:32-44 */
    orc_boundaries1(p, s->u);
    orc_winds(p, s->u);
    orc_calc_prims(p, s->prim, s->u);
}

static int read_table(const char *path, double **T, double **L) {
    FILE *f = fopen(path, "r");
    int n = 0;
    if (!f || fscanf(f, "%d", &n) != 1 || n < 2) {
        if (f) fclose(f);
        return 0;
    }
    char rest[256];
    if (!fgets(rest, sizeof rest, f)) n = 0;  /* remainder of the first line */
    *T = malloc(n * sizeof(double));
    *L = malloc(n * sizeof(double));
    for (int q = 0; q < n; ++q) {
        double a, b;
        if (fscanf(f, "%lf %lf", &a, &b) != 2) {
            n = 0;
            break;
        }
        (*T)[q] = a;
        (*L)[q] = -b;
    }
    fclose(f);
    return n;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <cooling table>\n", argv[0]);
        return 2;
    }
    double *tT = NULL, *tL = NULL;
    const int nt = read_table(argv[1], &tT, &tL);
    if (!nt) {
        fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    orc_set_threads(2);

    /* 1. A cgs problem with every source: mixed boundaries, two winds,
     *    cooling, a softened point mass; three full steps. */
    const double dx = 0.1 * PC;
    orc_wind winds[2] = {
        {3.0 * dx, 4.0 * dx, 4.3 * dx, 2e5, 0.0, -1e5, 2.5 * dx, 1e-5 * MSUN / YEAR, 3e7, 1e4, 1.0},
        {9.5 * dx, 2.0 * dx, 7.0 * dx, 0.0, 0.0, 0.0, 1.5 * dx, 1e-6 * MSUN / YEAR, 1e8, 5e3, 1.3}};
    const double mass[4] = {6.25 * dx, 5.0 * dx, 4.5 * dx, 1e4 * MSUN};
    orc_params p;
    memset(&p, 0, sizeof p);
    p.nx = 12; p.ny = 10; p.nz = 9;
    p.xl = 0.0; p.xr = 12 * dx; p.yl = 0.0; p.yr = 10 * dx; p.zl = 0.0; p.zr = 9 * dx;
    p.gamma = 5.0 / 3.0; p.mu0 = 1.0; p.cfl = 0.4; p.eta = 0.05; p.cooling_limit = 0.5;
    p.limiter = ORC_LIM_VANLEER;
    const int bc[6] = {ORC_BC_REFLECTIVE, ORC_BC_OUTFLOW, ORC_BC_PERIODIC, ORC_BC_PERIODIC,
                       ORC_BC_OUTFLOW, ORC_BC_REFLECTIVE};
    memcpy(p.bc, bc, sizeof bc);
    p.solver = ORC_SOLVER_HLLC;
    p.cooling = 1;
    p.n_winds = 2; p.winds = winds;
    p.n_masses = 1; p.masses = mass; p.softening = 2.0 * dx;
    p.n_table = nt; p.table_T = tT; p.table_L = tL;
    state s = alloc_state(&p);
    random_state(&p, s.u, 1.0);
    start(&p, &s);
    for (int q = 0; q < 3; ++q) orc_advance(&p, s.u, s.up, s.upp, s.prim, s.ff, s.gg, s.hh, s.src);
    check_finite("all sources", s.u, s.n);

    /* 2. Every limiter with HLLC, and HLL, one step each; both flux choices
     *    and the face solvers on their own. */
    p.cooling = 0; p.n_winds = 0; p.n_masses = 0;
    for (int solver = 0; solver <= 1; ++solver)
        for (int lim = 1; lim <= 6; ++lim) {
            p.solver = solver;
            p.limiter = lim;
            /* unlimited reconstruction of white noise can overshoot to a
             * negative density: a gentler state for the limiter sweep */
            random_state(&p, s.u, 0.1);
            start(&p, &s);
            orc_advance(&p, s.u, s.up, s.upp, s.prim, s.ff, s.gg, s.hh, NULL);
            check_finite("limiters", s.u, s.n);
        }
    orc_fluxes(&p, s.u, s.prim, s.ff, s.gg, s.hh, 1);
    orc_fluxes(&p, s.u, s.prim, s.ff, s.gg, s.hh, 2);
    orc_godunov(&p, s.up, s.u, s.ff, s.gg, s.hh, NULL, 1e3);
    orc_step(&p, s.u, s.up, s.upp, 1e3);
    check_finite("routines", s.u, s.n);
    {
        const double wl[5] = {1.0, 0.3, -0.2, 0.1, 1.0}, wr[5] = {0.125, -0.1, 0.0, 0.2, 0.1};
        double ul[5], ur[5], f[5], a[7], b[7], out[7];
        for (int e = 0; e < 5; ++e) { ul[e] = wl[e]; ur[e] = wr[e]; }
        for (int d = 0; d < 3; ++d) {
            orc_hllc_face(&p, wl, wr, ul, ur, d, f);
            check_finite("hllc face", f, 5);
            orc_hll_face(&p, wl, wr, ul, ur, d, f);
            check_finite("hll face", f, 5);
        }
        for (int q = 0; q < 7; ++q) { a[q] = q - 3.0; b[q] = 2.0 - q; }
        for (int lim = 1; lim <= 6; ++lim) {
            orc_limiter_avg(lim, a, b, out, 7);
            check_finite("limiter avg", out, 7);
        }
    }
    free_state(&s);

    /* 3. Self-gravity: multigrid with multipole boundaries, two steps. */
    orc_params g;
    memset(&g, 0, sizeof g);
    g.nx = g.ny = g.nz = 16;
    g.xl = g.yl = g.zl = -PC; g.xr = g.yr = g.zr = PC;
    g.gamma = 5.0 / 3.0; g.mu0 = 1.0; g.cfl = 0.4; g.eta = 0.005; g.cooling_limit = 0.5;
    g.limiter = ORC_LIM_VANLEER;
    for (int q = 0; q < 6; ++q) g.bc[q] = ORC_BC_OUTFLOW;
    state t = alloc_state(&g);
    random_state(&g, t.u, 1.0);
    start(&g, &t);
    const int levels = 3;
    const size_t nphi = (size_t)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    double *phi = calloc(nphi, sizeof(double)), *dens = calloc(nphi, sizeof(double));
    double *work = calloc(orc_sg_work_len(&g, levels), sizeof(double));
    orc_sg sg = {levels, 2, 6.67430e-8, phi, work, 0, 0, 0.0, 0.0};
    orc_density(&g, t.u, dens);
    orc_multigrid(&g, &sg, dens);
    for (int q = 0; q < 2; ++q) {
        sg.it = q;
        orc_advance_sg(&g, t.u, t.up, t.upp, t.prim, t.ff, t.gg, t.hh, t.src, &sg, dens);
    }
    check_finite("self-gravity state", t.u, t.n);
    check_finite("self-gravity phi", phi, nphi);
    double c0[5], cm[16], sm[16];
    orc_moments(&g, dens, c0, cm, sm);
    orc_phi_boundaries(&g, sg.gconst, c0, cm, sm, phi);
    orc_gravity_source_phi(&g, t.prim, phi, t.src);
    check_finite("multipole faces", phi, nphi);
    free(phi); free(dens); free(work);
    free_state(&t);

    /* 4. Cooling to convergence, its exact integration and the surrogate network
     * (forward, loss, gradient). */
    p.n_table = nt; p.table_T = tT; p.table_L = tL; p.cooling = 1;
    for (int q = 0; q < 20; ++q) {
        int ns = 0;
        double T1 = orc_cool_subcycled(&p, pow(10.0, 4.0 * urand() - 2.0),
                                       pow(10.0, 4.0 + 4.5 * urand()), 1e12 * urand(), 0.01,
                                       100000, &ns);
        check_finite("subcycled", &T1, 1);
        /* exact integration: below, inside and above the table, to the gate */
        double T2 = orc_cool_exact_int(&p, pow(10.0, 4.0 * urand() - 2.0),
                                       pow(10.0, 3.5 + 5.5 * urand()), 1e14 * urand(), 0.995);
        check_finite("exact_int", &T2, 1);
    }
    {
        double lam = orc_find_coef(&p, 3e5);
        check_finite("find_coef", &lam, 1);
    }
    const int np = orc_mlp_nparams();
    float *theta = malloc(np * sizeof(float));
    double *grad = malloc(np * sizeof(double));
    for (int q = 0; q < np; ++q) theta[q] = (float)(0.3 * (2.0 * urand() - 1.0));
    double x[3 * 8], y[8], out[8];
    for (int b = 0; b < 8; ++b) {
        x[3 * b] = 7.0 * urand() - 3.0;
        x[3 * b + 1] = 4.0 + 5.0 * urand();
        x[3 * b + 2] = 4.0 * urand() - 2.0;
        y[b] = -urand();
    }
    for (int q = 0; q <= 1; ++q) {
        orc_mlp_set_quant(q);
        double loss = orc_mlp(theta, x, y, 8, out, grad);
        check_finite("mlp out", out, 8);
        check_finite("mlp loss", &loss, 1);
        check_finite("mlp grad", grad, (size_t)np);
    }
    free(theta); free(grad); free(tT); free(tL);
    if (g_bad) return 1;
    printf("sanitize_oracle: every routine ran; all results finite\n");
    return 0;
}
