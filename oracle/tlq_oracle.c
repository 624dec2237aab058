/*
 * tlq_oracle.c -- FP64 CPU restatement of the tlaloque HydroCore time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline
 * timed by bench.py.  Nothing in the product (tlaloque_amd, libtlq.so) calls
 * it.  Each routine cites the reference routine it follows; arithmetic keeps
 * the reference's operand order so that a build of the reference with the
 * same compiler settings agrees to the last bits (libm aside).  Build with
 * -ffp-contract=off (no fused multiply-add), as in oracle/Makefile.
 *
 * Layout: see tlq_oracle.h (Fortran (e,i,j,k), e fastest, two ghost layers).
 */
#include "tlq_oracle.h"

#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants of source/constants.f90:6-15 (cgs). */
#define ORC_AMU 1.660538782e-24
#define ORC_KB 1.380650400e-16
#define ORC_GR 6.67430e-8
#define ORC_PI 3.14159265358979

/* Fortran (e, i, j, k) with i, j, k in -1..n+2 (offset 0 at index -1) -> flat
 * offset of component 0. */
static inline size_t cell(const orc_params *p, int i, int j, int k) {
    const size_t sx = (size_t)p->nx + 4, sy = (size_t)p->ny + 4;
    return 5 * ((size_t)(i + 1) + sx * (size_t)(j + 1) + sx * sy * (size_t)(k + 1));
}

size_t orc_array_len(const orc_params *p) {
    return (size_t)5 * (p->nx + 4) * (p->ny + 4) * (p->nz + 4);
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Cell-centre coordinates, source/HydroCore.f90:35-75 (single rank: coords = 0). */
static inline double xcoord(const orc_params *p, int i) {
    const double dx = (p->xr - p->xl) / p->nx;
    return (p->xl + dx * (double)i) - dx / 2.0;
}
static inline double ycoord(const orc_params *p, int j) {
    const double dy = (p->yr - p->yl) / p->ny;
    return (p->yl + dy * (double)j) - dy / 2.0;
}
static inline double zcoord(const orc_params *p, int k) {
    const double dz = (p->zr - p->zl) / p->nz;
    return (p->zl + dz * (double)k) - dz / 2.0;
}

/* calc_prims, source/HydroCore.f90:1-33: all cells, ghosts included. */
void orc_calc_prims(const orc_params *p, double *prim, const double *uu) {
    const double g1 = p->gamma - 1.0;
    const size_t ncell = orc_array_len(p) / 5;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < ncell; ++c) {
        const double *u = uu + 5 * c;
        double *w = prim + 5 * c;
        w[0] = u[0];
        w[1] = u[1] / u[0];
        w[2] = u[2] / u[0];
        w[3] = u[3] / u[0];
        double pr = g1 * (u[4] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0]);
        w[4] = pr > 1e-20 ? pr : 1e-20; /* HydroCore.f90:21 */
    }
}

/* csound, source/HydroCore.f90:157-168. */
static inline double csound(double gamma, double pres, double den) {
    return sqrt(gamma * pres / den);
}

/* time_step, source/HydroCore.f90:113-155 (serial branch, :148). */
double orc_time_step(const orc_params *p, const double *prim) {
    double mx = 0.0, my = 0.0, mz = 0.0;
#pragma omp parallel for reduction(max : mx, my, mz) schedule(static)
    for (int k = 1; k <= p->nz; ++k)
        for (int j = 1; j <= p->ny; ++j)
            for (int i = 1; i <= p->nx; ++i) {
                const double *w = prim + cell(p, i, j, k);
                const double cs = csound(p->gamma, w[4], w[0]);
                const double ax = fabs(w[1]) + cs, ay = fabs(w[2]) + cs, az = fabs(w[3]) + cs;
                if (ax > mx) mx = ax;
                if (ay > my) my = ay;
                if (az > mz) mz = az;
            }
    const double dx = (p->xr - p->xl) / p->nx, dy = (p->yr - p->yl) / p->ny,
                 dz = (p->zr - p->zl) / p->nz;
    double m = dx / mx;
    if (dy / my < m) m = dy / my;
    if (dz / mz < m) m = dz / mz;
    return p->cfl * m / sqrt(3.0); /* HydroCore.f90:144 */
}

/* prim2f, source/HydroCore.f90:213-251.  dim = 0, 1, 2 for x, y, z. */
static inline void prim2f(double gamma, const double *w, int dim, double *f) {
    const int n = dim + 1;
    const double vn = w[n];
    f[0] = w[0] * vn;
    for (int c = 1; c <= 3; ++c) {
        const int a = c < n ? c : n, b = c < n ? n : c;
        f[c] = w[0] * w[a] * w[b];
        if (c == n) f[c] = f[c] + w[4];
    }
    f[4] = vn * (0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) +
                 (gamma / (gamma - 1.0)) * w[4]);
}

/* average(), source/HLLC.f90:586-636 (identical in HLL.f90:345-395). */
static inline double lim_avg(int opt, double a, double b) {
    double s, c, d, r;
    switch (opt) {
    case ORC_LIM_NONE:
        return 0.5 * (a + b);
    case ORC_LIM_VANLEER:
        if (a * b <= 0.0) return 0.0;
        return a * b * (a + b) / (a * a + b * b);
    case ORC_LIM_MINMOD:
        s = copysign(1.0, a);
        r = fabs(a) < s * b ? fabs(a) : s * b;
        return s * (r > 0.0 ? r : 0.0);
    case ORC_LIM_UMIST:
        s = copysign(1.0, a);
        c = 0.25 * a + 0.75 * b;
        d = 0.75 * a + 0.25 * b;
        r = 2.0 * fabs(a);
        if (2.0 * s * b < r) r = 2.0 * s * b;
        if (s * c < r) r = s * c;
        if (s * d < r) r = s * d;
        return s * (r > 0.0 ? r : 0.0);
    case ORC_LIM_WOODWARD:
        s = copysign(1.0, a);
        c = 0.5 * (a + b);
        r = 2.0 * fabs(a);
        if (2.0 * s * b < r) r = 2.0 * s * b;
        if (s * c < r) r = s * c;
        return s * (r > 0.0 ? r : 0.0);
    case ORC_LIM_SUPERBEE:
        s = copysign(1.0, b);
        c = 2.0 * fabs(b) < s * a ? 2.0 * fabs(b) : s * a;
        d = fabs(b) < 2.0 * s * a ? fabs(b) : 2.0 * s * a;
        r = 0.0;
        if (c > r) r = c;
        if (d > r) r = d;
        return s * r;
    default:
        return NAN;
    }
}

void orc_limiter_avg(int lim, const double *a, const double *b, double *out, int n) {
    for (int t = 0; t < n; ++t) out[t] = lim_avg(lim, a[t], b[t]);
}

/* limiter, source/HLLC.f90:560-582: reconstructs pl (in place) from
 * (pll, pl, pr) and pr from (pl, pr, prr). */
static inline void limiter(int lim, const double *pll, double *pl, double *pr, const double *prr) {
    for (int e = 0; e < 5; ++e) {
        const double dl = pl[e] - pll[e], dm = pr[e] - pl[e], dr = prr[e] - pr[e];
        const double al = lim_avg(lim, dl, dm), ar = lim_avg(lim, dm, dr);
        pl[e] = pl[e] + 0.5 * al;
        pr[e] = pr[e] - 0.5 * ar;
    }
}

/* wavespeed, source/HLLC.f90:498-518 (Davis estimates + contact speed). */
static inline void wavespeed(double gamma, const double *wl, const double *wr, int dim, double *sl,
                             double *sr, double *sst) {
    const int n = dim + 1;
    const double csl = csound(gamma, wl[4], wl[0]), csr = csound(gamma, wr[4], wr[0]);
    const double a = wl[n] - csl, b = wr[n] - csr;
    const double c = wl[n] + csl, d = wr[n] + csr;
    *sl = a < b ? a : b;
    *sr = c > d ? c : d;
    *sst = (wr[4] - wl[4] + wl[0] * wl[n] * (*sl - wl[n]) - wr[0] * wr[n] * (*sr - wr[n])) /
           (wl[0] * (*sl - wl[n]) - wr[0] * (*sr - wr[n]));
}

/* One HLLC face: source/HLLC.f90:99-135 (choice 1) and 275-314 (choice 2).
 * pl/pr are the (reconstructed) primitive states, ul/ur the cell-centred
 * conserved states of the two cells that share the face. */
void orc_hllc_face(const orc_params *p, const double *pl, const double *pr, const double *ul,
                   const double *ur, int dim, double *f) {
    const int n = dim + 1;
    double sl, sr, sst, fs[5], us[5];
    wavespeed(p->gamma, pl, pr, dim, &sl, &sr, &sst);
    if (sl >= 0) {
        prim2f(p->gamma, pl, dim, f);
    } else if (sr <= 0) {
        prim2f(p->gamma, pr, dim, f);
    } else if (sst >= 0) {
        prim2f(p->gamma, pl, dim, fs);
        const double rs = pl[0] * (sl - pl[n]) / (sl - sst);
        us[0] = rs;
        for (int c = 1; c <= 3; ++c) us[c] = c == n ? rs * sst : rs * pl[c];
        us[4] = rs * (ul[4] / pl[0] + (sst - pl[n]) * (sst + pl[4] / (pl[0] * (sl - pl[n]))));
        for (int e = 0; e < 5; ++e) f[e] = fs[e] + sl * (us[e] - ul[e]);
    } else if (sst <= 0) {
        prim2f(p->gamma, pr, dim, fs);
        const double rs = pr[0] * (sr - pr[n]) / (sr - sst);
        us[0] = rs;
        for (int c = 1; c <= 3; ++c) us[c] = c == n ? rs * sst : rs * pr[c];
        us[4] = rs * (ur[4] / pr[0] + (sst - pr[n]) * (sst + pr[4] / (pr[0] * (sr - pr[n]))));
        for (int e = 0; e < 5; ++e) f[e] = fs[e] + sr * (us[e] - ur[e]);
    }
    /* A NaN speed matches no branch and leaves f unassigned, as in the reference. */
}

/* One HLL face: source/HLL.f90:106-128 and 197-223 (wavespeed 299-317). */
void orc_hll_face(const orc_params *p, const double *pl, const double *pr, const double *ul,
                  const double *ur, int dim, double *f) {
    double sl, sr, sst, fl[5], fr[5];
    wavespeed(p->gamma, pl, pr, dim, &sl, &sr, &sst);
    (void)sst;
    if (sl >= 0) {
        prim2f(p->gamma, pl, dim, f);
    } else if (sr <= 0) {
        prim2f(p->gamma, pr, dim, f);
    } else {
        prim2f(p->gamma, pl, dim, fl);
        prim2f(p->gamma, pr, dim, fr);
        for (int e = 0; e < 5; ++e)
            f[e] = (sr * fl[e] - sl * fr[e] + sl * sr * (ur[e] - ul[e])) / (sr - sl);
    }
}

/* HLLCFluxes / HLLFluxes, source/HLLC.f90:64-496, source/HLL.f90:71-297.
 * Faces i = 0..nx, j = 0..ny, k = 0..nz in all three directions (HLLC.f90:94-96). */
void orc_fluxes(const orc_params *p, const double *uu, const double *prim, double *ff,
                double *gg, double *hh, int choice) {
    double *out[3] = {ff, gg, hh};
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= p->nz; ++k)
        for (int j = 0; j <= p->ny; ++j)
            for (int i = 0; i <= p->nx; ++i) {
                for (int dim = 0; dim < 3; ++dim) {
                    const int di = dim == 0, dj = dim == 1, dk = dim == 2;
                    const size_t c0 = cell(p, i, j, k), c1 = cell(p, i + di, j + dj, k + dk);
                    double wl[5], wr[5];
                    memcpy(wl, prim + c0, sizeof wl);
                    memcpy(wr, prim + c1, sizeof wr);
                    if (choice == 2) {
                        const double *wll = prim + cell(p, i - di, j - dj, k - dk);
                        const double *wrr = prim + cell(p, i + 2 * di, j + 2 * dj, k + 2 * dk);
                        limiter(p->limiter, wll, wl, wr, wrr);
                    }
                    double *f = out[dim] + c0;
                    if (p->solver == ORC_SOLVER_HLL)
                        orc_hll_face(p, wl, wr, uu + c0, uu + c1, dim, f);
                    else
                        orc_hllc_face(p, wl, wr, uu + c0, uu + c1, dim, f);
                }
            }
}

/* Godunov, source/HLLC.f90:1-62: interior update; src (if non-NULL) is the
 * S_Solver source added as dt*S (HLLC.f90:39-41). */
void orc_godunov(const orc_params *p, double *uout, const double *uu, const double *ff,
                 const double *gg, const double *hh, const double *src, double dt) {
    const double dx = (p->xr - p->xl) / p->nx, dy = (p->yr - p->yl) / p->ny,
                 dz = (p->zr - p->zl) / p->nz;
    const double dtdx = dt / dx, dtdy = dt / dy, dtdz = dt / dz;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= p->nz; ++k)
        for (int j = 1; j <= p->ny; ++j)
            for (int i = 1; i <= p->nx; ++i) {
                const size_t c = cell(p, i, j, k);
                const size_t cx = cell(p, i - 1, j, k), cy = cell(p, i, j - 1, k),
                             cz = cell(p, i, j, k - 1);
                for (int e = 0; e < 5; ++e) {
                    uout[c + e] = uu[c + e] - dtdx * (ff[c + e] - ff[cx + e]) -
                                  dtdy * (gg[c + e] - gg[cy + e]) - dtdz * (hh[c + e] - hh[cz + e]);
                    if (src) uout[c + e] = uout[c + e] + dt * src[c + e];
                }
            }
}

/* Copy one full ghost plane: dst layer <- src layer along axis, with the
 * normal momentum component (1 + axis) negated when sign < 0. */
static void copy_plane(const orc_params *p, double *u, int axis, int dst, int src, int negate) {
    const int lo = -1;
    const int hi0 = p->nx + 2, hi1 = p->ny + 2, hi2 = p->nz + 2;
    if (axis == 0) {
        for (int k = lo; k <= hi2; ++k)
            for (int j = lo; j <= hi1; ++j) {
                double *d = u + cell(p, dst, j, k);
                const double *s = u + cell(p, src, j, k);
                for (int e = 0; e < 5; ++e) d[e] = (negate && e == 1) ? -s[e] : s[e];
            }
    } else if (axis == 1) {
        for (int k = lo; k <= hi2; ++k)
            for (int i = lo; i <= hi0; ++i) {
                double *d = u + cell(p, i, dst, k);
                const double *s = u + cell(p, i, src, k);
                for (int e = 0; e < 5; ++e) d[e] = (negate && e == 2) ? -s[e] : s[e];
            }
    } else {
        for (int j = lo; j <= hi1; ++j)
            for (int i = lo; i <= hi0; ++i) {
                double *d = u + cell(p, i, j, dst);
                const double *s = u + cell(p, i, j, src);
                for (int e = 0; e < 5; ++e) d[e] = (negate && e == 3) ? -s[e] : s[e];
            }
    }
}

/* boundariesI serial branch, source/boundaries.f90:179-259: one ghost layer,
 * order left, right (x), top, down (y), back, front (z). */
void orc_boundaries1(const orc_params *p, double *u) {
    const int n[3] = {p->nx, p->ny, p->nz};
    /* (axis, side) in the reference's order; side 0 = low face, 1 = high face. */
    const int order[6][2] = {{0, 0}, {0, 1}, {1, 1}, {1, 0}, {2, 1}, {2, 0}};
    for (int s = 0; s < 6; ++s) {
        const int axis = order[s][0], side = order[s][1];
        const int bc = p->bc[2 * axis + side];
        const int g = side ? n[axis] + 1 : 0;          /* ghost layer */
        const int in = side ? n[axis] : 1;             /* adjacent interior layer */
        const int wrap = side ? 1 : n[axis];           /* periodic partner */
        if (bc == ORC_BC_OUTFLOW) copy_plane(p, u, axis, g, in, 0);
        else if (bc == ORC_BC_REFLECTIVE) copy_plane(p, u, axis, g, in, 1);
        else if (bc == ORC_BC_PERIODIC) copy_plane(p, u, axis, g, wrap, 0);
        /* BC_INFLOW is a no-op in the reference (boundaries.f90:191-193). */
    }
}

/* boundariesII serial branch, source/boundaries.f90:432-533: two ghost
 * layers, order left, right, down, top, front, back. */
void orc_boundaries2(const orc_params *p, double *u) {
    const int n[3] = {p->nx, p->ny, p->nz};
    for (int axis = 0; axis < 3; ++axis)
        for (int side = 0; side < 2; ++side) {
            const int bc = p->bc[2 * axis + side];
            const int N = n[axis];
            if (side == 0) {
                if (bc == ORC_BC_PERIODIC) {
                    copy_plane(p, u, axis, -1, N - 1, 0);
                    copy_plane(p, u, axis, 0, N, 0);
                } else if (bc == ORC_BC_OUTFLOW) {
                    copy_plane(p, u, axis, 0, 1, 0);
                    copy_plane(p, u, axis, -1, 1, 0);
                } else if (bc == ORC_BC_REFLECTIVE) {
                    copy_plane(p, u, axis, 0, 1, 1);
                    copy_plane(p, u, axis, -1, 2, 1);
                }
            } else {
                if (bc == ORC_BC_PERIODIC) {
                    copy_plane(p, u, axis, N + 1, 1, 0);
                    copy_plane(p, u, axis, N + 2, 2, 0);
                } else if (bc == ORC_BC_OUTFLOW) {
                    copy_plane(p, u, axis, N + 1, N, 0);
                    copy_plane(p, u, axis, N + 2, N, 0);
                } else if (bc == ORC_BC_REFLECTIVE) {
                    copy_plane(p, u, axis, N + 1, N, 1);
                    copy_plane(p, u, axis, N + 2, N - 1, 1);
                }
            }
        }
}

/* imposeSphericalWind for every wind, source/winds.f90:32-90 called from
 * user_boundaries, user/user.f90:57-70: all cells, ghosts included. */
void orc_winds(const orc_params *p, double *u) {
    for (int w = 0; w < p->n_winds; ++w) {
        const orc_wind *wd = &p->winds[w];
#pragma omp parallel for schedule(static)
        for (int k = -1; k <= p->nz + 2; ++k)
            for (int j = -1; j <= p->ny + 2; ++j)
                for (int i = -1; i <= p->nx + 2; ++i) {
                    const double x = xcoord(p, i), y = ycoord(p, j), z = zcoord(p, k);
                    const double dist = sqrt((x - wd->xc) * (x - wd->xc) + (y - wd->yc) * (y - wd->yc) +
                                             (z - wd->zc) * (z - wd->zc));
                    if (dist < wd->radius) {
                        const double dens = wd->mdot / (wd->vinf * dist * dist * 4.0 * ORC_PI);
                        const double vx = wd->vinf * (x - wd->xc) / dist + wd->vx;
                        const double vy = wd->vinf * (y - wd->yc) / dist + wd->vy;
                        const double vz = wd->vinf * (z - wd->zc) / dist + wd->vz;
                        const double pres = dens * ORC_KB * wd->temp / (wd->mu * ORC_AMU);
                        double *c = u + cell(p, i, j, k);
                        c[0] = dens;
                        c[1] = dens * vx;
                        c[2] = dens * vy;
                        c[3] = dens * vz;
                        c[4] = 0.5 * dens * (vx * vx + vy * vy + vz * vz) + pres / (p->gamma - 1.0);
                    }
                }
    }
}

/* step, source/HydroCore.f90:170-211: UPP = U; interior U = UP + eta*Lap(UPP);
 * every ghost cell of U takes UP's value.  Time/iteration bookkeeping
 * (:208-209) is the caller's. */
void orc_step(const orc_params *p, double *u, double *up, double *upp, double dt) {
    (void)dt;
    const size_t len = orc_array_len(p);
    memcpy(upp, u, len * sizeof(double));
    const double eta = p->eta;
#pragma omp parallel for schedule(static)
    for (int k = -1; k <= p->nz + 2; ++k)
        for (int j = -1; j <= p->ny + 2; ++j)
            for (int i = -1; i <= p->nx + 2; ++i) {
                const size_t c = cell(p, i, j, k);
                const int interior = i >= 1 && i <= p->nx && j >= 1 && j <= p->ny && k >= 1 &&
                                     k <= p->nz;
                if (!interior) {
                    for (int e = 0; e < 5; ++e) u[c + e] = up[c + e];
                    continue;
                }
                const size_t xp = cell(p, i + 1, j, k), xm = cell(p, i - 1, j, k);
                const size_t yp = cell(p, i, j + 1, k), ym = cell(p, i, j - 1, k);
                const size_t zp = cell(p, i, j, k + 1), zm = cell(p, i, j, k - 1);
                for (int e = 0; e < 5; ++e)
                    u[c + e] = up[c + e] + eta * (upp[xp + e] + upp[xm + e] + upp[yp + e] +
                                                  upp[ym + e] + upp[zp + e] + upp[zm + e] -
                                                  6.0 * upp[c + e]);
            }
}

/* find_coef, source/cooling.f90:121-144: first row with T_i > temp, linear
 * interpolation of log10 Lambda in log10 T.  The reference leaves the result
 * unassigned for temp >= T_max (the loop falls through); here that case
 * returns the last row's Lambda. */
double orc_find_coef(const orc_params *p, double temp) {
    for (int i = 1; i < p->n_table; ++i) {
        if (p->table_T[i] > temp) {
            const double T0 = log10(p->table_T[i - 1]), C0 = p->table_L[i - 1];
            const double T1 = log10(p->table_T[i]), C1 = p->table_L[i];
            return pow(10.0, (C1 - C0) / (T1 - T0) * (log10(temp) - T0) + C0);
        }
    }
    return pow(10.0, p->table_L[p->n_table - 1]);
}

/* apply_cooling, source/cooling.f90:47-119 with calcTemp (:146-159). */
void orc_cooling(const orc_params *p, double *prim, double *u, double dt) {
    const double Tmin = p->table_T[0]; /* cooling.f90:42 */
    const double g1 = p->gamma - 1.0;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= p->nz; ++k)
        for (int j = 1; j <= p->ny; ++j)
            for (int i = 1; i <= p->nx; ++i) {
                const size_t c = cell(p, i, j, k);
                double *w = prim + c;
                const double temp = w[4] * p->mu0 * ORC_AMU / (w[0] * ORC_KB);
                if (!(temp >= Tmin)) continue;
                double eth = w[4] / g1;
                const double vel2 = w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
                const double ek = 0.5 * w[0] * vel2;
                const double aloss = orc_find_coef(p, temp);
                const double numdens = w[0] / (p->mu0 * ORC_AMU);
                const double ce = aloss * (numdens * numdens) / eth;
                double f = exp(-ce * dt);
                if (f < 1.0 - p->cooling_limit) f = 1.0 - p->cooling_limit;
                const double new_temp = temp * f;
                if (new_temp < 10.0) f = 10.0 / temp;
                w[4] = w[4] * f;
                eth = w[4] / g1;
                u[c + 4] = ek + eth;
            }
}

/* Point-mass potential at a cell centre: pointsource_gravity.f90:12-39
 * (Phi -= G*M/r per mass, summed in order) with optional Plummer softening. */
static inline double potential(const orc_params *p, int i, int j, int k) {
    const double x = xcoord(p, i), y = ycoord(p, j), z = zcoord(p, k);
    const double eps2 = p->softening * p->softening;
    double phi = 0.0;
    for (int s = 0; s < p->n_masses; ++s) {
        const double *m = p->masses + 4 * s;
        const double r = sqrt((m[0] - x) * (m[0] - x) + (m[1] - y) * (m[1] - y) +
                              (m[2] - z) * (m[2] - z) + eps2);
        phi = phi - ORC_GR * m[3] / r;
    }
    return phi;
}

/* PHI(0:nx+1, 0:ny+1, 0:nz+1) of source/globals.f90:19 (i fastest). */
static inline size_t phc(const orc_params *p, int i, int j, int k) {
    return (size_t)i + (size_t)(p->nx + 2) * ((size_t)j + (size_t)(p->ny + 2) * (size_t)k);
}

/* PHIT = PHIP + PHI (HLLC.f90:22) at a cell: the point masses' potential
 * plus, when phi is given, the self-gravity potential. */
static inline double phit(const orc_params *p, const double *phi, int i, int j, int k) {
    const double pp = potential(p, i, j, k);
    return phi ? pp + phi[phc(p, i, j, k)] : pp;
}

/* S_Solver, source/sources.f90:872-897, with Phi = PHIT. */
void orc_gravity_source_phi(const orc_params *p, const double *prim, const double *phi,
                            double *src) {
    const double dx = (p->xr - p->xl) / p->nx, dy = (p->yr - p->yl) / p->ny,
                 dz = (p->zr - p->zl) / p->nz;
    memset(src, 0, orc_array_len(p) * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= p->nz; ++k)
        for (int j = 1; j <= p->ny; ++j)
            for (int i = 1; i <= p->nx; ++i) {
                const double g0 = (phit(p, phi, i + 1, j, k) - phit(p, phi, i - 1, j, k)) / (2.0 * dx);
                const double g1 = (phit(p, phi, i, j + 1, k) - phit(p, phi, i, j - 1, k)) / (2.0 * dy);
                const double g2 = (phit(p, phi, i, j, k + 1) - phit(p, phi, i, j, k - 1)) / (2.0 * dz);
                const size_t c = cell(p, i, j, k);
                const double *w = prim + c;
                double *s = src + c;
                s[0] = 0.0;
                s[1] = -w[0] * g0;
                s[2] = -w[0] * g1;
                s[3] = -w[0] * g2;
                s[4] = -(w[1] * w[0] * g0) - w[2] * w[0] * g1 - w[3] * w[0] * g2;
            }
}

/* S_Solver with Phi the point-mass potential only. */
void orc_gravity_source(const orc_params *p, const double *prim, double *src) {
    orc_gravity_source_phi(p, prim, NULL, src);
}

/* One iteration of the main loop, source/mainHD.f90:74-99. */
double orc_advance(const orc_params *p, double *u, double *up, double *upp, double *prim,
                   double *ff, double *gg, double *hh, double *src) {
    const int grav = p->n_masses > 0 && src != NULL;
    const double dt = orc_time_step(p, prim);            /* :74 */
    orc_fluxes(p, u, prim, ff, gg, hh, 1);               /* :76 */
    const double dth = dt / 2.0;                         /* :78 */
    if (grav) orc_gravity_source(p, prim, src);
    orc_godunov(p, upp, u, ff, gg, hh, grav ? src : NULL, dth); /* :80 */
    orc_boundaries2(p, upp);                             /* :82 */
    orc_winds(p, upp);
    orc_calc_prims(p, prim, upp);                        /* :84 */
    orc_fluxes(p, upp, prim, ff, gg, hh, 2);             /* :86 */
    if (grav) orc_gravity_source(p, prim, src);
    orc_godunov(p, up, u, ff, gg, hh, grav ? src : NULL, dt); /* :88 */
    orc_boundaries1(p, up);                              /* :90 */
    orc_winds(p, up);
    orc_step(p, u, up, upp, dt);                         /* :92 */
    orc_calc_prims(p, prim, u);                          /* :94 */
    if (p->cooling) orc_cooling(p, prim, u, dt);         /* :96-98 */
    return dt;
}

/* ------------------------------------------------------------------------
 * Self-gravity: MultiGrid and its parts, source/sources.f90:99-676, and the
 * step with the potential solved inside Godunov (HLLC.f90:15-24).
 *
 * Level l holds Error(l), Residue(l), ret(l) (globals.f90:22-28), each
 * (nxl+2)(nyl+2)(nzl+2) with nxl = nx / 2^l (init.f90:94-116), i fastest.
 * Loops run in the reference's order (i outermost, k innermost), so the
 * Gauss-Seidel sweeps see the same mix of new and old values. */

typedef struct { int nx, ny, nz; double *e, *r, *t; } mg_level;

static inline size_t lidx(const mg_level *L, int i, int j, int k) {
    return (size_t)i + (size_t)(L->nx + 2) * ((size_t)j + (size_t)(L->ny + 2) * (size_t)k);
}
static inline size_t lsize(int nx, int ny, int nz) {
    return (size_t)(nx + 2) * (ny + 2) * (nz + 2);
}

size_t orc_sg_work_len(const orc_params *p, int levels) {
    size_t n = 0;
    for (int l = 0; l <= levels; ++l) n += 3 * lsize(p->nx >> l, p->ny >> l, p->nz >> l);
    return n;
}

static void mg_levels(const orc_params *p, int levels, double *work, mg_level *L) {
    for (int l = 0; l <= levels; ++l) {
        L[l].nx = p->nx >> l;
        L[l].ny = p->ny >> l;
        L[l].nz = p->nz >> l;
        const size_t n = lsize(L[l].nx, L[l].ny, L[l].nz);
        L[l].e = work;
        L[l].r = work + n;
        L[l].t = work + 2 * n;
        work += 3 * n;
    }
}

/* restriction, sources.f90:207-228: 8-cell average onto the coarse level. */
static void mg_restriction(const mg_level *F, const double *var, const mg_level *C, double *vt) {
    for (int i = 1; i <= C->nx; ++i)
        for (int j = 1; j <= C->ny; ++j)
            for (int k = 1; k <= C->nz; ++k) {
                const int a = 2 * i - 1, b = 2 * j - 1, c = 2 * k - 1;
#define V(ii, jj, kk) var[lidx(F, ii, jj, kk)]
                const double s = V(a, b, c) + V(a + 1, b, c) + V(a, b + 1, c) + V(a, b, c + 1) +
                                 V(a + 1, b + 1, c) + V(a + 1, b, c + 1) + V(a, b + 1, c + 1) +
                                 V(a + 1, b + 1, c + 1);
#undef V
                vt[lidx(C, i, j, k)] = (1.0 / 8.0) * s;
            }
}

/* prolongation, sources.f90:230-268: trilinear 27/9/3/1 weights from the
 * coarse level (ghosts included) onto the 8 children. */
static void mg_prolongation(const mg_level *C, const double *var, const mg_level *F, double *vt) {
    for (int i = 1; i <= C->nx; ++i)
        for (int j = 1; j <= C->ny; ++j)
            for (int k = 1; k <= C->nz; ++k) {
                const int a = 2 * i - 1, b = 2 * j - 1, c = 2 * k - 1;
#define V(ii, jj, kk) var[lidx(C, ii, jj, kk)]
                for (int dc = 0; dc < 2; ++dc)
                    for (int db = 0; db < 2; ++db)
                        for (int da = 0; da < 2; ++da) {
                            const int si = da ? 1 : -1, sj = db ? 1 : -1, sk = dc ? 1 : -1;
                            const double s = 27.0 * V(i, j, k) + 9.0 * V(i + si, j, k) +
                                             9.0 * V(i, j + sj, k) + 9.0 * V(i, j, k + sk) +
                                             3.0 * V(i + si, j + sj, k) + 3.0 * V(i + si, j, k + sk) +
                                             3.0 * V(i, j + sj, k + sk) + V(i + si, j + sj, k + sk);
                            vt[lidx(F, a + da, b + db, c + dc)] = (1.0 / 64.0) * s;
                        }
#undef V
            }
}

/* resid, sources.f90:270-300: resi = source - Laplacian(var) (ghosts: the
 * Laplacian is 0 there, so resi = source). */
static void mg_resid(const orc_params *p, const mg_level *L, const double *var, const double *src,
                     double *resi) {
    const double dxl = (p->xr - p->xl) / (double)L->nx;
    const size_t n = lsize(L->nx, L->ny, L->nz);
    for (size_t q = 0; q < n; ++q) resi[q] = src[q];
    for (int i = 1; i <= L->nx; ++i)
        for (int j = 1; j <= L->ny; ++j)
            for (int k = 1; k <= L->nz; ++k) {
                const size_t c = lidx(L, i, j, k);
                const size_t sx = 1, sy = (size_t)(L->nx + 2), sz = sy * (size_t)(L->ny + 2);
                const double g = (var[c - sx] + var[c + sx] + var[c - sy] + var[c + sy] +
                                  var[c - sz] + var[c + sz] - 6.0 * var[c]) /
                                 (dxl * dxl);
                resi[c] = src[c] - g;
            }
}

/* The correction levels' boundary (sources.f90:337-344 / 400-407, serial
 * branch): each ghost face is minus the first interior plane, in the order
 * x-low, y-low, z-low, x-high, y-high, z-high, whole planes. */
static void mg_bc(const mg_level *L, double *v) {
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j) v[lidx(L, 0, j, k)] = -v[lidx(L, 1, j, k)];
    for (int k = 0; k <= nz + 1; ++k)
        for (int i = 0; i <= nx + 1; ++i) v[lidx(L, i, 0, k)] = -v[lidx(L, i, 1, k)];
    for (int j = 0; j <= ny + 1; ++j)
        for (int i = 0; i <= nx + 1; ++i) v[lidx(L, i, j, 0)] = -v[lidx(L, i, j, 1)];
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j) v[lidx(L, nx + 1, j, k)] = -v[lidx(L, nx, j, k)];
    for (int k = 0; k <= nz + 1; ++k)
        for (int i = 0; i <= nx + 1; ++i) v[lidx(L, i, ny + 1, k)] = -v[lidx(L, i, ny, k)];
    for (int j = 0; j <= ny + 1; ++j)
        for (int i = 0; i <= nx + 1; ++i) v[lidx(L, i, j, nz + 1)] = -v[lidx(L, i, j, nz)];
}

/* One Gauss-Seidel update (w = 1), sources.f90:362-365; returns |change|. */
static inline double mg_update(const mg_level *L, double *v, const double *src, int i, int j, int k,
                               double dxl) {
    const size_t c = lidx(L, i, j, k);
    const size_t sy = (size_t)(L->nx + 2), sz = sy * (size_t)(L->ny + 2);
    const double w = 1.0;
    const double d1 = v[c];
    v[c] = ((1.0 / 6.0) * (v[c + 1] + v[c - 1] + v[c + sy] + v[c - sy] + v[c + sz] + v[c - sz] -
                           src[c] * (dxl * dxl))) * w + (1.0 - w) * v[c];
    return fabs(v[c] - d1);
}

/* Iterations cap of the coarsest-level solve: the reference iterates until
 * Dif <= 0.001 with no cap (sources.f90:347-376). */
#define ORC_MG_SOLVE_CAP 1000000

/* poisson_solver, sources.f90:302-485, serial branch.  relsolv 1 = Solve
 * (iterate to Dif <= 0.001), 2 = Relax (it-1 iterations of two sweeps, the
 * column's start plane depending on (i, j) parity as written). */
static void mg_poisson(const orc_params *p, const mg_level *L, const double *src, double *v,
                       int relsolv, int it, double *dif, int level) {
    const double dxl = (p->xr - p->xl) / (double)L->nx;
    if (level > 0) mg_bc(L, v);
    if (relsolv == 1) {
        const double er = 0.001;
        *dif = 1000.0;
        int iter = 0;
        while (*dif > er && iter < ORC_MG_SOLVE_CAP) {
            double difp = 0.0;
            for (int i = 1; i <= L->nx; ++i)
                for (int j = 1; j <= L->ny; ++j)
                    for (int k = 1; k <= L->nz; ++k) difp += mg_update(L, v, src, i, j, k, dxl);
            ++iter;
            *dif = difp;
        }
        return;
    }
    for (int iter = 1; iter <= it - 1; ++iter) {
        double difp = 0.0;
        for (int i = 1; i <= L->nx; ++i)
            for (int j = 1; j <= L->ny; ++j) {
                const int start = ((i % 2 == 0) && (j % 2 == 1)) ? 1 : 2;
                for (int k = start; k <= L->nz; k += 2) difp += mg_update(L, v, src, i, j, k, dxl);
            }
        for (int i = 1; i <= L->nx; ++i)
            for (int j = 1; j <= L->ny; ++j) {
                const int start = ((i % 2 == 0) && (j % 2 == 1)) ? 2 : 1;
                for (int k = start; k <= L->nz; k += 2) difp += mg_update(L, v, src, i, j, k, dxl);
            }
        *dif = difp;
    }
}

/* Legendre / aLegendre, sources.f90:9-97 (associated functions with the
 * Condon-Shortley phase). */
static double legendre(double x, int l) {
    switch (l) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return 0.5 * (3.0 * (x * x) - 1.0);
    case 3: return 0.5 * (5.0 * (x * x * x) - 3.0 * x);
    default: return (1.0 / 8.0) * (35.0 * (x * x * x * x) - 30.0 * (x * x) + 3.0);
    }
}
static double alegendre(double x, int l, int m) {
    const double s = pow(1.0 - x * x, 0.5);
    switch (l * 8 + m) {
    case 0: return 1.0;
    case 8: return x;
    case 9: return -s;
    case 16: return 0.5 * (3.0 * (x * x) - 1.0);
    case 17: return -3.0 * x * s;
    case 18: return 3.0 * (1.0 - x * x);
    case 24: return 0.5 * (5.0 * (x * x * x) - 3.0 * x);
    case 25: return (3.0 / 2.0) * (1.0 - 5.0 * (x * x)) * s;
    case 26: return 15.0 * x * (1.0 - x * x);
    case 27: return -15.0 * pow(1.0 - x * x, 1.5);
    case 32: return (1.0 / 8.0) * (35.0 * (x * x * x * x) - 30.0 * (x * x) + 3.0);
    case 33: return -(5.0 / 2.0) * (7.0 * (x * x * x) - 3.0 * x) * s;
    case 34: return (15.0 / 2.0) * (7.0 * (x * x) - 1.0) * (1.0 - x * x);
    case 35: return -105.0 * x * pow(1.0 - x * x, 1.5);
    default: return 105.0 * ((1.0 - x * x) * (1.0 - x * x));
    }
}

static inline double ipow(double r, int l) {
    double v = 1.0;
    for (int q = 0; q < l; ++q) v *= r;
    return v;
}

/* Spherical angles of a cell centre: pangle, aangle, distance
 * (sources.f90:556-597). */
static inline void sph(const orc_params *p, int i, int j, int k, double *r, double *ct, double *fi) {
    const double x = xcoord(p, i), y = ycoord(p, j), z = zcoord(p, k);
    const double rr = sqrt(x * x + y * y + z * z);
    const double the = acos(z / rr);
    double f = atan2(y, x);
    if (f < 0) f = f + 2. * ORC_PI;
    *r = rr;
    *ct = cos(the);
    *fi = f;
}

/* Moments, sources.f90:599-676, with the sums the MPI build reduces (one
 * rank: the local sums).  The serial build reads cmoment, smoment and
 * c0moment uninitialised (sources.f90:642-652).  c0[0..4]; cm, sm [l-1][m-1]. */
void orc_moments(const orc_params *p, const double *rho, double *c0, double *cm, double *sm) {
    const double dx = (p->xr - p->xl) / p->nx, dy = (p->yr - p->yl) / p->ny,
                 dz = (p->zr - p->zl) / p->nz;
    for (int q = 0; q < 16; ++q) cm[q] = sm[q] = 0.0;
    for (int l = 0; l <= 4; ++l) c0[l] = 0.0;
    for (int l = 1; l <= 4; ++l)
        for (int m = 1; m <= l; ++m) {
            double c0mp = 0.0, cmp = 0.0, smp = 0.0;
            const double c = 2.0 * tgamma(l - m + 1.0) / tgamma(l + m + 1.0);
            for (int i = 1; i <= p->nx; ++i)
                for (int j = 1; j <= p->ny; ++j)
                    for (int k = 1; k <= p->nz; ++k) {
                        double r, ct, fi;
                        sph(p, i, j, k, &r, &ct, &fi);
                        const double rl = ipow(r, l), d = rho[phc(p, i, j, k)];
                        c0mp = c0mp + legendre(ct, l) * rl * d * dx * dy * dz;
                        cmp = cmp + alegendre(ct, l, m) * cos(m * fi) * rl * d * dx * dy * dz;
                        smp = smp + alegendre(ct, l, m) * sin(m * fi) * rl * d * dx * dy * dz;
                    }
            cm[(l - 1) * 4 + (m - 1)] = c * cmp;
            sm[(l - 1) * 4 + (m - 1)] = c * smp;
            c0[l] = c0mp;
        }
    double c0mp = 0.0;
    for (int i = 1; i <= p->nx; ++i)
        for (int j = 1; j <= p->ny; ++j)
            for (int k = 1; k <= p->nz; ++k) {
                double r, ct, fi;
                sph(p, i, j, k, &r, &ct, &fi);
                c0mp = c0mp + legendre(ct, 0) * ipow(r, 0) * rho[phc(p, i, j, k)] * dx * dy * dz;
            }
    c0[0] = c0mp;
}

/* The multipole potential at one ghost cell, phi_[xyz]boundaries
 * (sources.f90:678-787): the m > 0 terms carry a factor 2 on top of the
 * one already in cm, sm (as written there). */
static double multipole(const orc_params *p, double gconst, const double *c0, const double *cm,
                        const double *sm, int i, int j, int k) {
    double rb, ct, fi;
    sph(p, i, j, k, &rb, &ct, &fi);
    double phi = 0.0;
    phi = phi - c0[0] / rb;
    for (int l = 1; l <= 4; ++l) {
        phi = phi - c0[l] * legendre(ct, l) / ipow(rb, l + 1);
        for (int m = 1; m <= l; ++m)
            phi = phi - 2.0 * (cm[(l - 1) * 4 + m - 1] * cos(m * fi) +
                               sm[(l - 1) * 4 + m - 1] * sin(m * fi)) *
                            alegendre(ct, l, m) / ipow(rb, l + 1);
    }
    return gconst * phi;
}

/* The six faces of phi (face cells only, edges untouched), in MultiGrid's
 * order: x-low, x-high, y-low, y-high, z-low, z-high (sources.f90:117-136). */
void orc_phi_boundaries(const orc_params *p, double gconst, const double *c0, const double *cm,
                        const double *sm, double *phi) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    for (int side = 0; side < 2; ++side) {
        const int ib = side ? nx + 1 : 0;
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= nz; ++k) phi[phc(p, ib, j, k)] = multipole(p, gconst, c0, cm, sm, ib, j, k);
    }
    for (int side = 0; side < 2; ++side) {
        const int jb = side ? ny + 1 : 0;
        for (int i = 1; i <= nx; ++i)
            for (int k = 1; k <= nz; ++k) phi[phc(p, i, jb, k)] = multipole(p, gconst, c0, cm, sm, i, jb, k);
    }
    for (int side = 0; side < 2; ++side) {
        const int kb = side ? nz + 1 : 0;
        for (int i = 1; i <= nx; ++i)
            for (int j = 1; j <= ny; ++j) phi[phc(p, i, j, kb)] = multipole(p, gconst, c0, cm, sm, i, j, kb);
    }
}

/* MultiGrid, sources.f90:99-205.  dens and sg->phi are (0:n+1)^3 arrays.
 * boundary 2 refreshes phi's faces from the multipole moments when
 * mod(it, 10) = 0 (the MPI build at the domain edge); boundary 1 keeps
 * them (the serial build, whose neighbour ranks are never negative).
 * Returns the number of V-cycles. */
int orc_multigrid(const orc_params *p, orc_sg *sg, const double *dens) {
    mg_level L[32];
    const int lv = sg->levels;
    mg_levels(p, lv, sg->work, L);
    if (sg->boundary == 2 && sg->it % 10 == 0) {
        double c0[5], cm[16], sm[16];
        orc_moments(p, dens, c0, cm, sm);
        orc_phi_boundaries(p, sg->gconst, c0, cm, sm, sg->phi);
    }
    const size_t n0 = lsize(p->nx, p->ny, p->nz);
    const double fpg = 4.0 * ORC_PI * sg->gconst;
    for (size_t q = 0; q < n0; ++q) L[0].r[q] = fpg * dens[q];
    memcpy(L[0].e, sg->phi, n0 * sizeof(double));
    double dif = 1.e28;
    int iter = 0;
    const int its = 6;
    const double tol = sg->tol > 0.0 ? sg->tol : 10.0;   /* sources.f90:140 uses 10 */
    while (dif > tol && iter < 31) {
        ++iter;
        for (int l = 1; l <= lv; ++l)
            memset(L[l].e, 0, lsize(L[l].nx, L[l].ny, L[l].nz) * sizeof(double));
        mg_poisson(p, &L[0], L[0].r, L[0].e, 2, its, &dif, 0);
        for (int l = 1; l <= lv; ++l) {
            mg_resid(p, &L[l - 1], L[l - 1].e, L[l - 1].r, L[l - 1].t);
            mg_restriction(&L[l - 1], L[l - 1].t, &L[l], L[l].r);
            mg_poisson(p, &L[l], L[l].r, L[l].e, l == lv ? 1 : 2, its, &dif, l);
        }
        for (int l = lv - 1; l >= 0; --l) {
            const size_t n = lsize(L[l].nx, L[l].ny, L[l].nz);
            memset(L[l].t, 0, n * sizeof(double));
            mg_prolongation(&L[l + 1], L[l + 1].e, &L[l], L[l].t);
            for (size_t q = 0; q < n; ++q) L[l].e[q] = L[l].e[q] + L[l].t[q];
            mg_poisson(p, &L[l], L[l].r, L[l].e, 2, its, &dif, l);
        }
    }
    memcpy(sg->phi, L[0].e, n0 * sizeof(double));
    sg->cycles = iter;
    sg->dif = dif;
    return iter;
}

/* dens = UU(1, 0:nx+1, 0:ny+1, 0:nz+1) (HLLC.f90:16, mainHD.f90:51). */
void orc_density(const orc_params *p, const double *u, double *dens) {
    for (int k = 0; k <= p->nz + 1; ++k)
        for (int j = 0; j <= p->ny + 1; ++j)
            for (int i = 0; i <= p->nx + 1; ++i) dens[phc(p, i, j, k)] = u[cell(p, i, j, k)];
}

/* mainHD.f90:74-99 with self-gravity: each Godunov call (:80, :88) solves
 * the potential from U^n's density, starting from the last solution, and
 * adds S_Solver's source with PHIT = PHIP + PHI (HLLC.f90:15-24).  dens is
 * scratch of (nx+2)(ny+2)(nz+2) doubles.  Returns dt; sg->it advances. */
double orc_advance_sg(const orc_params *p, double *u, double *up, double *upp, double *prim,
                      double *ff, double *gg, double *hh, double *src, orc_sg *sg, double *dens) {
    const double dt = orc_time_step(p, prim);            /* :74 */
    orc_fluxes(p, u, prim, ff, gg, hh, 1);               /* :76 */
    const double dth = dt / 2.0;                         /* :78 */
    orc_density(p, u, dens);
    orc_multigrid(p, sg, dens);
    orc_gravity_source_phi(p, prim, sg->phi, src);
    orc_godunov(p, upp, u, ff, gg, hh, src, dth);        /* :80 */
    orc_boundaries2(p, upp);                             /* :82 */
    orc_winds(p, upp);
    orc_calc_prims(p, prim, upp);                        /* :84 */
    orc_fluxes(p, upp, prim, ff, gg, hh, 2);             /* :86 */
    orc_density(p, u, dens);
    orc_multigrid(p, sg, dens);
    orc_gravity_source_phi(p, prim, sg->phi, src);
    orc_godunov(p, up, u, ff, gg, hh, src, dt);          /* :88 */
    orc_boundaries1(p, up);                              /* :90 */
    orc_winds(p, up);
    orc_step(p, u, up, upp, dt);                         /* :92 */
    orc_calc_prims(p, prim, u);                          /* :94 */
    if (p->cooling) orc_cooling(p, prim, u, dt);         /* :96-98 */
    sg->it += 1;
    return dt;
}

/* ------------------------------------------------------------------------
 * Cooling integrated to convergence: apply_cooling's per-cell formula
 * (cooling.f90:67-112) applied in substeps h <= eps * t_cool(T), with
 * Lambda re-read from the table at the current temperature every substep,
 * until dt is used up or T drops below the table's first temperature (the
 * reference's gate).  With eps >= dt/t_cool it is one substep, i.e. exactly
 * apply_cooling.  Works on (n, T): p = n KB T, E_th = p/(gamma-1).
 * Returns the new temperature; *nsub receives the substep count. */
double orc_cool_subcycled(const orc_params *p, double n, double T, double dt, double eps,
                          int max_sub, int *nsub) {
    const double Tmin = p->table_T[0];
    double left = dt;
    int s = 0;
    while (left > 0.0 && T >= Tmin && s < max_sub) {
        const double lam = orc_find_coef(p, T);
        const double eth = n * ORC_KB * T / (p->gamma - 1.0);
        const double ce = lam * (n * n) / eth;
        double h = eps / ce;
        if (h > left) h = left;
        double f = exp(-ce * h);
        if (f < 1.0 - p->cooling_limit) f = 1.0 - p->cooling_limit;
        if (T * f < 10.0) f = 10.0 / T;
        T = T * f;
        left -= h;
        ++s;
    }
    if (nsub) *nsub = s;
    return T;
}

/* ------------------------------------------------------------------------
 * Exact integration of the cooling equation dT/dt = -(gamma-1) n Lambda(T)/k_B
 * (the per-cell loss of apply_cooling, cooling.f90:67-112, as a rate) over dt,
 * for Lambda as find_coef interpolates it: a power law on each table segment
 * and the last row's value above the table.  Then the time to cool from T0 to
 * T1 is k_B/((gamma-1) n) (Y(T0) - Y(T1)) with Y(T) = int_{T_min}^{T} dT'/Lambda
 * in closed form, and T(dt) follows by inverting Y on the segment where
 * Y(T0) - dt (gamma-1) n / k_B falls (Townsend 2009, ApJS 181, 391, "exact
 * integration").  It is the limit of orc_cool_subcycled as eps -> 0.  When
 * the cooling time to the gate T_min is at most dt the gas stops at the gate,
 * as the reference's loop does (cooling.f90:71): the result is then
 * gate_frac T_min (the integrator's last substep ends within 1% below it).
 * Returns T0 unchanged below T_min. */
static double orc_seg_slope(const orc_params *p, int i) {
    return (p->table_L[i] - p->table_L[i - 1]) /
           (log10(p->table_T[i]) - log10(p->table_T[i - 1]));
}

/* Y inside segment i (from T_{i-1} to T <= T_i) */
static double orc_seg_Y(const orc_params *p, int i, double T) {
    const double a = 1.0 - orc_seg_slope(p, i);
    const double coef = p->table_T[i - 1] / pow(10.0, p->table_L[i - 1]);
    const double r = T / p->table_T[i - 1];
    return fabs(a) > 1e-12 ? coef * (pow(r, a) - 1.0) / a : coef * log(r);
}

double orc_cool_exact_int(const orc_params *p, double n, double T0, double dt, double gate_frac) {
    const int nt = p->n_table;
    const double Tmin = p->table_T[0], Tmax = p->table_T[nt - 1];
    if (!(T0 >= Tmin)) return T0;
    /* Y at the nodes */
    double ycum[4096];
    if (nt > 4096) return NAN;
    ycum[0] = 0.0;
    for (int i = 1; i < nt; ++i) ycum[i] = ycum[i - 1] + orc_seg_Y(p, i, p->table_T[i]);
    const double lam_top = pow(10.0, p->table_L[nt - 1]);
    /* Y(T0): the segment as find_coef picks it (first row with T_i > T0) */
    double y0;
    if (T0 >= Tmax) {
        y0 = ycum[nt - 1] + (T0 - Tmax) / lam_top;
    } else {
        int i = 1;
        while (!(p->table_T[i] > T0)) ++i;
        y0 = ycum[i - 1] + orc_seg_Y(p, i, T0);
    }
    const double y1 = y0 - dt * (p->gamma - 1.0) * n / ORC_KB;
    if (!(y1 > 0.0)) return gate_frac * Tmin;
    if (y1 >= ycum[nt - 1]) return Tmax + (y1 - ycum[nt - 1]) * lam_top;
    int i = 1;
    while (!(ycum[i] > y1)) ++i;
    /* invert Y on segment i: (T/T_{i-1})^a = 1 + a (y1 - Y_{i-1}) / coef */
    const double a = 1.0 - orc_seg_slope(p, i);
    const double coef = p->table_T[i - 1] / pow(10.0, p->table_L[i - 1]);
    const double s = (y1 - ycum[i - 1]) / coef;
    const double r = fabs(a) > 1e-12 ? pow(1.0 + a * s, 1.0 / a) : exp(s);
    return p->table_T[i - 1] * r;
}

/* ------------------------------------------------------------------------
 * The cooling surrogate: an MLP 3 -> 64 -> 64 -> 1 with GELU (tanh form),
 * inputs (log10 n, log10 T, log10(dt/t_cool)) normalised to [-1, 1] over the
 * training box, output r = log10(T_new / T).  Parameters, flat float array:
 *   W1[64][3] | b1[64] | W2[64][64] (row = output) | b2[64] | W3[64] | b3
 * The GPU kernels round W1..b3 to bf16 for the forward pass, the layer-1
 * activations and the layer-2 pre-activation gradients to bf16 as MFMA
 * operands, and accumulate in FP32.  This FP64 restatement applies the same
 * bf16 roundings at the same points, so its gradients match the GPU's to
 * FP32 accumulation error. */
#define MLP_H 64
#define MLP_NP (MLP_H * 3 + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1)
static const double MLP_C[3] = {0.5, 6.5, 0.0}, MLP_S[3] = {3.5, 2.5, 2.0};

int orc_mlp_nparams(void) { return MLP_NP; }

static double bf16r(double x) {
    float f = (float)x;
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double gelu(double x) {
    const double u = 2.0 * 0.7978845608028654 * (x + 0.044715 * x * x * x);
    return x / (1.0 + exp(-u));
}

static double gelu_d(double x) {
    const double u = 2.0 * 0.7978845608028654 * (x + 0.044715 * x * x * x);
    const double s = 1.0 / (1.0 + exp(-u));
    const double du = 2.0 * 0.7978845608028654 * (1.0 + 3.0 * 0.044715 * x * x);
    return s + x * s * (1.0 - s) * du;
}

/* Forward (out, B values) and, when y != NULL, the mean-squared-error loss
 * and its gradient with respect to the parameters (grad, MLP_NP values). */
static int g_quant = 1;
void orc_mlp_set_quant(int q) { g_quant = q; }
#define bf16r(v) (g_quant ? bf16r(v) : (double)(v))

double orc_mlp(const float *theta, const double *x, const double *y, int B, double *out,
               double *grad) {
    const float *W1 = theta, *b1 = W1 + 3 * MLP_H, *W2 = b1 + MLP_H, *b2 = W2 + MLP_H * MLP_H,
                *W3 = b2 + MLP_H, *b3 = W3 + MLP_H;
    double *gW1 = 0, *gb1 = 0, *gW2 = 0, *gb2 = 0, *gW3 = 0, *gb3 = 0;
    if (grad) {
        memset(grad, 0, sizeof(double) * MLP_NP);
        gW1 = grad;
        gb1 = gW1 + 3 * MLP_H;
        gW2 = gb1 + MLP_H;
        gb2 = gW2 + MLP_H * MLP_H;
        gW3 = gb2 + MLP_H;
        gb3 = gW3 + MLP_H;
    }
    double loss = 0.0;
    for (int b = 0; b < B; ++b) {
        double xn[3], z1[MLP_H], h1q[MLP_H], z2[MLP_H], h2[MLP_H];
        for (int c = 0; c < 3; ++c) xn[c] = (x[3 * b + c] - MLP_C[c]) / MLP_S[c];
        for (int i = 0; i < MLP_H; ++i) {
            double z = bf16r(b1[i]);
            for (int c = 0; c < 3; ++c) z += bf16r(W1[3 * i + c]) * xn[c];
            z1[i] = z;
            h1q[i] = bf16r(gelu(z));
        }
        double o = bf16r(b3[0]);
        for (int j = 0; j < MLP_H; ++j) {
            double z = 0.0;
            for (int i = 0; i < MLP_H; ++i) z += bf16r(W2[MLP_H * j + i]) * h1q[i];
            z2[j] = z + bf16r(b2[j]);
            h2[j] = gelu(z2[j]);
            o += bf16r(W3[j]) * h2[j];
        }
        if (out) out[b] = o;
        if (!y) continue;
        const double d = o - y[b];
        loss += d * d / B;
        if (!grad) continue;
        const double dout = 2.0 * d / B;
        gb3[0] += dout;
        double dz2q[MLP_H];
        for (int j = 0; j < MLP_H; ++j) {
            gW3[j] += dout * h2[j];
            const double dz2 = dout * bf16r(W3[j]) * gelu_d(z2[j]);
            dz2q[j] = bf16r(dz2);
            gb2[j] += dz2q[j];
            for (int i = 0; i < MLP_H; ++i) gW2[MLP_H * j + i] += dz2q[j] * h1q[i];
        }
        for (int i = 0; i < MLP_H; ++i) {
            double dh1 = 0.0;
            for (int j = 0; j < MLP_H; ++j) dh1 += bf16r(W2[MLP_H * j + i]) * dz2q[j];
            const double dz1 = dh1 * gelu_d(z1[i]);
            gb1[i] += dz1;
            for (int c = 0; c < 3; ++c) gW1[3 * i + c] += dz1 * xn[c];
        }
    }
    return loss;
}
#undef bf16r
