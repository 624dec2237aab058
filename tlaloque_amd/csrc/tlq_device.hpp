// tlq_device.hpp -- device math of the HydroCore step for gfx950.
//
// Every routine is templated on the sweep type R: double (FP64 mode) or float
// (FP32 mode).  In FP32 mode the pressure recovery E - |m|^2/(2 rho) and the
// accumulation of the conserved update and of the viscosity term run in
// double, and the wind states are computed in double; the reconstruction,
// the Riemann fluxes and the cooling and point-gravity source terms run in
// float (the *_fast routines).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tlq {

// source/constants.f90:6-15 (cgs).
constexpr double AMU = 1.660538782e-24;
constexpr double KB = 1.380650400e-16;
constexpr double GR = 6.67430e-8;
constexpr double PI_REF = 3.14159265358979;

enum { BC_OUTFLOW = 1, BC_REFLECTIVE = 2, BC_PERIODIC = 3, BC_INFLOW = 4 };
enum { LIM_NONE = 1, LIM_VANLEER = 2, LIM_MINMOD = 3, LIM_UMIST = 4, LIM_WOODWARD = 5,
       LIM_SUPERBEE = 6 };
enum { SOLVER_HLLC = 0, SOLVER_HLL = 1 };

// Structure-of-arrays geometry.  Cell (i, j, k), i in -1..nx+2, sits at
// base + i + sx*j + sxy*k inside each field; field e starts at e*nc.  The
// x pitch sx and the base are chosen so that i = 1 of every row is aligned
// to 32 elements (128 B in FP32).
struct Grid {
    int nx, ny, nz;
    int sx;          // elements per x row (padded)
    long long sxy;   // elements per z plane
    long long nc;    // elements per field
    long long base;  // offset of cell (0, 0, 0)
    __host__ __device__ long long idx(int i, int j, int k) const {
        return base + (long long)i + (long long)sx * j + sxy * k;
    }
};

struct WindDev {  // spherical_wind_type, source/winds.f90:5-13
    double xc, yc, zc, vx, vy, vz, radius, mdot, vinf, temp, mu;
};

constexpr int MAX_WINDS = 64;
constexpr int MAX_MASSES = 64;
constexpr int MAX_TABLE = 1024;

// Problem description shared by all kernels (lives in device memory).
struct Problem {
    Grid g;
    double xl, yl, zl, dx, dy, dz;
    double gamma, mu0, cfl, eta, cooling_limit, softening;
    int limiter, solver;
    int bc[6];  // x-, x+, y-, y+, z-, z+
    int nwinds, nmasses, ntable;
    int cooling;  // 0 off, 1 exact, 2 surrogate
    double tmin;  // first table temperature (cooling.f90:42)
    // surrogate mode: cells with dt / t_cool below this (and above 0.01) take
    // the converged integrator, which costs fewer instructions there
    float net_min_dtc;
    WindDev winds[MAX_WINDS];
    double masses[4 * MAX_MASSES];
    double tab_logT[MAX_TABLE];  // log10 T_i
    double tab_T[MAX_TABLE];     // T_i
    double tab_L[MAX_TABLE];     // log10 Lambda_i
    // FP32-mode constants (host-filled): cooling ...
    float f_cT;           // mu0 AMU / KB   (T = p f_cT / rho)
    float f_inv_mu_amu;   // 1 / (mu0 AMU)  (n = rho / (mu0 AMU))
    float f_inv_gm1;      // 1 / (gamma - 1)
    float f_tmin, f_tmax, f_floor_f;  // T_0, T_last, 1 - cooling_limit
    int tab_uniform;      // table equally spaced in log10 T (direct indexing)
    float f_lgT0, f_inv_dlg;
    float tab_Tf[MAX_TABLE], tab_lgTf[MAX_TABLE], tab_Lf[MAX_TABLE], tab_slopef[MAX_TABLE];
    // Y(T) = int_{T_0}^{T} dT'/Lambda(T') at the nodes (double: differences of
    // nearby values are taken), and T_{i-1}/Lambda_{i-1} of segment i (Lambda
    // is a power law on each segment of the table).
    double tab_Ycum[MAX_TABLE];
    float tab_ycoef[MAX_TABLE];
    // ... and gravity in units of L = dx: cell centre X_i = g_X0[m] + i (same for Y, Z
    // with steps g_dY, g_dZ); acceleration factor g_A[m] = 2 G M / L^2.
    float g_A[MAX_MASSES], g_X0[MAX_MASSES], g_Y0[MAX_MASSES], g_Z0[MAX_MASSES];
    float g_dY, g_dZ, g_e2;
    // Self-gravity potential PHI(0:nx+1, 0:ny+1, 0:nz+1) (i fastest), or null.
    const double *phi;
    // FP32 mode: gravity_fast's point-mass acceleration of every interior
    // cell, three fields in the grid's layout, computed when the masses are
    // set (they do not move), or null.
    const float *gacc;
};

// Block barrier for LDS hand-off: waits for this wave's LDS operations only,
// so prefetched global loads are not drained.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// PHI at cell (i, j, k), 0 <= i <= nx+1 (likewise j, k): a global load (a
// generic one would also count against the LDS counter that the fused
// kernels' barriers wait on).
__device__ __forceinline__ double phi_at(const Problem &P, int i, int j, int k) {
    typedef const __attribute__((address_space(1))) double gdouble;
    return ((gdouble *)P.phi)[(long long)i + (long long)(P.g.nx + 2) * ((long long)j + (long long)(P.g.ny + 2) * k)];
}

// ---------------------------------------------------------------------------
// Arithmetic helpers.  FP64 keeps IEEE division and square root (the oracle
// comparison); FP32 uses the hardware reciprocal and square root (1 ulp).
__device__ __forceinline__ double rdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float rdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double rsqrt_(double a) { return sqrt(a); }
__device__ __forceinline__ float rsqrt_(float a) { return __builtin_amdgcn_sqrtf(a); }

template <class R> __device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }
template <class R> __device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }

// Cell-centre coordinates, source/HydroCore.f90:35-75 (single rank).
__device__ __forceinline__ double xcoord(const Problem &P, int i) {
    return (P.xl + P.dx * (double)i) - P.dx / 2.0;
}
__device__ __forceinline__ double ycoord(const Problem &P, int j) {
    return (P.yl + P.dy * (double)j) - P.dy / 2.0;
}
__device__ __forceinline__ double zcoord(const Problem &P, int k) {
    return (P.zl + P.dz * (double)k) - P.dz / 2.0;
}

// Pressure recovery of calc_prims (HydroCore.f90:18-21), always in FP64:
// p = (gamma-1) (E - |m|^2/(2 rho)), floored at 1e-20.
__device__ __forceinline__ double pressure64(double gamma, double r, double mx, double my,
                                             double mz, double e) {
    double p = (gamma - 1.0) * (e - 0.5 * (mx * mx + my * my + mz * mz) / r);
    return p > 1e-20 ? p : 1e-20;
}

// calc_prims for one cell (HydroCore.f90:14-21): w = (rho, u, v, w, p).
template <class R>
__device__ __forceinline__ void cons2prim(double gamma, const R u[5], R w[5]) {
    w[0] = u[0];
    const R inv = rdiv(R(1), u[0]);
    if constexpr (sizeof(R) == 8) {
        w[1] = u[1] / u[0];
        w[2] = u[2] / u[0];
        w[3] = u[3] / u[0];
    } else {
        w[1] = u[1] * inv;
        w[2] = u[2] * inv;
        w[3] = u[3] * inv;
    }
    w[4] = (R)pressure64(gamma, (double)u[0], (double)u[1], (double)u[2], (double)u[3],
                         (double)u[4]);
}

// prim2f, HydroCore.f90:213-251, for direction D (0 x, 1 y, 2 z).
template <int D, class R>
__device__ __forceinline__ void prim2f(R gamma_ratio, const R w[5], R f[5]) {
    constexpr int n = D + 1;
    const R vn = w[n];
    f[0] = w[0] * vn;
#pragma unroll
    for (int c = 1; c <= 3; ++c) {
        const int a = c < n ? c : n, b = c < n ? n : c;
        f[c] = w[0] * w[a] * w[b];
        if (c == n) f[c] = f[c] + w[4];
    }
    f[4] = vn * (R(0.5) * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) + gamma_ratio * w[4]);
}

// average(), HLLC.f90:586-636: the slope of a cell from its two one-sided
// differences a (left) and b (right).
template <class R> __device__ __forceinline__ R lim_avg(int opt, R a, R b) {
    R s, c, d, r;
    switch (opt) {
    case LIM_NONE:
        return R(0.5) * (a + b);
    case LIM_VANLEER:
        if constexpr (sizeof(R) == 8) {
            if (a * b <= 0.0) return 0.0;
            return a * b * (a + b) / (a * a + b * b);
        } else {
            // No contraction here: fma(a, 1/b, b/a) would round the two terms
            // of s differently and lose the symmetry below.
#pragma clang fp contract(off)
            // a b (a+b)/(a^2+b^2) = (a+b) / s with s = a/b + b/a: ratios, so
            // no product underflows for cgs density differences (~1e-25) and
            // none overflows.  Both ratios are formed, so swapping a and b
            // swaps the two terms of s and the result is bitwise symmetric in
            // (a, b), as the product form is (and odd, as it is).  At a
            // reflective wall the ghost's slope of the normal velocity is
            // lim_avg(b, a) against the cell's lim_avg(a, b), and only equal
            // slopes give mirror-image face states: with r = a/b and 1/r in
            // place of b/a the two differed by an ulp, and HLLC left a 1e-5
            // pressure error beside the wall.  Where a b < 0, s < 0 and the
            // max gives the zero slope; a = 0 or b = 0 makes s infinite (1/s =
            // 0), and a = b = 0 a NaN that the max takes to 0.  Three hardware
            // reciprocals, no division, no select.
            const R s = a * __builtin_amdgcn_rcpf(b) + b * __builtin_amdgcn_rcpf(a);
            return (a + b) * fmaxf(__builtin_amdgcn_rcpf(s), 0.f);
        }
    case LIM_MINMOD:
        s = __builtin_copysign(R(1), a);
        r = rmin(R(__builtin_fabs(a)), s * b);
        return s * rmax(r, R(0));
    case LIM_UMIST:
        s = __builtin_copysign(R(1), a);
        c = R(0.25) * a + R(0.75) * b;
        d = R(0.75) * a + R(0.25) * b;
        r = R(2) * R(__builtin_fabs(a));
        r = rmin(r, R(2) * s * b);
        r = rmin(r, s * c);
        r = rmin(r, s * d);
        return s * rmax(r, R(0));
    case LIM_WOODWARD:
        s = __builtin_copysign(R(1), a);
        c = R(0.5) * (a + b);
        r = R(2) * R(__builtin_fabs(a));
        r = rmin(r, R(2) * s * b);
        r = rmin(r, s * c);
        return s * rmax(r, R(0));
    case LIM_SUPERBEE:
        s = __builtin_copysign(R(1), b);
        c = rmin(R(2) * R(__builtin_fabs(b)), s * a);
        d = rmin(R(__builtin_fabs(b)), R(2) * s * a);
        return s * rmax(rmax(R(0), c), d);
    default:
        return R(0);
    }
}

// Sound speed of a primitive state, sqrt(gamma p / rho) (csound,
// HydroCore.f90:157-168, as wavespeed forms it).  A function of the state
// alone: a kernel whose face states are cell states computes it once per
// cell and hands it to the *_c solvers below with the same result.
template <class R> __device__ __forceinline__ R sound_speed(R gamma, const R w[5]) {
    return rsqrt_(rdiv(gamma * w[4], w[0]));
}

// wavespeed, HLLC.f90:498-518 (Davis estimates and the contact speed), from
// the two states and their sound speeds.  (wavespeed below is the same with
// the sound speeds formed in place: the source order the compiler schedules
// best in the reconstructed-state kernel.)
template <int D, class R>
__device__ __forceinline__ void wavespeed_c(const R wl[5], const R wr[5], R csl, R csr, R &sl,
                                            R &sr, R &sst) {
    constexpr int n = D + 1;
    if constexpr (sizeof(R) == 8) {
        sl = rmin(wl[n] - csl, wr[n] - csr);
        sr = rmax(wl[n] + csl, wr[n] + csr);
        sst = rdiv(wr[4] - wl[4] + wl[0] * wl[n] * (sl - wl[n]) - wr[0] * wr[n] * (sr - wr[n]),
                   wl[0] * (sl - wl[n]) - wr[0] * (sr - wr[n]));
    } else {  // the hardware min/max (equal for the finite values)
        sl = fminf(wl[n] - csl, wr[n] - csr);
        sr = fmaxf(wl[n] + csl, wr[n] + csr);
        // The numerator in the reference's order with no contraction: at a
        // face between mirror-image states (a reflective wall) its two
        // momentum terms are then rounded alike and cancel exactly, S* = 0 as
        // in the reference, and the region choice at S* = 0 (the left star
        // state) is the reference's.  A fused multiply-add rounds one term
        // and not the other, and S* = +-1 ulp picks either side.
        R num;
        {
#pragma clang fp contract(off)
            num = wr[4] - wl[4] + wl[0] * wl[n] * (sl - wl[n]) - wr[0] * wr[n] * (sr - wr[n]);
        }
        sst = rdiv(num, wl[0] * (sl - wl[n]) - wr[0] * (sr - wr[n]));
    }
}

template <int D, class R>
__device__ __forceinline__ void wavespeed(R gamma, const R wl[5], const R wr[5], R &sl, R &sr,
                                          R &sst) {
    const R csl = rsqrt_(rdiv(gamma * wl[4], wl[0]));
    const R csr = rsqrt_(rdiv(gamma * wr[4], wr[0]));
    wavespeed_c<D>(wl, wr, csl, csr, sl, sr, sst);
}

// One HLLC face (HLLC.f90:99-135, 275-314): wl/wr reconstructed primitive
// states, ul/ur the cell-centred conserved states of the two cells.  The
// region is selected without branches (supersonic left, supersonic right,
// star left, star right: the reference's if-chain), so a wavefront whose
// lanes fall in different regions runs one path.  FP64 keeps the reference's
// arithmetic; FP32 writes the star energy
//   rs (E/rho + (S*-u)(S* + p/(rho (S-u))))
//     = q E + rs (S*-u) S* + (S*-u) p / (S - S*),  q = (S-u)/(S-S*), rs = rho q
// with one reciprocal.
// `speeds(sl, sr, sst)` supplies the wave speeds.
template <int D, class R, class WS>
__device__ __forceinline__ void hllc_body(R gratio, const R wl[5], const R wr[5], const R ul[5],
                                          const R ur[5], WS &&speeds, R f[5]) {
    constexpr int n = D + 1;
    R sl, sr, sst;
    speeds(sl, sr, sst);
    const bool supl = sl >= R(0), supr = sr <= R(0);
    const bool left = supl || (!supr && sst >= R(0));
    const bool star = !supl && !supr;
    R w[5], u[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) {
        w[e] = left ? wl[e] : wr[e];
        u[e] = left ? ul[e] : ur[e];
    }
    const R s = left ? sl : sr;
    R fs[5], us[5];
    prim2f<D>(gratio, w, fs);
    const R sw = s - w[n];
    if constexpr (sizeof(R) == 8) {
        const R rs = w[0] * sw / (s - sst);
        us[0] = rs;
#pragma unroll
        for (int c = 1; c <= 3; ++c) us[c] = c == n ? rs * sst : rs * w[c];
        us[4] = rs * (u[4] / w[0] + (sst - w[n]) * (sst + w[4] / (w[0] * sw)));
    } else {
        const R inv = __builtin_amdgcn_rcpf(s - sst);
        const R q = sw * inv, rs = w[0] * q, du = sst - w[n];
        us[0] = rs;
#pragma unroll
        for (int c = 1; c <= 3; ++c) us[c] = c == n ? rs * sst : rs * w[c];
        us[4] = q * u[4] + rs * du * sst + du * w[4] * inv;
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) f[e] = star ? fs[e] + s * (us[e] - u[e]) : fs[e];
}

template <int D, class R>
__device__ __forceinline__ void hllc_face(R gamma, R gratio, const R wl[5], const R wr[5],
                                          const R ul[5], const R ur[5], R f[5]) {
    hllc_body<D>(gratio, wl, wr, ul, ur,
                 [&](R &sl, R &sr, R &sst) { wavespeed<D>(gamma, wl, wr, sl, sr, sst); }, f);
}

// The same with the sound speeds of the two states given.
template <int D, class R>
__device__ __forceinline__ void hllc_face_c(R gratio, const R wl[5], const R wr[5], const R ul[5],
                                            const R ur[5], R csl, R csr, R f[5]) {
    hllc_body<D>(gratio, wl, wr, ul, ur,
                 [&](R &sl, R &sr, R &sst) { wavespeed_c<D>(wl, wr, csl, csr, sl, sr, sst); }, f);
}

// One HLL face (HLL.f90:106-128, 197-223).
template <int D, class R, class WS>
__device__ __forceinline__ void hll_body(R gratio, const R wl[5], const R wr[5], const R ul[5],
                                         const R ur[5], WS &&speeds, R f[5]) {
    R sl, sr, sst;
    speeds(sl, sr, sst);
    if (sl >= R(0)) {
        prim2f<D>(gratio, wl, f);
    } else if (sr <= R(0)) {
        prim2f<D>(gratio, wr, f);
    } else {
        R fl[5], fr[5];
        prim2f<D>(gratio, wl, fl);
        prim2f<D>(gratio, wr, fr);
        const R inv = rdiv(R(1), sr - sl);
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            const R num = sr * fl[e] - sl * fr[e] + sl * sr * (ur[e] - ul[e]);
            if constexpr (sizeof(R) == 8) f[e] = num / (sr - sl);
            else f[e] = num * inv;
        }
    }
}

template <int D, class R>
__device__ __forceinline__ void hll_face(R gamma, R gratio, const R wl[5], const R wr[5],
                                         const R ul[5], const R ur[5], R f[5]) {
    hll_body<D>(gratio, wl, wr, ul, ur,
                [&](R &sl, R &sr, R &sst) { wavespeed<D>(gamma, wl, wr, sl, sr, sst); }, f);
}

template <int D, class R>
__device__ __forceinline__ void hll_face_c(R gratio, const R wl[5], const R wr[5], const R ul[5],
                                           const R ur[5], R csl, R csr, R f[5]) {
    hll_body<D>(gratio, wl, wr, ul, ur,
                [&](R &sl, R &sr, R &sst) { wavespeed_c<D>(wl, wr, csl, csr, sl, sr, sst); }, f);
}

template <int D, class R>
__device__ __forceinline__ void riemann(int solver, R gamma, R gratio, const R wl[5],
                                        const R wr[5], const R ul[5], const R ur[5], R f[5]) {
    if (solver == SOLVER_HLL) hll_face<D>(gamma, gratio, wl, wr, ul, ur, f);
    else hllc_face<D>(gamma, gratio, wl, wr, ul, ur, f);
}

// Spherical wind state at a point (winds.f90:69-83); returns false outside
// the wind's radius.  Computed in FP64.
__device__ __noinline__ bool wind_state(const Problem &P, const WindDev &wd, double x, double y,
                                        double z, double out[5]) {
    const double dxw = x - wd.xc, dyw = y - wd.yc, dzw = z - wd.zc;
    const double dist = sqrt(dxw * dxw + dyw * dyw + dzw * dzw);
    if (!(dist < wd.radius)) return false;
    const double dens = wd.mdot / (wd.vinf * dist * dist * 4.0 * PI_REF);
    const double vx = wd.vinf * dxw / dist + wd.vx;
    const double vy = wd.vinf * dyw / dist + wd.vy;
    const double vz = wd.vinf * dzw / dist + wd.vz;
    const double pres = dens * KB * wd.temp / (wd.mu * AMU);
    out[0] = dens;
    out[1] = dens * vx;
    out[2] = dens * vy;
    out[3] = dens * vz;
    out[4] = 0.5 * dens * (vx * vx + vy * vy + vz * vz) + pres / (P.gamma - 1.0);
    return true;
}

// user_boundaries applies the winds in order (user.f90:64-68), so the last
// wind that covers a cell sets it.  Returns true if any wind covers (i,j,k).
__device__ __forceinline__ bool winds_at(const Problem &P, int i, int j, int k, double out[5]) {
    if (P.nwinds == 0) return false;
    const double x = xcoord(P, i), y = ycoord(P, j), z = zcoord(P, k);
    for (int w = P.nwinds - 1; w >= 0; --w)
        if (wind_state(P, P.winds[w], x, y, z, out)) return true;
    return false;
}

// Point-mass potential at a cell centre (pointsource_gravity.f90:30-35 with
// optional Plummer softening).
__device__ __forceinline__ double potential(const Problem &P, int i, int j, int k) {
    const double x = xcoord(P, i), y = ycoord(P, j), z = zcoord(P, k);
    const double e2 = P.softening * P.softening;
    double phi = 0.0;
    for (int s = 0; s < P.nmasses; ++s) {
        const double *m = P.masses + 4 * s;
        const double r = sqrt((m[0] - x) * (m[0] - x) + (m[1] - y) * (m[1] - y) +
                              (m[2] - z) * (m[2] - z) + e2);
        phi = phi - GR * m[3] / r;
    }
    return phi;
}

// PHIT = PHIP + PHI at a cell (HLLC.f90:22): the point masses' potential plus
// the self-gravity potential when there is one.
__device__ __forceinline__ double phit(const Problem &P, int i, int j, int k) {
    const double pp = potential(P, i, j, k);
    return P.phi ? pp + phi_at(P, i, j, k) : pp;
}

// S_Solver (sources.f90:872-897) for one interior cell from its primitives.
__device__ __forceinline__ void gravity_source(const Problem &P, int i, int j, int k,
                                               double rho, double u, double v, double w,
                                               double s[5]) {
    const double g0 = (phit(P, i + 1, j, k) - phit(P, i - 1, j, k)) / (2.0 * P.dx);
    const double g1 = (phit(P, i, j + 1, k) - phit(P, i, j - 1, k)) / (2.0 * P.dy);
    const double g2 = (phit(P, i, j, k + 1) - phit(P, i, j, k - 1)) / (2.0 * P.dz);
    s[0] = 0.0;
    s[1] = -rho * g0;
    s[2] = -rho * g1;
    s[3] = -rho * g2;
    s[4] = -(u * rho * g0) - v * rho * g1 - w * rho * g2;
}

// find_coef, cooling.f90:121-144: first row with T_i > T, then linear
// interpolation of log10 Lambda in log10 T.  Rows are located by bisection
// (same row as the reference's linear scan).  T >= T_max returns the last row.
__device__ __forceinline__ double find_coef(const Problem &P, double temp) {
    int lo = 0, hi = P.ntable - 1;  // invariant: T[lo] <= temp < T[hi] when found
    if (!(P.tab_T[hi] > temp)) return pow(10.0, P.tab_L[hi]);
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (P.tab_T[mid] > temp) hi = mid;
        else lo = mid;
    }
    // hi is the first index with T_hi > temp (T_lo <= temp, given temp >= T_0).
    const double T0 = P.tab_logT[hi - 1], C0 = P.tab_L[hi - 1];
    const double T1 = P.tab_logT[hi], C1 = P.tab_L[hi];
    return pow(10.0, (C1 - C0) / (T1 - T0) * (log10(temp) - T0) + C0);
}

// apply_cooling for one cell (cooling.f90:67-112) on FP64 primitives.
// Returns the cooled pressure; *cooled tells whether the T >= T_min gate
// passed, and *ce_dt (when given) receives dt / t_cool = Lambda n^2 dt / e_th.
__device__ __forceinline__ double cool_exact(const Problem &P, double rho, double p, double dt,
                                             bool *cooled, double *ce_dt = nullptr) {
    const double temp = p * P.mu0 * AMU / (rho * KB);
    if (!(temp >= P.tmin)) {
        *cooled = false;
        return p;
    }
    *cooled = true;
    const double eth = p / (P.gamma - 1.0);
    const double aloss = find_coef(P, temp);
    const double numdens = rho / (P.mu0 * AMU);
    const double ce = aloss * (numdens * numdens) / eth;
    if (ce_dt) *ce_dt = ce * dt;
    double f = exp(-ce * dt);
    if (f < 1.0 - P.cooling_limit) f = 1.0 - P.cooling_limit;
    if (temp * f < 10.0) f = 10.0 / temp;
    return p * f;
}

// ---------------------------------------------------------------------------
// FP32-mode fast paths (the FP64 mode keeps the reference's arithmetic).

// 1/r in FP64 from the FP32 hardware reciprocal and one Newton step
// (relative error ~1e-14).
__device__ __forceinline__ double inv64(float r) {
    const double rd = r;
    double y = (double)__builtin_amdgcn_rcpf(r);
    const double e = __builtin_fma(-rd, y, 1.0);
    return __builtin_fma(y, e, y);
}

// calc_prims in FP32 with the FP64-guarded pressure recovery.
__device__ __forceinline__ void cons2prim_fast(double gm1, const float u[5], float w[5]) {
    // No contraction of these products into the caller's sums: a cell's
    // primitives then round alike wherever they are used (a tile's first
    // planes and the same planes inside the neighbouring tile's march).
#pragma clang fp contract(off)
    const float inv = __builtin_amdgcn_rcpf(u[0]);
    w[0] = u[0];
    w[1] = u[1] * inv;
    w[2] = u[2] * inv;
    w[3] = u[3] * inv;
    const double m1 = u[1], m2 = u[2], m3 = u[3];
    const double m2s = __builtin_fma(m3, m3, __builtin_fma(m2, m2, m1 * m1));
    const double p = gm1 * __builtin_fma(-0.5 * m2s, inv64(u[0]), (double)u[4]);
    w[4] = p > 1e-20 ? (float)p : 1e-20f;
}

// The FP32 cooling table as one float4 per row (T, log10 T, log10 Lambda,
// slope of the segment ending at the row), the form a kernel stages in LDS.
struct TabRow {
    float T, lgT, L, slope;
};

// Fills tab[0..ntable) from the Problem's FP32 table (all threads of a block
// call it; the caller synchronises before use).
__device__ __forceinline__ void stage_table(const Problem &P, TabRow *tab, int tid, int nthreads) {
    for (int q = tid; q < P.ntable; q += nthreads)
        tab[q] = TabRow{P.tab_Tf[q], P.tab_lgTf[q], P.tab_Lf[q], P.tab_slopef[q]};
}

// The FP32 cooling constants of a Problem, read once by a kernel.  The fused
// kernels' block barriers clobber memory, so a Problem field read inside
// their plane loop is reloaded (a scalar load and its wait) on every plane;
// held in a local, it is read once.
struct CoolF32 {
    float cT, inv_mu_amu, inv_gm1, tmin, tmax, floor_f, lgT0, inv_dlg;
    int ntable, uniform;
};

__device__ __forceinline__ CoolF32 cool_consts(const Problem &P) {
    return CoolF32{P.f_cT,   P.f_inv_mu_amu, P.f_inv_gm1, P.f_tmin,  P.f_tmax,
                   P.f_floor_f, P.f_lgT0,    P.f_inv_dlg, P.ntable, P.tab_uniform};
}

// apply_cooling for one cell in FP32 (cooling.f90:67-112) with the table
// rows in LDS: the row is found by direct indexing when the table is equally
// spaced in log10 T (Z1.0.dat is: 0.01 dex), else by bisection; 10^x and exp
// via exp2/log2.  *ce_dt (when given) receives dt / t_cool.
__device__ __forceinline__ float cool_fast_tab(const CoolF32 &C, const TabRow *tab, float rho,
                                               float p, float dt, bool *cooled,
                                               float *ce_dt = nullptr) {
    const float rinv = __builtin_amdgcn_rcpf(rho);
    const float temp = p * C.cT * rinv;
    if (!(temp >= C.tmin)) {
        *cooled = false;
        return p;
    }
    *cooled = true;
    const float lg = __builtin_amdgcn_logf(temp) * 0.301029995663981f;  // log10 T
    float lgl;
    if (!(temp < C.tmax)) {
        lgl = tab[C.ntable - 1].L;
    } else {
        if (C.uniform) {
            // The row index h0 from log10 T is corrected by at most one row
            // either way (rounding of the index): hi = h0 + 1 when T[h0] <=
            // temp (and h0 < n-1), else h0 - 1 when T[h0-1] > temp (and h0 >
            // 1), else h0.  The four rows that (hi-1, hi) can be are read at
            // once, so the lookup costs one LDS round trip instead of three
            // dependent ones.
            int h0 = (int)((lg - C.lgT0) * C.inv_dlg) + 1;
            h0 = h0 < 1 ? 1 : (h0 > C.ntable - 1 ? C.ntable - 1 : h0);
            const TabRow r0 = tab[h0 > 1 ? h0 - 2 : 0], r1 = tab[h0 - 1], r2 = tab[h0],
                         r3 = tab[h0 < C.ntable - 1 ? h0 + 1 : h0];
            const bool up = !(r2.T > temp) && h0 < C.ntable - 1;
            const bool down = !up && h0 > 1 && r1.T > temp;
            const float aL = up ? r2.L : (down ? r0.L : r1.L);
            const float alg = up ? r2.lgT : (down ? r0.lgT : r1.lgT);
            const float bs = up ? r3.slope : (down ? r1.slope : r2.slope);
            lgl = aL + bs * (lg - alg);
        } else {
            int lo = 0, hi = C.ntable - 1;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (tab[mid].T > temp) hi = mid;
                else lo = mid;
            }
            const TabRow a = tab[hi - 1];
            lgl = a.L + tab[hi].slope * (lg - a.lgT);
        }
    }
    const float lam = __builtin_amdgcn_exp2f(lgl * 3.32192809488736f);  // 10^lgl
    const float n = rho * C.inv_mu_amu;
    const float eth = p * C.inv_gm1;
    const float ce = lam * n * n * __builtin_amdgcn_rcpf(eth);
    if (ce_dt) *ce_dt = ce * dt;
    float f = __builtin_amdgcn_exp2f(-ce * dt * 1.44269504088896f);  // exp(-ce dt)
    if (f < C.floor_f) f = C.floor_f;
    if (temp * f < 10.f) f = 10.f * __builtin_amdgcn_rcpf(temp);
    return p * f;
}

// S_Solver's central-difference gradient of the point-mass potential
// (sources.f90:886-888) without cancellation, in FP32: for Phi = -GM/r,
// (Phi(x+dx) - Phi(x-dx)) / (2dx) = 2 G M X / (r+ r- (r+ + r-)), X = x - x_c,
// r+- the (softened) distances of the two neighbour centres -- the same
// quantity as the reference's difference, rewritten with r-^2 - r+^2 = -4 dx X.
__device__ __forceinline__ void gravity_fast(const Problem &P, int i, int j, int k, float g[3]) {
    g[0] = g[1] = g[2] = 0.f;
    for (int m = 0; m < P.nmasses; ++m) {
        const float X = P.g_X0[m] + (float)i, Y = P.g_Y0[m] + P.g_dY * (float)j,
                    Z = P.g_Z0[m] + P.g_dZ * (float)k;
        const float X2 = X * X, Y2 = Y * Y, Z2 = Z * Z, e2 = P.g_e2;
        const float dY = P.g_dY, dZ = P.g_dZ;
        float c[3];
        {
            const float a = (X + 1.f) * (X + 1.f) + Y2 + Z2 + e2, b = (X - 1.f) * (X - 1.f) + Y2 + Z2 + e2;
            const float ra = __builtin_amdgcn_sqrtf(a), rb = __builtin_amdgcn_sqrtf(b);
            c[0] = X * __builtin_amdgcn_rcpf(ra * rb * (ra + rb));
        }
        {
            const float a = X2 + (Y + dY) * (Y + dY) + Z2 + e2, b = X2 + (Y - dY) * (Y - dY) + Z2 + e2;
            const float ra = __builtin_amdgcn_sqrtf(a), rb = __builtin_amdgcn_sqrtf(b);
            c[1] = Y * __builtin_amdgcn_rcpf(ra * rb * (ra + rb));
        }
        {
            const float a = X2 + Y2 + (Z + dZ) * (Z + dZ) + e2, b = X2 + Y2 + (Z - dZ) * (Z - dZ) + e2;
            const float ra = __builtin_amdgcn_sqrtf(a), rb = __builtin_amdgcn_sqrtf(b);
            c[2] = Z * __builtin_amdgcn_rcpf(ra * rb * (ra + rb));
        }
        g[0] += P.g_A[m] * c[0];
        g[1] += P.g_A[m] * c[1];
        g[2] += P.g_A[m] * c[2];
    }
}

// Winds whose sphere can reach the box of cell indices [i0,i1]x[j0,j1]x[k0,k1]
// (bit w set); lets blocks far from every wind skip the per-cell test.
__device__ __forceinline__ unsigned long long winds_mask(const Problem &P, int i0, int i1, int j0,
                                                         int j1, int k0, int k1) {
    unsigned long long m = 0;
    const double x0 = xcoord(P, i0), x1 = xcoord(P, i1), y0 = ycoord(P, j0), y1 = ycoord(P, j1),
                 z0 = zcoord(P, k0), z1 = zcoord(P, k1);
    for (int w = 0; w < P.nwinds; ++w) {
        const WindDev &wd = P.winds[w];
        const double cx = fmin(fmax(wd.xc, x0), x1), cy = fmin(fmax(wd.yc, y0), y1),
                     cz = fmin(fmax(wd.zc, z0), z1);
        const double d2 = (cx - wd.xc) * (cx - wd.xc) + (cy - wd.yc) * (cy - wd.yc) +
                          (cz - wd.zc) * (cz - wd.zc);
        if (d2 <= wd.radius * wd.radius * (1.0 + 1e-9)) m |= 1ull << w;
    }
    return m;
}

// winds_at restricted to the winds of `mask` (last wind wins).
__device__ __forceinline__ bool winds_at_mask(const Problem &P, unsigned long long mask, int i,
                                              int j, int k, double out[5]) {
    if (!mask) return false;
    const double x = xcoord(P, i), y = ycoord(P, j), z = zcoord(P, k);
    for (int w = 63 - __builtin_clzll(mask); w >= 0; --w)
        if (((mask >> w) & 1ull) && wind_state(P, P.winds[w], x, y, z, out)) return true;
    return false;
}

// Replaces v by the state of the last wind of `mask` that covers (i, j, k).
// The wind state passes through a private array only inside the branch, so
// v itself stays in registers (no scratch traffic when no wind is near).
template <class V>
__device__ __forceinline__ bool wind_override(const Problem &P, unsigned long long mask, int i,
                                              int j, int k, V v[5]) {
    if (!mask) return false;
    double o[5];
    if (!winds_at_mask(P, mask, i, j, k, o)) return false;
#pragma unroll
    for (int e = 0; e < 5; ++e) v[e] = (V)o[e];
    return true;
}

// Order-preserving map of a non-negative double onto uint64 for atomicMax.
__device__ __forceinline__ unsigned long long ord64(double x) {
    return (unsigned long long)__double_as_longlong(x);
}

}  // namespace tlq
