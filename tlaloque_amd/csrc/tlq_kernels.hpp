// tlq_kernels.hpp -- HIP kernels of the HydroCore step for gfx950.
//
// Two families, both templated on the sweep type R (double / float):
//
//  * per-routine kernels: one kernel per reference subroutine, acting on the
//    device copies of the reference's global arrays (U, UP, UPP, UPrim, F, G,
//    H, S).  They keep the reference's data flow exactly and serve the
//    per-routine C-ABI entry points.
//
//  * the fused step: k_stage1 (first-order HLLC predictor, U^{1/2} ghost fill
//    and winds) and k_stage2 (limited HLLC corrector, winds, artificial
//    viscosity, cooling, U^{n+1} ghost fill and the next step's CFL maxima).
//    A block owns a TX x TY column of cells and marches along z through a
//    chunk of ZC planes: the z-direction stencil lives in registers, the
//    in-plane stencil in LDS.  Only face ghosts are written and read: every
//    interior stencil of the scheme (first-order faces, 4-point
//    reconstruction, 7-point viscosity) reaches ghosts along one axis only.
#pragma once

#include "tlq_device.hpp"

namespace tlq {

// ---------------------------------------------------------------------------
// Small helpers

template <class R> struct Fields {  // 5 SoA fields with a shared geometry
    R *p;
    long long nc;
    __device__ __forceinline__ R &at(int e, long long c) const { return p[e * nc + c]; }
};

template <class R>
__device__ __forceinline__ void load5(const R *__restrict__ a, long long nc, long long c, R u[5]) {
#pragma unroll
    for (int e = 0; e < 5; ++e) u[e] = a[e * nc + c];
}
template <class R>
__device__ __forceinline__ void store5(R *__restrict__ a, long long nc, long long c, const R u[5]) {
#pragma unroll
    for (int e = 0; e < 5; ++e) a[e * nc + c] = u[e];
}

__device__ __forceinline__ double dt_from_maxima(const Problem &P,
                                                 const unsigned long long *__restrict__ mx) {
    // time_step, HydroCore.f90:144: cfl * min(dx/mx, dy/my, dz/mz) / sqrt(3)
    const double ax = __longlong_as_double((long long)mx[0]);
    const double ay = __longlong_as_double((long long)mx[1]);
    const double az = __longlong_as_double((long long)mx[2]);
    double m = P.dx / ax;
    if (P.dy / ay < m) m = P.dy / ay;
    if (P.dz / az < m) m = P.dz / az;
    return P.cfl * m / sqrt(3.0);
}

// Wave + block max reduction of three non-negative doubles, then atomicMax
// of their bit patterns (order-preserving for non-negative values).
template <int NT>
__device__ __forceinline__ void block_max3(double a, double b, double c,
                                           unsigned long long *__restrict__ out) {
    __shared__ double red[3][NT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a = fmax(a, __shfl_xor(a, off));
        b = fmax(b, __shfl_xor(b, off));
        c = fmax(c, __shfl_xor(c, off));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wv] = a;
        red[1][wv] = b;
        red[2][wv] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / 64; ++w) {
            a = fmax(a, red[0][w]);
            b = fmax(b, red[1][w]);
            c = fmax(c, red[2][w]);
        }
        atomicMax(out + 0, ord64(a));
        atomicMax(out + 1, ord64(b));
        atomicMax(out + 2, ord64(c));
    }
}

// ---------------------------------------------------------------------------
// Per-routine kernels

// calc_prims, HydroCore.f90:1-33, over every cell of the (n+4)^3 box.
template <class R>
__global__ void k_calc_prims(const Problem *__restrict__ Pp, const R *__restrict__ U,
                             R *__restrict__ W) {
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const long long ncell = (long long)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < ncell;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % (g.nx + 4)) - 1;
        const long long r = t / (g.nx + 4);
        const int j = (int)(r % (g.ny + 4)) - 1, k = (int)(r / (g.ny + 4)) - 1;
        const long long c = g.idx(i, j, k);
        R u[5], w[5];
        load5(U, g.nc, c, u);
        cons2prim(P.gamma, u, w);
        store5(W, g.nc, c, w);
    }
}

// time_step maxima, HydroCore.f90:127-143, from primitives (prims=true) or
// from conserved variables (prims=false, the fused path's initial maxima).
template <class R, int NT>
__global__ __launch_bounds__(NT) void k_maxima(const Problem *__restrict__ Pp,
                                                const R *__restrict__ A, int prims,
                                                unsigned long long *__restrict__ out) {
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const long long n = (long long)g.nx * g.ny * g.nz;
    double mx = 0.0, my = 0.0, mz = 0.0;
    for (long long t = blockIdx.x * (long long)NT + threadIdx.x; t < n;
         t += (long long)gridDim.x * NT) {
        const int i = (int)(t % g.nx) + 1;
        const long long r = t / g.nx;
        const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
        const long long c = g.idx(i, j, k);
        double w[5];
        if (prims) {
#pragma unroll
            for (int e = 0; e < 5; ++e) w[e] = (double)A[e * g.nc + c];
        } else {
            R u[5], wr[5];
            load5(A, g.nc, c, u);
            cons2prim(P.gamma, u, wr);
#pragma unroll
            for (int e = 0; e < 5; ++e) w[e] = (double)wr[e];
        }
        const double cs = sqrt(P.gamma * w[4] / w[0]);
        mx = fmax(mx, fabs(w[1]) + cs);
        my = fmax(my, fabs(w[2]) + cs);
        mz = fmax(mz, fabs(w[3]) + cs);
    }
    block_max3<NT>(mx, my, mz, out);
}

// HLLCFluxes / HLLFluxes (HLLC.f90:64-496): faces i=0..nx, j=0..ny, k=0..nz
// in the three directions; choice 2 reconstructs with the limiter.
template <int D, class R>
__device__ __forceinline__ void face_from_arrays(const Problem &P, const R *__restrict__ U,
                                                 const R *__restrict__ W, int i, int j, int k,
                                                 int choice, R *__restrict__ out) {
    const Grid &g = P.g;
    constexpr int di = D == 0, dj = D == 1, dk = D == 2;
    const long long c0 = g.idx(i, j, k), c1 = g.idx(i + di, j + dj, k + dk);
    R wl[5], wr[5], ul[5], ur[5];
    load5(W, g.nc, c0, wl);
    load5(W, g.nc, c1, wr);
    load5(U, g.nc, c0, ul);
    load5(U, g.nc, c1, ur);
    if (choice == 2) {
        R wll[5], wrr[5];
        load5(W, g.nc, g.idx(i - di, j - dj, k - dk), wll);
        load5(W, g.nc, g.idx(i + 2 * di, j + 2 * dj, k + 2 * dk), wrr);
#pragma unroll
        for (int e = 0; e < 5; ++e) {  // limiter, HLLC.f90:574-582
            const R dl = wl[e] - wll[e], dm = wr[e] - wl[e], dr = wrr[e] - wr[e];
            const R al = lim_avg(P.limiter, dl, dm), ar = lim_avg(P.limiter, dm, dr);
            wl[e] = wl[e] + R(0.5) * al;
            wr[e] = wr[e] - R(0.5) * ar;
        }
    }
    R f[5];
    const R gam = (R)P.gamma, gr = (R)(P.gamma / (P.gamma - 1.0));
    riemann<D>(P.solver, gam, gr, wl, wr, ul, ur, f);
    store5(out, g.nc, c0, f);
}

template <class R>
__global__ void k_fluxes(const Problem *__restrict__ Pp, const R *__restrict__ U,
                         const R *__restrict__ W, R *__restrict__ F, R *__restrict__ G,
                         R *__restrict__ H, int choice) {
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const long long n = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % (g.nx + 1));
        const long long r = t / (g.nx + 1);
        const int j = (int)(r % (g.ny + 1)), k = (int)(r / (g.ny + 1));
        face_from_arrays<0>(P, U, W, i, j, k, choice, F);
        face_from_arrays<1>(P, U, W, i, j, k, choice, G);
        face_from_arrays<2>(P, U, W, i, j, k, choice, H);
    }
}

// Godunov, HLLC.f90:1-62 (interior), plus dt*S when S is given.
template <class R>
__global__ void k_godunov(const Problem *__restrict__ Pp, R *__restrict__ out,
                          const R *__restrict__ U, const R *__restrict__ F,
                          const R *__restrict__ G, const R *__restrict__ H,
                          const R *__restrict__ S, double dt) {
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const double dtdx = dt / P.dx, dtdy = dt / P.dy, dtdz = dt / P.dz;
    const long long n = (long long)g.nx * g.ny * g.nz;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.nx) + 1;
        const long long r = t / g.nx;
        const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
        const long long c = g.idx(i, j, k);
        const long long cx = c - 1, cy = c - g.sx, cz = c - g.sxy;
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            const long long o = e * g.nc;
            double v = (double)U[o + c] - dtdx * ((double)F[o + c] - (double)F[o + cx]) -
                       dtdy * ((double)G[o + c] - (double)G[o + cy]) -
                       dtdz * ((double)H[o + c] - (double)H[o + cz]);
            if (S) v = v + dt * (double)S[o + c];
            out[o + c] = (R)v;
        }
    }
}

// One plane copy of the serial ghost fill (boundaries.f90:179-259, 432-533):
// layer dst <- layer src along `axis` over the full extent of the other two
// axes; the normal momentum is negated for reflective walls.
template <class R>
__global__ void k_bc_plane(const Problem *__restrict__ Pp, R *__restrict__ U, int axis, int dst,
                           int src, int negate) {
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const int na = axis == 0 ? g.ny + 4 : g.nx + 4;  // fastest of the two other axes
    const int nb = axis == 2 ? g.ny + 4 : g.nz + 4;
    const long long n = (long long)na * nb;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(t % na) - 1, b = (int)(t / na) - 1;
        long long cd, cs;
        if (axis == 0) {
            cd = g.idx(dst, a, b);
            cs = g.idx(src, a, b);
        } else if (axis == 1) {
            cd = g.idx(a, dst, b);
            cs = g.idx(a, src, b);
        } else {
            cd = g.idx(a, b, dst);
            cs = g.idx(a, b, src);
        }
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            const R v = U[e * g.nc + cs];
            U[e * g.nc + cd] = (negate && e == axis + 1) ? -v : v;
        }
    }
}

// imposeSphericalWind for all winds (winds.f90:32-90 via user.f90:64-68)
// over a box of cells [lo, hi] (clipped to the array); last wind wins.
template <class R>
__global__ void k_winds(const Problem *__restrict__ Pp, R *__restrict__ U, int i0, int j0, int k0,
                        int ni, int nj, int nk) {
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const long long n = (long long)ni * nj * nk;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = i0 + (int)(t % ni);
        const long long r = t / ni;
        const int j = j0 + (int)(r % nj), k = k0 + (int)(r / nj);
        double ws[5];
        if (winds_at(P, i, j, k, ws)) {
            const long long c = g.idx(i, j, k);
#pragma unroll
            for (int e = 0; e < 5; ++e) U[e * g.nc + c] = (R)ws[e];
        }
    }
}

// Whole-array copy (UPP = U, HydroCore.f90:182).
template <class R>
__global__ void k_copy(R *__restrict__ dst, const R *__restrict__ src, long long n) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

// step, HydroCore.f90:184-206: interior U = UP + eta*Lap(UPP); every ghost
// cell of U takes UP's value.  UPP already holds the old U.
template <class R>
__global__ void k_step(const Problem *__restrict__ Pp, R *__restrict__ U, const R *__restrict__ UP,
                       const R *__restrict__ UPP) {
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const long long ncell = (long long)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < ncell;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % (g.nx + 4)) - 1;
        const long long r = t / (g.nx + 4);
        const int j = (int)(r % (g.ny + 4)) - 1, k = (int)(r / (g.ny + 4)) - 1;
        const long long c = g.idx(i, j, k);
        const bool interior = i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz;
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            const long long o = e * g.nc;
            if (!interior) {
                U[o + c] = UP[o + c];
            } else {
                const R lap = UPP[o + c + 1] + UPP[o + c - 1] + UPP[o + c + g.sx] +
                              UPP[o + c - g.sx] + UPP[o + c + g.sxy] + UPP[o + c - g.sxy] -
                              R(6) * UPP[o + c];
                U[o + c] = (R)((double)UP[o + c] + P.eta * (double)lap);
            }
        }
    }
}

// apply_cooling over the interior (cooling.f90:47-119): writes UPrim(5), U(5).
template <class R>
__global__ void k_cooling(const Problem *__restrict__ Pp, R *__restrict__ W, R *__restrict__ U,
                          double dt) {
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const long long n = (long long)g.nx * g.ny * g.nz;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.nx) + 1;
        const long long r = t / g.nx;
        const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
        const long long c = g.idx(i, j, k);
        const double rho = W[c], u = W[g.nc + c], v = W[2 * g.nc + c], w = W[3 * g.nc + c];
        const double p = W[4 * g.nc + c];
        bool cooled;
        const double pn = cool_exact(P, rho, p, dt, &cooled);
        if (!cooled) continue;
        const double vel2 = u * u + v * v + w * w;
        const double ek = 0.5 * rho * vel2;
        const R pr = (R)pn;
        W[4 * g.nc + c] = pr;
        U[4 * g.nc + c] = (R)(ek + (double)pr / (P.gamma - 1.0));
    }
}

// S_Solver (sources.f90:872-897) from UPrim; S is zero outside the interior.
template <class R>
__global__ void k_gravity(const Problem *__restrict__ Pp, const R *__restrict__ W,
                          R *__restrict__ S) {
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const long long n = (long long)g.nx * g.ny * g.nz;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.nx) + 1;
        const long long r = t / g.nx;
        const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
        const long long c = g.idx(i, j, k);
        double s[5];
        gravity_source(P, i, j, k, W[c], W[g.nc + c], W[2 * g.nc + c], W[3 * g.nc + c], s);
#pragma unroll
        for (int e = 0; e < 5; ++e) S[e * g.nc + c] = (R)s[e];
    }
}

// Fortran AoS real(8) <-> device SoA for planes k0 .. k0+nk-1 (index -1 based).
template <class R>
__global__ void k_aos_to_soa(const Problem *__restrict__ Pp, const double *__restrict__ aos,
                             R *__restrict__ soa, int k0, int nk) {
    const Grid &g = Pp->g;
    const long long plane = (long long)(g.nx + 4) * (g.ny + 4);
    const long long n = plane * nk;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % (g.nx + 4)) - 1;
        const long long r = t / (g.nx + 4);
        const int j = (int)(r % (g.ny + 4)) - 1, k = k0 + (int)(r / (g.ny + 4));
        const long long c = g.idx(i, j, k);
#pragma unroll
        for (int e = 0; e < 5; ++e) soa[e * g.nc + c] = (R)aos[5 * t + e];
    }
}

template <class R>
__global__ void k_soa_to_aos(const Problem *__restrict__ Pp, const R *__restrict__ soa,
                             double *__restrict__ aos, int k0, int nk) {
    const Grid &g = Pp->g;
    const long long plane = (long long)(g.nx + 4) * (g.ny + 4);
    const long long n = plane * nk;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % (g.nx + 4)) - 1;
        const long long r = t / (g.nx + 4);
        const int j = (int)(r % (g.ny + 4)) - 1, k = k0 + (int)(r / (g.ny + 4));
        const long long c = g.idx(i, j, k);
#pragma unroll
        for (int e = 0; e < 5; ++e) aos[5 * t + e] = (double)soa[e * g.nc + c];
    }
}

// Count interior cells whose density is not finite and positive or whose
// energy is not finite.
template <class R>
__global__ void k_check(const Problem *__restrict__ Pp, const R *__restrict__ U,
                        unsigned long long *__restrict__ bad) {
    const Grid &g = Pp->g;
    const long long n = (long long)g.nx * g.ny * g.nz;
    unsigned long long cnt = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.nx) + 1;
        const long long r = t / g.nx;
        const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
        const long long c = g.idx(i, j, k);
        const double rho = U[c], e = U[4 * g.nc + c];
        if (!(rho > 0.0) || !isfinite(rho) || !isfinite(e)) ++cnt;
    }
    if (cnt) atomicAdd(bad, cnt);
}

// ---------------------------------------------------------------------------
// Control block shared by the fused kernels (device memory).
struct StepState {
    unsigned long long maxima[2][4];  // CFL maxima (x, y, z) for step parity 0/1
    double time[2];                   // t for step parity 0/1
    double dt;                        // last dt
    long long it;                     // steps taken
};

}  // namespace tlq
