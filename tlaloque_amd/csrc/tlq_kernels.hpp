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
// Fused step

// Launch geometry of the fused kernels: blocks cover TX x TY columns and ZC
// planes.  Block ids are remapped so that the 8 XCDs (which receive blocks
// round-robin) each get a contiguous range of tiles, keeping tiles that share
// halos in one L2.
struct Tiling {
    int nbx, nby, nbz;   // tiles per axis
    int ntiles;          // nbx*nby*nbz
    int per_xcd;         // ceil(ntiles/8); the grid has 8*per_xcd blocks
};

__device__ __forceinline__ bool tile_of_block(const Tiling &T, int &bx, int &by, int &bz) {
    const int b = blockIdx.x;
    const int L = (b & 7) * T.per_xcd + (b >> 3);
    if (L >= T.ntiles) return false;
    bx = L % T.nbx;
    const int r = L / T.nbx;
    by = r % T.nby;
    bz = r / T.nby;
    return true;
}

// Face-ghost images of an interior cell.  For a cell (i, j, k) along `axis`
// (index a in 1..n), `f(layer, negate)` is called for every ghost layer whose
// boundary fill copies this cell: two layers (boundariesII) or one layer
// (boundariesI).  bcl/bch are the low/high boundary ids of the axis.
template <class Fn>
__device__ __forceinline__ void ghost_images(int a, int n, int bcl, int bch, bool two, Fn &&f) {
    // low face
    if (bcl == BC_OUTFLOW) {
        if (a == 1) {
            f(0, false);
            if (two) f(-1, false);
        }
    } else if (bcl == BC_REFLECTIVE) {
        if (a == 1) f(0, true);
        if (two && a == 2) f(-1, true);
    } else if (bcl == BC_PERIODIC) {
        if (a == n) f(0, false);
        if (two && a == n - 1) f(-1, false);
    }
    // high face
    if (bch == BC_OUTFLOW) {
        if (a == n) {
            f(n + 1, false);
            if (two) f(n + 2, false);
        }
    } else if (bch == BC_REFLECTIVE) {
        if (a == n) f(n + 1, true);
        if (two && a == n - 1) f(n + 2, true);
    } else if (bch == BC_PERIODIC) {
        if (a == 1) f(n + 1, false);
        if (two && a == 2) f(n + 2, false);
    }
}

// Writes the pre-wind value v of interior cell (i, j, k) into the face ghosts
// that the boundary fill copies from it, then lets the winds override ghosts
// inside a wind sphere (boundaries then user_boundaries, boundaries.f90:11-24).
template <class R>
__device__ __forceinline__ void write_face_ghosts(const Problem &P, R *__restrict__ out, int i,
                                                  int j, int k, const double v[5], bool two) {
    const Grid &g = P.g;
    if (i <= 2 || i >= g.nx - 1) {
        ghost_images(i, g.nx, P.bc[0], P.bc[1], two, [&](int gi, bool neg) {
            double w[5];
            const long long c = g.idx(gi, j, k);
            if (!winds_at(P, gi, j, k, w)) {
#pragma unroll
                for (int e = 0; e < 5; ++e) w[e] = v[e];
                if (neg) w[1] = -w[1];
            }
#pragma unroll
            for (int e = 0; e < 5; ++e) out[e * g.nc + c] = (R)w[e];
        });
    }
    if (j <= 2 || j >= g.ny - 1) {
        ghost_images(j, g.ny, P.bc[2], P.bc[3], two, [&](int gj, bool neg) {
            double w[5];
            const long long c = g.idx(i, gj, k);
            if (!winds_at(P, i, gj, k, w)) {
#pragma unroll
                for (int e = 0; e < 5; ++e) w[e] = v[e];
                if (neg) w[2] = -w[2];
            }
#pragma unroll
            for (int e = 0; e < 5; ++e) out[e * g.nc + c] = (R)w[e];
        });
    }
    if (k <= 2 || k >= g.nz - 1) {
        ghost_images(k, g.nz, P.bc[4], P.bc[5], two, [&](int gk, bool neg) {
            double w[5];
            const long long c = g.idx(i, j, gk);
            if (!winds_at(P, i, j, gk, w)) {
#pragma unroll
                for (int e = 0; e < 5; ++e) w[e] = v[e];
                if (neg) w[3] = -w[3];
            }
#pragma unroll
            for (int e = 0; e < 5; ++e) out[e * g.nc + c] = (R)w[e];
        });
    }
}

// Loads cell (i, j, k) (conserved) if it lies inside the array, else a
// harmless uniform state; returns conserved u and primitives w.
template <class R>
__device__ __forceinline__ void load_cell(const Problem &P, const R *__restrict__ A, int i, int j,
                                          int k, R u[5], R w[5]) {
    const Grid &g = P.g;
    if (i >= -1 && i <= g.nx + 2 && j >= -1 && j <= g.ny + 2 && k >= -1 && k <= g.nz + 2) {
        load5(A, g.nc, g.idx(i, j, k), u);
    } else {
        u[0] = R(1);
        u[1] = u[2] = u[3] = R(0);
        u[4] = R(1);
    }
    cons2prim(P.gamma, u, w);
}

template <class R>
__device__ __forceinline__ void load_cons(const Problem &P, const R *__restrict__ A, int i, int j,
                                          int k, R u[5]) {
    const Grid &g = P.g;
    if (i >= -1 && i <= g.nx + 2 && j >= -1 && j <= g.ny + 2 && k >= -1 && k <= g.nz + 2) {
        load5(A, g.nc, g.idx(i, j, k), u);
    } else {
        u[0] = R(1);
        u[1] = u[2] = u[3] = R(0);
        u[4] = R(1);
    }
}

// Control block shared by the fused kernels (device memory).
struct StepState {
    unsigned long long maxima[2][4];  // CFL maxima (x, y, z) for step parity 0/1
    double time[2];                   // t for step parity 0/1
    double dt;                        // last dt
    long long it;                     // steps taken
};

// ---------------------------------------------------------------------------
// Stage 1: U^{1/2} = U^n - dt/2 div F1(U^n) (+ dt/2 S), HLLC.f90:92-219 and
// Godunov(UPP, U, ..., dth); then the two-layer face ghosts of U^{1/2}
// (boundariesII) and the winds (user_boundaries).
template <class R, int TX, int TY, int ZC>
__global__ __launch_bounds__(TX *TY) void k_stage1(const Problem *__restrict__ Pp, Tiling T,
                                                    const R *__restrict__ U,
                                                    R *__restrict__ Uh,
                                                    StepState *__restrict__ st, int parity) {
    constexpr int NT = TX * TY;
    constexpr int LX = TX + 2, LY = TY + 2;  // one-cell halo
    __shared__ R sW[5][LY][LX];
    __shared__ R sU[5][LY][LX];
    __shared__ R sF[5][TY][TX + 1];
    __shared__ R sG[5][TY + 1][TX];

    const Problem &P = *Pp;
    const Grid &g = P.g;
    int bx, by, bz;
    if (!tile_of_block(T, bx, by, bz)) return;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int x0 = 1 + bx * TX, y0 = 1 + by * TY, z0 = 1 + bz * ZC;
    const int z1 = min(z0 + ZC - 1, g.nz);
    const int i = x0 + tx, j = y0 + ty;
    const bool active = i <= g.nx && j <= g.ny;

    const double dt = dt_from_maxima(P, st->maxima[parity]);
    const double dth = dt / 2.0;
    if (blockIdx.x == 0 && tid < 4) st->maxima[parity ^ 1][tid] = 0ull;
    const double dtdx = dth / P.dx, dtdy = dth / P.dy, dtdz = dth / P.dz;
    const R gam = (R)P.gamma, gr = (R)(P.gamma / (P.gamma - 1.0));
    const int solver = P.solver;
    const bool grav = P.nmasses > 0;

    R uc[5], wc[5], un[5], wn[5], hm[5], hp[5];
    {
        R up[5], wp[5];
        load_cell(P, U, i, j, z0 - 1, up, wp);
        load_cell(P, U, i, j, z0, uc, wc);
        riemann<2>(solver, gam, gr, wp, wc, up, uc, hm);
    }
    for (int k = z0; k <= z1; ++k) {
        load_cell(P, U, i, j, k + 1, un, wn);
        riemann<2>(solver, gam, gr, wc, wn, uc, un, hp);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            sW[e][ty + 1][tx + 1] = wc[e];
            sU[e][ty + 1][tx + 1] = uc[e];
        }
        // halo ring: x columns (2*TY cells) then y rows (2*TX cells)
        for (int h = tid; h < 2 * (TX + TY); h += NT) {
            int lx, ly;
            if (h < 2 * TY) {
                ly = 1 + (h >> 1);
                lx = (h & 1) ? TX + 1 : 0;
            } else {
                const int q = h - 2 * TY;
                lx = 1 + (q >> 1);
                ly = (q & 1) ? TY + 1 : 0;
            }
            R uh[5], wh[5];
            load_cell(P, U, x0 + lx - 1, y0 + ly - 1, k, uh, wh);
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                sW[e][ly][lx] = wh[e];
                sU[e][ly][lx] = uh[e];
            }
        }
        __syncthreads();
        for (int q = tid; q < (TX + 1) * TY + TX * (TY + 1); q += NT) {
            R wl[5], wr[5], ul[5], ur[5], f[5];
            if (q < (TX + 1) * TY) {
                const int fy = q / (TX + 1), fx = q % (TX + 1);
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    wl[e] = sW[e][fy + 1][fx];
                    wr[e] = sW[e][fy + 1][fx + 1];
                    ul[e] = sU[e][fy + 1][fx];
                    ur[e] = sU[e][fy + 1][fx + 1];
                }
                riemann<0>(solver, gam, gr, wl, wr, ul, ur, f);
#pragma unroll
                for (int e = 0; e < 5; ++e) sF[e][fy][fx] = f[e];
            } else {
                const int r = q - (TX + 1) * TY;
                const int fy = r / TX, fx = r % TX;
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    wl[e] = sW[e][fy][fx + 1];
                    wr[e] = sW[e][fy + 1][fx + 1];
                    ul[e] = sU[e][fy][fx + 1];
                    ur[e] = sU[e][fy + 1][fx + 1];
                }
                riemann<1>(solver, gam, gr, wl, wr, ul, ur, f);
#pragma unroll
                for (int e = 0; e < 5; ++e) sG[e][fy][fx] = f[e];
            }
        }
        __syncthreads();
        if (active) {
            double v[5];
#pragma unroll
            for (int e = 0; e < 5; ++e)
                v[e] = (double)uc[e] - dtdx * ((double)sF[e][ty][tx + 1] - (double)sF[e][ty][tx]) -
                       dtdy * ((double)sG[e][ty + 1][tx] - (double)sG[e][ty][tx]) -
                       dtdz * ((double)hp[e] - (double)hm[e]);
            if (grav) {
                double s[5];
                gravity_source(P, i, j, k, (double)wc[0], (double)wc[1], (double)wc[2],
                               (double)wc[3], s);
#pragma unroll
                for (int e = 0; e < 5; ++e) v[e] = v[e] + dth * s[e];
            }
            double o[5];
            if (!winds_at(P, i, j, k, o)) {
#pragma unroll
                for (int e = 0; e < 5; ++e) o[e] = v[e];
            }
            const long long c = g.idx(i, j, k);
#pragma unroll
            for (int e = 0; e < 5; ++e) Uh[e * g.nc + c] = (R)o[e];
            write_face_ghosts(P, Uh, i, j, k, v, true);
        }
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            hm[e] = hp[e];
            uc[e] = un[e];
            wc[e] = wn[e];
        }
    }
}

// Limited reconstruction of one face from four primitive states along the
// axis (limiter, HLLC.f90:574-582): returns the left and right face states.
template <class R>
__device__ __forceinline__ void reconstruct(int lim, const R a[5], const R b[5], const R c[5],
                                            const R d[5], R wl[5], R wr[5]) {
#pragma unroll
    for (int e = 0; e < 5; ++e) {
        const R dl = b[e] - a[e], dm = c[e] - b[e], dr = d[e] - c[e];
        wl[e] = b[e] + R(0.5) * lim_avg(lim, dl, dm);
        wr[e] = c[e] - R(0.5) * lim_avg(lim, dm, dr);
    }
}

// ---------------------------------------------------------------------------
// Stage 2: UP = U^n - dt div F2(U^{1/2}) (+ dt S), HLLC.f90:221-491 and
// Godunov(UP, U, ..., dt); winds (borders(UP,1)); U^{n+1} = UP + eta Lap(U^n)
// with ghosts from UP (step, HydroCore.f90:170-211); cooling (cooling.f90);
// CFL maxima of the cooled state for the next step (time_step).
template <class R, int TX, int TY, int ZC>
__global__ __launch_bounds__(TX *TY) void k_stage2(const Problem *__restrict__ Pp, Tiling T,
                                                    const R *__restrict__ U,
                                                    const R *__restrict__ Uh,
                                                    R *__restrict__ Unew,
                                                    StepState *__restrict__ st, int parity) {
    constexpr int NT = TX * TY;
    constexpr int WX = TX + 4, WY = TY + 4;  // two-cell halo (primitives of U^{1/2})
    constexpr int CX = TX + 2, CY = TY + 2;  // one-cell halo (conserved U^{1/2}, U^n)
    __shared__ R sW[5][WY][WX];
    __shared__ R sU[5][CY][CX];
    __shared__ R sN[5][CY][CX];
    __shared__ R sF[5][TY][TX + 1];
    __shared__ R sG[5][TY + 1][TX];

    const Problem &P = *Pp;
    const Grid &g = P.g;
    int bx, by, bz;
    const bool have_tile = tile_of_block(T, bx, by, bz);
    if (!have_tile) return;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int x0 = 1 + bx * TX, y0 = 1 + by * TY, z0 = 1 + bz * ZC;
    const int z1 = min(z0 + ZC - 1, g.nz);
    const int i = x0 + tx, j = y0 + ty;
    const bool active = i <= g.nx && j <= g.ny;

    const double dt = dt_from_maxima(P, st->maxima[parity]);
    const double dtdx = dt / P.dx, dtdy = dt / P.dy, dtdz = dt / P.dz;
    const R gam = (R)P.gamma, gr = (R)(P.gamma / (P.gamma - 1.0));
    const int solver = P.solver, lim = P.limiter;
    const bool grav = P.nmasses > 0;
    const bool cool = P.cooling != 0 && P.ntable > 1;
    const double g1 = P.gamma - 1.0;
    const double eta = P.eta;

    if (blockIdx.x == 0 && tid == 0) {
        st->time[parity ^ 1] = st->time[parity] + dt;
        st->dt = dt;
        st->it = st->it + 1;
    }

    // z pipeline of U^{1/2}: primitives at k-1..k+2, conserved at k..k+2,
    // slope at k; U^n at k-1..k+1.
    R wm[5], wc[5], wp[5], wq[5], uc[5], up[5], uq[5], hm[5], hp[5];
    R nm[5], nc[5], np[5];
    {
        R wmm[5], um[5], t5[5];
        load_cell(P, Uh, i, j, z0 - 2, t5, wmm);
        load_cell(P, Uh, i, j, z0 - 1, um, wm);
        load_cell(P, Uh, i, j, z0, uc, wc);
        load_cell(P, Uh, i, j, z0 + 1, up, wp);
        R wl[5], wr[5];
        reconstruct(lim, wmm, wm, wc, wp, wl, wr);
        riemann<2>(solver, gam, gr, wl, wr, um, uc, hm);
        load_cons(P, U, i, j, z0 - 1, nm);
        load_cons(P, U, i, j, z0, nc);
    }
    double mx = 0.0, my = 0.0, mz = 0.0;
    for (int k = z0; k <= z1; ++k) {
        load_cell(P, Uh, i, j, k + 2, uq, wq);
        load_cons(P, U, i, j, k + 1, np);
        {
            R wl[5], wr[5];
            reconstruct(lim, wm, wc, wp, wq, wl, wr);
            riemann<2>(solver, gam, gr, wl, wr, uc, up, hp);
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            sW[e][ty + 2][tx + 2] = wc[e];
            sU[e][ty + 1][tx + 1] = uc[e];
            sN[e][ty + 1][tx + 1] = nc[e];
        }
        // Halo of U^{1/2}: columns x = -2,-1,TX,TX+1 for TY rows, then rows
        // y = -2,-1,TY,TY+1 for TX columns (no corners); U^n: 1-cell ring.
        constexpr int NWH = 4 * TY + 4 * TX, NNH = 2 * TY + 2 * TX;
        for (int h = tid; h < NWH + NNH; h += NT) {
            if (h < NWH) {
                int lx, ly;  // offsets relative to the tile origin
                if (h < 4 * TY) {
                    ly = h >> 2;
                    const int s = h & 3;
                    lx = s < 2 ? s - 2 : TX + (s - 2);
                } else {
                    const int q = h - 4 * TY;
                    lx = q >> 2;
                    const int s = q & 3;
                    ly = s < 2 ? s - 2 : TY + (s - 2);
                }
                R uh[5], wh[5];
                load_cell(P, Uh, x0 + lx, y0 + ly, k, uh, wh);
#pragma unroll
                for (int e = 0; e < 5; ++e) sW[e][ly + 2][lx + 2] = wh[e];
                if (lx >= -1 && lx <= TX && ly >= -1 && ly <= TY) {
#pragma unroll
                    for (int e = 0; e < 5; ++e) sU[e][ly + 1][lx + 1] = uh[e];
                }
            } else {
                const int q = h - NWH;
                int lx, ly;
                if (q < 2 * TY) {
                    ly = q >> 1;
                    lx = (q & 1) ? TX : -1;
                } else {
                    const int r = q - 2 * TY;
                    lx = r >> 1;
                    ly = (r & 1) ? TY : -1;
                }
                R un[5];
                load_cons(P, U, x0 + lx, y0 + ly, k, un);
#pragma unroll
                for (int e = 0; e < 5; ++e) sN[e][ly + 1][lx + 1] = un[e];
            }
        }
        __syncthreads();
        for (int q = tid; q < (TX + 1) * TY + TX * (TY + 1); q += NT) {
            R a[5], b[5], c[5], d[5], ul[5], ur[5], wl[5], wr[5], f[5];
            if (q < (TX + 1) * TY) {
                // x face between tile cells fx-1 and fx of row fy
                const int fy = q / (TX + 1), fx = q % (TX + 1);
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    a[e] = sW[e][fy + 2][fx];
                    b[e] = sW[e][fy + 2][fx + 1];
                    c[e] = sW[e][fy + 2][fx + 2];
                    d[e] = sW[e][fy + 2][fx + 3];
                    ul[e] = sU[e][fy + 1][fx];
                    ur[e] = sU[e][fy + 1][fx + 1];
                }
                reconstruct(lim, a, b, c, d, wl, wr);
                riemann<0>(solver, gam, gr, wl, wr, ul, ur, f);
#pragma unroll
                for (int e = 0; e < 5; ++e) sF[e][fy][fx] = f[e];
            } else {
                // y face between tile rows fy-1 and fy of column fx
                const int r = q - (TX + 1) * TY;
                const int fy = r / TX, fx = r % TX;
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    a[e] = sW[e][fy][fx + 2];
                    b[e] = sW[e][fy + 1][fx + 2];
                    c[e] = sW[e][fy + 2][fx + 2];
                    d[e] = sW[e][fy + 3][fx + 2];
                    ul[e] = sU[e][fy][fx + 1];
                    ur[e] = sU[e][fy + 1][fx + 1];
                }
                reconstruct(lim, a, b, c, d, wl, wr);
                riemann<1>(solver, gam, gr, wl, wr, ul, ur, f);
#pragma unroll
                for (int e = 0; e < 5; ++e) sG[e][fy][fx] = f[e];
            }
        }
        __syncthreads();
        if (active) {
            double v[5];
#pragma unroll
            for (int e = 0; e < 5; ++e)
                v[e] = (double)nc[e] - dtdx * ((double)sF[e][ty][tx + 1] - (double)sF[e][ty][tx]) -
                       dtdy * ((double)sG[e][ty + 1][tx] - (double)sG[e][ty][tx]) -
                       dtdz * ((double)hp[e] - (double)hm[e]);
            if (grav) {
                double s[5];
                gravity_source(P, i, j, k, (double)wc[0], (double)wc[1], (double)wc[2],
                               (double)wc[3], s);
#pragma unroll
                for (int e = 0; e < 5; ++e) v[e] = v[e] + dt * s[e];
            }
            // UP ghosts: boundary fill of the pre-wind UP, then the winds.
            write_face_ghosts(P, Unew, i, j, k, v, false);
            double o[5];
            if (!winds_at(P, i, j, k, o)) {
#pragma unroll
                for (int e = 0; e < 5; ++e) o[e] = v[e];
            }
            // step: U^{n+1} = UP + eta * Lap(U^n)
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                const R lap = sN[e][ty + 1][tx + 2] + sN[e][ty + 1][tx] + sN[e][ty + 2][tx + 1] +
                              sN[e][ty][tx + 1] + np[e] + nm[e] - R(6) * nc[e];
                o[e] = o[e] + eta * (double)lap;
            }
            // calc_prims + cooling + time_step of the new state
            R oq[5];
#pragma unroll
            for (int e = 0; e < 5; ++e) oq[e] = (R)o[e];
            const double rho = (double)oq[0];
            double vx, vy, vz;
            if constexpr (sizeof(R) == 8) {
                vx = oq[1] / oq[0];
                vy = oq[2] / oq[0];
                vz = oq[3] / oq[0];
            } else {
                const double inv = 1.0 / rho;
                vx = (double)oq[1] * inv;
                vy = (double)oq[2] * inv;
                vz = (double)oq[3] * inv;
            }
            double p = pressure64(P.gamma, rho, (double)oq[1], (double)oq[2], (double)oq[3],
                                  (double)oq[4]);
            if (cool) {
                bool cooled;
                const double pn = cool_exact(P, rho, p, dt, &cooled);
                if (cooled) {
                    const double ek = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
                    const R pr = (R)pn;
                    oq[4] = (R)(ek + (double)pr / g1);
                    p = (double)pr;
                }
            }
            const long long cc = g.idx(i, j, k);
#pragma unroll
            for (int e = 0; e < 5; ++e) Unew[e * g.nc + cc] = oq[e];
            const double cs = sqrt(P.gamma * p / rho);
            mx = fmax(mx, fabs(vx) + cs);
            my = fmax(my, fabs(vy) + cs);
            mz = fmax(mz, fabs(vz) + cs);
        }
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            hm[e] = hp[e];
            wm[e] = wc[e];
            wc[e] = wp[e];
            wp[e] = wq[e];
            uc[e] = up[e];
            up[e] = uq[e];
            nm[e] = nc[e];
            nc[e] = np[e];
        }
    }
    block_max3<NT>(mx, my, mz, st->maxima[parity ^ 1]);
}

}  // namespace tlq
