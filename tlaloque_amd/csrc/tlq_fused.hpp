// tlq_fused.hpp -- the fused time step: two kernels per iteration of the
// reference's main loop (source/mainHD.f90:74-99).
//
//   k_stage1: U^{1/2} = U^n - dt/2 div F1(U^n) (+ dt/2 S)   HLLC.f90:92-219, Godunov
//             + two-layer face ghosts of U^{1/2}             boundariesII
//             + winds                                        user_boundaries
//   k_stage2: UP = U^n - dt div F2(U^{1/2}) (+ dt S)         HLLC.f90:221-491, Godunov
//             + one-layer face ghosts of U^{n+1} from UP     boundariesI, step
//             + winds, U^{n+1} = UP + eta Lap(U^n)           step
//             + cooling of U^{n+1}                           cooling.f90
//             + CFL maxima of the cooled state               time_step (next step)
//
// A block owns TX x TY columns of cells and marches along z through ZC
// planes.  The z stencil of each column lives in registers; the in-plane
// stencil goes through LDS, one plane at a time.  Global loads for plane
// k+1 are issued at the start of plane k (prefetch) and the block barriers
// wait on LDS traffic only, so those loads stay in flight across the plane's
// compute.  Every interior stencil reaches ghosts along one axis only, so
// only face ghosts are read and written.
#pragma once

#include "tlq_cooling.hpp"
#include "tlq_kernels.hpp"

namespace tlq {

// Launch geometry: blocks cover TX x TY columns and ZC planes.  Block ids are
// remapped so that each of the 8 XCDs (which receive blocks round-robin) gets
// a contiguous range of tiles, keeping tiles that share halos in one L2.
struct Tiling {
    int nbx, nby, nbz;  // tiles per axis
    int ntiles;         // nbx*nby*nbz
    int per_xcd;        // ceil(ntiles/8); the grid has 8*per_xcd blocks
};

__device__ __forceinline__ bool tile_of_block(const Tiling &T, int &bx, int &by, int &bz, int &L) {
    const int b = blockIdx.x;
    L = (b & 7) * T.per_xcd + (b >> 3);
    if (L >= T.ntiles) return false;
    bx = L % T.nbx;
    const int r = L / T.nbx;
    by = r % T.nby;
    bz = r / T.nby;
    return true;
}

// Face-ghost images of an interior cell along one axis (index a in 1..n):
// calls f(ghost_layer, negate) for every ghost layer the boundary fill copies
// from this cell -- two layers (boundariesII) or one (boundariesI).
template <class Fn>
__device__ __forceinline__ void ghost_images(int a, int n, int bcl, int bch, bool two, Fn &&f) {
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
// the boundary fill copies from it, then lets the winds of `wmask` override
// ghosts inside a wind sphere (boundaries, then user_boundaries:
// boundaries.f90:11-24).
template <class R, class V>
__device__ __forceinline__ void write_face_ghosts(const Problem &P, unsigned long long wmask,
                                                  R *__restrict__ out, int i, int j, int k,
                                                  const V v[5], bool two) {
    const Grid &g = P.g;
    auto put = [&](int gi, int gj, int gk, int comp, bool neg) {
        double w[5];
#pragma unroll
        for (int e = 0; e < 5; ++e) w[e] = (neg && e == comp) ? -v[e] : v[e];
        wind_override(P, wmask, gi, gj, gk, w);
        const long long c = g.idx(gi, gj, gk);
#pragma unroll
        for (int e = 0; e < 5; ++e) out[e * g.nc + c] = (R)w[e];
    };
    if (i <= 2 || i >= g.nx - 1)
        ghost_images(i, g.nx, P.bc[0], P.bc[1], two, [&](int q, bool neg) { put(q, j, k, 1, neg); });
    if (j <= 2 || j >= g.ny - 1)
        ghost_images(j, g.ny, P.bc[2], P.bc[3], two, [&](int q, bool neg) { put(i, q, k, 2, neg); });
    if (k <= 2 || k >= g.nz - 1)
        ghost_images(k, g.nz, P.bc[4], P.bc[5], two, [&](int q, bool neg) { put(i, j, q, 3, neg); });
}

// calc_prims of one cell: the reference's arithmetic in FP64 mode, the
// guarded fast path in FP32 mode.  pg is prims_param's value, read from the
// Problem once per kernel (a field read inside a plane loop would be
// reloaded after every block barrier, whose memory clobber covers it).
template <class R>
__device__ __forceinline__ double prims_param(const Problem &P) {
    return sizeof(R) == 8 ? P.gamma : P.gamma - 1.0;
}
template <class R>
__device__ __forceinline__ void prims(double pg, const R u[5], R w[5]) {
    if constexpr (sizeof(R) == 8) cons2prim(pg, u, w);  // pg = gamma
    else cons2prim_fast(pg, u, w);                      // pg = gamma - 1
}

// Addressing of the z march.  A thread computes the byte offset of its
// column (i, j) inside a plane once; every load and store of the march then
// adds it to a block-uniform plane address held in scalar registers, with no
// per-access range check.  Cell (i, j, k) of field e sits at byte
//   e*fbytes + (k+1)*pbytes + off(i, j),  off(i, j) = idx(i, j, -1)*sizeof(R),
// which is below pbytes for -1 <= i <= nx+2, -1 <= j <= ny+2.
struct Column {
    unsigned off;  // byte offset of (i, j) inside its plane (0 outside the array)
};

struct Planes {
    long long fbytes;  // bytes per field
    long long pbytes;  // bytes per z plane
};

template <class R>
__device__ __forceinline__ Planes planes_of(const Grid &g) {
    return Planes{g.nc * (long long)sizeof(R), g.sxy * (long long)sizeof(R)};
}

template <class R>
__device__ __forceinline__ Column column_of(const Grid &g, int i, int j) {
    const bool in = i >= -1 && i <= g.nx + 2 && j >= -1 && j <= g.ny + 2;
    return Column{in ? (unsigned)((g.base - g.sxy + i + (long long)g.sx * j) * (long long)sizeof(R)) : 0u};
}

// Address of field e, plane k (block-uniform).
template <class T>
__device__ __forceinline__ T *plane_ptr(T *A, const Planes &pl, int e, int k) {
    return A + (pl.pbytes * (k + 1) + e * pl.fbytes);
}

// Conserved state of cell (i, j, k) of column c.  A column outside the array
// (only tiles that overhang the mesh have them) reads the first element of
// the plane instead: in bounds, and it only feeds faces of cells that are not
// written.  The z marches only read planes -1 <= k <= nz+2.
template <class R>
__device__ __forceinline__ void load_col(const R *__restrict__ A, const Planes &pl, Column c, int k,
                                         R u[5]) {
#pragma unroll
    for (int e = 0; e < 5; ++e)
        u[e] = *reinterpret_cast<const R *>(plane_ptr(reinterpret_cast<const char *>(A), pl, e, k) +
                                            c.off);
}

// The precomputed point-mass acceleration of cell (i, j, k) of column c
// (FP32 mode; no load when there is no field).  The field's pointer comes
// from the Problem record, so the loads are made global explicitly: as
// generic (flat) loads they would also count against the LDS counter, and
// every LDS barrier would wait for the prefetch of the next plane.
__device__ __forceinline__ void load_gacc(const float *gacc, const Planes &pl, Column c, int k,
                                          float g[3]) {
    if (!gacc) return;
    typedef const __attribute__((address_space(1))) char gchar;
    typedef const __attribute__((address_space(1))) float gfloat;
#pragma unroll
    for (int e = 0; e < 3; ++e)
        g[e] = *(gfloat *)((gchar *)plane_ptr(reinterpret_cast<const char *>(gacc), pl, e, k) + c.off);
}

// Stores the five fields of cell (i, j, k) of column c (which lies inside the array).
template <class R>
__device__ __forceinline__ void store_col(R *__restrict__ A, const Planes &pl, Column c, int k,
                                          const R u[5]) {
#pragma unroll
    for (int e = 0; e < 5; ++e)
        *reinterpret_cast<R *>(plane_ptr(reinterpret_cast<char *>(A), pl, e, k) + c.off) = u[e];
}

// Slope average with the limiter fixed at compile time (LIM > 0) or read
// at run time (LIM == 0).
template <int LIM, class R> __device__ __forceinline__ R avg(int lim, R a, R b) {
    return lim_avg<R>(LIM > 0 ? LIM : lim, a, b);
}

// b + s/2 and b - s/2 (the face states of a cell of value b and slope s) as
// one fused operation each: s/2 is exact, so the result is that of the
// separate product and sum.
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <class R> __device__ __forceinline__ R plus_half(R b, R s) { return fma_(R(0.5), s, b); }
template <class R> __device__ __forceinline__ R minus_half(R b, R s) { return fma_(R(-0.5), s, b); }

// Limited reconstruction of the face between b and c from the four states
// a, b, c, d along the axis (limiter, HLLC.f90:574-582).
template <int LIM, class R>
__device__ __forceinline__ void reconstruct(int lim, const R a[5], const R b[5], const R c[5],
                                            const R d[5], R wl[5], R wr[5]) {
#pragma unroll
    for (int e = 0; e < 5; ++e) {
        const R dl = b[e] - a[e], dm = c[e] - b[e], dr = d[e] - c[e];
        wl[e] = plus_half(b[e], avg<LIM>(lim, dl, dm));
        wr[e] = minus_half(c[e], avg<LIM>(lim, dm, dr));
    }
}

template <int D, int SOL, class R>
__device__ __forceinline__ void riemann_t(R gamma, R gratio, const R wl[5], const R wr[5],
                                          const R ul[5], const R ur[5], R f[5]) {
    if constexpr (SOL == SOLVER_HLL) hll_face<D>(gamma, gratio, wl, wr, ul, ur, f);
    else hllc_face<D>(gamma, gratio, wl, wr, ul, ur, f);
}

// The same with the sound speeds of the two states given.
template <int D, int SOL, class R>
__device__ __forceinline__ void riemann_c(R gratio, const R wl[5], const R wr[5], const R ul[5],
                                          const R ur[5], R csl, R csr, R f[5]) {
    if constexpr (SOL == SOLVER_HLL) hll_face_c<D>(gratio, wl, wr, ul, ur, csl, csr, f);
    else hllc_face_c<D>(gratio, wl, wr, ul, ur, csl, csr, f);
}

// Godunov update of one cell: U - dtdx dF - dtdy dG - dtdz dH.  FP64 mode
// keeps the reference's operation order (HLLC.f90:37-38); FP32 mode forms
// the flux-difference sum in FP32 and accumulates it onto U in FP64.
template <class R>
__device__ __forceinline__ double godunov1(R u, R fp, R fm, R gp, R gm, R hp, R hm, double dtdx,
                                           double dtdy, double dtdz) {
    if constexpr (sizeof(R) == 8) {
        return u - dtdx * (fp - fm) - dtdy * (gp - gm) - dtdz * (hp - hm);
    } else {
        const float d = (float)dtdx * (fp - fm) + (float)dtdy * (gp - gm) + (float)dtdz * (hp - hm);
        return (double)u - (double)d;
    }
}

// The same plus the time-scaled source src (when grav), in the sweep type:
// FP64 mode adds the source after the update (the reference's order); FP32
// mode forms the whole increment in FP32 and applies it with one rounding
// (which is what the FP64 sum of the two FP32 values rounds to as well).
template <class R>
__device__ __forceinline__ R godunov_src(R u, R fp, R fm, R gp, R gm, R hp, R hm, double dtdx,
                                         double dtdy, double dtdz, bool grav, R src) {
    if constexpr (sizeof(R) == 8) {
        const double v = godunov1<R>(u, fp, fm, gp, gm, hp, hm, dtdx, dtdy, dtdz);
        return grav ? v + src : v;
    } else {
        const float d = (float)dtdx * (fp - fm) + (float)dtdy * (gp - gm) + (float)dtdz * (hp - hm);
        return u - (d - src);
    }
}

// Gravity source S (sources.f90:886-893) from the cell's primitives.  In
// FP32 mode `gpre`, when not null, holds the cell's point-mass acceleration
// as gravity_fast gives it (read from the precomputed field), and `phi` is
// the Problem's self-gravity potential (null without one).
template <class R>
__device__ __forceinline__ void gravity_src(const Problem &P, int i, int j, int k, const R w[5],
                                            const float *gpre, const double *phi, double s[5]) {
    if constexpr (sizeof(R) == 8) {
        gravity_source(P, i, j, k, w[0], w[1], w[2], w[3], s);
    } else {
        float g[3];
        if (gpre) {
            g[0] = gpre[0];
            g[1] = gpre[1];
            g[2] = gpre[2];
        } else {
            gravity_fast(P, i, j, k, g);
        }
        if (phi) {  // central differences of the self-gravity potential (FP64)
            g[0] += (float)((phi_at(P, i + 1, j, k) - phi_at(P, i - 1, j, k)) / (2.0 * P.dx));
            g[1] += (float)((phi_at(P, i, j + 1, k) - phi_at(P, i, j - 1, k)) / (2.0 * P.dy));
            g[2] += (float)((phi_at(P, i, j, k + 1) - phi_at(P, i, j, k - 1)) / (2.0 * P.dz));
        }
        const float rg0 = w[0] * g[0], rg1 = w[0] * g[1], rg2 = w[0] * g[2];
        s[0] = 0.0;
        s[1] = -rg0;
        s[2] = -rg1;
        s[3] = -rg2;
        s[4] = -(w[1] * rg0 + w[2] * rg1 + w[3] * rg2);
    }
}

// The gravity source times the step h, in the sweep type.
template <class R>
__device__ __forceinline__ void gravity_src_h(const Problem &P, int i, int j, int k, const R w[5],
                                              const float *gpre, const double *phi, double h,
                                              R s[5]) {
    double sd[5];
    gravity_src<R>(P, i, j, k, w, gpre, phi, sd);
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int e = 0; e < 5; ++e) s[e] = h * sd[e];
    } else {
#pragma unroll
        for (int e = 0; e < 5; ++e) s[e] = (float)h * (float)sd[e];
    }
}

// Halo bookkeeping: thread `tid` loads one cell of the ring around the tile.
//   stage 1: ring of width 1 (no corners), conserved -> sW (prims) and sU
//   stage 2: ring of width 2 of U^{1/2} (no corners) -> sW, and its inner
//            ring -> sU; then the width-1 ring of U^n -> sN
struct HaloTask {
    int lx, ly;  // offset from the tile origin (cell x0+lx, y0+ly)
    int kind;    // 0 none, 1 U^{1/2} (sW [+ sU if inner]), 2 U^n (sN)
};

template <int TX, int TY>
__device__ __forceinline__ HaloTask halo_task1(int h) {
    HaloTask t{0, 0, 0};
    if (h < 2 * TY) {
        t.ly = h >> 1;
        t.lx = (h & 1) ? TX : -1;
        t.kind = 1;
    } else if (h < 2 * (TX + TY)) {
        const int q = h - 2 * TY;
        t.lx = q >> 1;
        t.ly = (q & 1) ? TY : -1;
        t.kind = 1;
    }
    return t;
}

template <int TX, int TY>
__device__ __forceinline__ HaloTask halo_task2(int h) {
    constexpr int NWH = 4 * TY + 4 * TX, NNH = 2 * TY + 2 * TX;
    HaloTask t{0, 0, 0};
    if (h < 4 * TY) {
        t.ly = h >> 2;
        const int s = h & 3;
        t.lx = s < 2 ? s - 2 : TX + (s - 2);
        t.kind = 1;
    } else if (h < NWH) {
        const int q = h - 4 * TY;
        t.lx = q >> 2;
        const int s = q & 3;
        t.ly = s < 2 ? s - 2 : TY + (s - 2);
        t.kind = 1;
    } else if (h < NWH + NNH) {
        const int q = h - NWH;
        if (q < 2 * TY) {
            t.ly = q >> 1;
            t.lx = (q & 1) ? TX : -1;
        } else {
            const int r = q - 2 * TY;
            t.lx = r >> 1;
            t.ly = (r & 1) ? TY : -1;
        }
        t.kind = 2;
    }
    return t;
}

// ---------------------------------------------------------------------------
template <class R, int TX, int TY, int ZC, int SOL, int WAVES>
__global__ __launch_bounds__(TX *TY, WAVES) void k_stage1(const Problem *__restrict__ Pp, Tiling T,
                                                    const R *__restrict__ U,
                                                    R *__restrict__ Uh,
                                                    StepState *__restrict__ st, int parity) {
    constexpr int NT = TX * TY;
    static_assert(2 * (TX + TY) <= NT, "one halo cell per thread");
    constexpr int XW = (TX + 63) / 64 * 64;  // first thread of the extra x faces
    static_assert(XW + TY <= NT, "extra x faces fit the block");
    constexpr int LX = TX + 2, LY = TY + 2;
    __shared__ R sW[5][LY][LX];
    __shared__ R sU[5][LY][LX];
    __shared__ R sC[LY][LX];  // sound speed of each cell, formed once
    __shared__ R sF[5][TY][TX + 1];
    __shared__ R sG[5][TY + 1][TX];
    __shared__ unsigned long long s_wmask;

    const Problem &P = *Pp;
    const Grid &g = P.g;
    int bx, by, bz, L;
    if (!tile_of_block(T, bx, by, bz, L)) return;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int x0 = 1 + bx * TX, y0 = 1 + by * TY, z0 = 1 + bz * ZC;
    const int z1 = min(z0 + ZC - 1, g.nz);
    const int i = x0 + tx, j = y0 + ty;
    const bool active = i <= g.nx && j <= g.ny;
    const HaloTask ht = halo_task1<TX, TY>(tid);
    const Planes pl = planes_of<R>(g);
    const Column cc = column_of<R>(g, i, j);
    const Column ch = column_of<R>(g, x0 + ht.lx, y0 + ht.ly);
    const float *const gacc = sizeof(R) == 4 ? P.gacc : nullptr;

    const double dt = dt_from_maxima(P, st->maxima[parity]);
    const double dth = dt / 2.0;
    if (blockIdx.x == 0 && tid < 4) st->maxima[parity ^ 1][tid] = 0ull;
    const double dtdx = dth / P.dx, dtdy = dth / P.dy, dtdz = dth / P.dz;
    const R gam = (R)P.gamma, gr = (R)(P.gamma / (P.gamma - 1.0));
    const bool grav = P.nmasses > 0 || P.phi != nullptr;
    if (tid == 0)
        s_wmask = P.nwinds ? winds_mask(P, x0 - 2, x0 + TX + 1, y0 - 2, y0 + TY + 1, z0 - 2, z1 + 2)
                           : 0ull;

    // The march starts one plane below the tile (k = z0 - 1), a step whose
    // update is not stored: it forms the z flux through the tile's lower
    // face with the code of every later step, so the two tiles that share
    // the face compute the same flux.
    R uc[5], wc[5], hm[5] = {}, pf_n[5], pf_h[5];
    float pf_g[3] = {};
    load_col(U, pl, cc, z0 - 1, uc);
    prims(prims_param<R>(P), uc, wc);
    R csc = sound_speed(gam, wc);
    load_col(U, pl, cc, z0, pf_n);
    if (ht.kind) load_col(U, pl, ch, z0 - 1, pf_h);

    for (int k = z0 - 1; k <= z1; ++k) {
        R un[5], wn[5], hr[5], hp[5];
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            un[e] = pf_n[e];
            hr[e] = pf_h[e];
        }
        // The plane index is block-uniform; readfirstlane lets the compiler
        // form the plane addresses in scalar registers.
        const int ku = __builtin_amdgcn_readfirstlane(k);
        const float gk[3] = {pf_g[0], pf_g[1], pf_g[2]};
        if (k < z1) {  // prefetch plane k+1 (planes z0-1 .. z1+1 <= nz+1)
            load_col(U, pl, cc, ku + 2, pf_n);
            if (ht.kind) load_col(U, pl, ch, ku + 1, pf_h);
            load_gacc(gacc, pl, cc, ku + 1, pf_g);
        }
        // (Stage 1 reads its few Problem fields inside the loop: held in
        // registers they cost spills at its three waves per SIMD.)
        prims(prims_param<R>(P), un, wn);
        const R csn = sound_speed(gam, wn);
        riemann_c<2, SOL>(gr, wc, wn, uc, un, csc, csn, hp);
        // No barrier before the plane's stores: sW and sU were last read by
        // the previous plane's faces, which its flux barrier closed, and sF
        // and sG are written only after this plane's next barrier.
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            sW[e][ty + 1][tx + 1] = wc[e];
            sU[e][ty + 1][tx + 1] = uc[e];
        }
        sC[ty + 1][tx + 1] = csc;
        if (ht.kind) {
            R wh[5];
            prims(prims_param<R>(P), hr, wh);
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                sW[e][ht.ly + 1][ht.lx + 1] = wh[e];
                sU[e][ht.ly + 1][ht.lx + 1] = hr[e];
            }
            sC[ht.ly + 1][ht.lx + 1] = sound_speed(gam, wh);
        }
        lds_barrier();
        // Faces: the left x face and the lower y face of every tile cell,
        // then the upper y faces of the last row (first wave) and the right
        // x faces of the last column (next wave), so that no wave runs both
        // directions in one pass.  The cell's own state comes from registers
        // (it is what this thread stored), and so do the fluxes of its own
        // two faces in the update.
        auto xface = [&](int fx, int fy, const R *wr, const R *ur, R csr, R f[5]) {  // cells fx-1 | fx
            R wl[5], ul[5];
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                wl[e] = sW[e][fy + 1][fx];
                ul[e] = sU[e][fy + 1][fx];
            }
            riemann_c<0, SOL>(gr, wl, wr, ul, ur, sC[fy + 1][fx], csr, f);
#pragma unroll
            for (int e = 0; e < 5; ++e) sF[e][fy][fx] = f[e];
        };
        auto yface = [&](int fx, int fy, const R *wr, const R *ur, R csr, R f[5]) {  // rows fy-1 | fy
            R wl[5], ul[5];
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                wl[e] = sW[e][fy][fx + 1];
                ul[e] = sU[e][fy][fx + 1];
            }
            riemann_c<1, SOL>(gr, wl, wr, ul, ur, sC[fy][fx + 1], csr, f);
#pragma unroll
            for (int e = 0; e < 5; ++e) sG[e][fy][fx] = f[e];
        };
        // A face whose upper cell (column fx, row fy in tile indices) is a halo cell.
        auto edge = [&](int fx, int fy, bool xdir) {
            R wr[5], ur[5], f[5];
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                wr[e] = sW[e][fy + 1][fx + 1];
                ur[e] = sU[e][fy + 1][fx + 1];
            }
            const R csr = sC[fy + 1][fx + 1];
            if (xdir) xface(fx, fy, wr, ur, csr, f);
            else yface(fx, fy, wr, ur, csr, f);
        };
        R fxm[5], gym[5];
        xface(tx, ty, wc, uc, csc, fxm);
        yface(tx, ty, wc, uc, csc, gym);
        if (tid < TX) edge(tid, TY, false);
        else if (tid >= XW && tid < XW + TY) edge(TX, tid - XW, true);
        lds_barrier();
        if (active && k >= z0) {
            R s[5] = {};
            if (grav) gravity_src_h<R>(P, i, j, k, wc, gacc ? gk : nullptr, P.phi, dth, s);
            R v[5];
#pragma unroll
            for (int e = 0; e < 5; ++e)
                v[e] = godunov_src<R>(uc[e], sF[e][ty][tx + 1], fxm[e], sG[e][ty + 1][tx], gym[e],
                                      hp[e], hm[e], dtdx, dtdy, dtdz, grav, s[e]);
            const unsigned long long wmask = s_wmask;
            R o[5];
#pragma unroll
            for (int e = 0; e < 5; ++e) o[e] = v[e];
            wind_override(P, wmask, i, j, k, o);
            store_col(Uh, pl, cc, ku, o);
            write_face_ghosts(P, wmask, Uh, i, j, k, v, true);
        }
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            hm[e] = hp[e];
            uc[e] = un[e];
            wc[e] = wn[e];
        }
        csc = csn;
    }
}

// ---------------------------------------------------------------------------
template <class R, int TX, int TY, int ZC, int LIM, int SOL, int WAVES>
__global__ __launch_bounds__(TX *TY, WAVES) void k_stage2(const Problem *__restrict__ Pp, Tiling T,
                                                    const R *__restrict__ U,
                                                    const R *__restrict__ Uh,
                                                    R *__restrict__ Unew,
                                                    StepState *__restrict__ st, int parity,
                                                    unsigned *__restrict__ cqueue,
                                                    unsigned *__restrict__ cseg) {
    constexpr int NT = TX * TY;
    static_assert(6 * (TX + TY) <= NT, "one halo cell per thread");
    constexpr int WX = TX + 4, WY = TY + 4;  // two-cell halo (primitives of U^{1/2})
    constexpr int CX = TX + 2, CY = TY + 2;  // one-cell halo (conserved U^{1/2}, U^n)
    // Every array has one writer phase and one reader phase per plane, and
    // U^n's plane alternates between two buffers, so a plane's stores need no
    // barrier against the previous plane's reads (see the loop).
    __shared__ R sW[5][WY][WX];
    __shared__ R sF[5][TY][TX + 1];
    __shared__ R sG[5][TY + 1][TX];
    __shared__ R sU[5][CY][CX];
    __shared__ R sN2[2][5][CY][CX];
    // Reconstructed face states of each cell: x faces of cells -1..TX in the
    // tile rows, y faces of rows -1..TY in the tile columns (m: the cell's
    // lower face, p: its upper face).
    __shared__ R sXm[5][TY][TX + 2], sXp[5][TY][TX + 2];
    __shared__ R sYm[5][TY + 2][TX], sYp[5][TY + 2][TX];
    __shared__ unsigned long long s_wmask;
    // cells of this tile queued for the dense cooling pass: for the network
    // (filled from the front of the tile's segment) and for the integrator
    // (from the back)
    __shared__ unsigned s_nq, s_ni;
    // FP32 cooling table (one-step cooling in FP32 mode)
    __shared__ TabRow sTab[sizeof(R) == 4 ? MAX_TABLE : 1];

    const Problem &P = *Pp;
    const Grid &g = P.g;
    int bx, by, bz, L;
    if (!tile_of_block(T, bx, by, bz, L)) return;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    // The first waves also form the halo face states (phase 2, threads
    // below 2 (TX + TY)) and the tile's upper-edge faces (phase 3): issuing
    // them first when they share a SIMD with another wave shortens the
    // phases that end at a barrier (about 1% of the FP32 kernel's time).
    if (tid < (2 * (TX + TY) + 63) / 64 * 64) __builtin_amdgcn_s_setprio(1);
    const int x0 = 1 + bx * TX, y0 = 1 + by * TY, z0 = 1 + bz * ZC;
    const int z1 = min(z0 + ZC - 1, g.nz);
    const int i = x0 + tx, j = y0 + ty;
    const bool active = i <= g.nx && j <= g.ny;
    const HaloTask ht = halo_task2<TX, TY>(tid);
    const R *__restrict__ hsrc = ht.kind == 2 ? U : Uh;
    const Planes pl = planes_of<R>(g);
    const Column cc = column_of<R>(g, i, j);
    const Column ch = column_of<R>(g, x0 + ht.lx, y0 + ht.ly);
    const float *const gacc = sizeof(R) == 4 ? P.gacc : nullptr;
    // Problem fields the plane loop uses, read once (see prims_param).
    const double pg = prims_param<R>(P);
    const double gamma_d = P.gamma;
    const double *const phi = P.phi;
    const CoolF32 cc32 = cool_consts(P);
    const bool netmode = P.cooling == 2;
    const float net_min = P.net_min_dtc;
    const int nz = g.nz;
    const bool gedge = i <= 2 || i >= g.nx - 1 || j <= 2 || j >= g.ny - 1;
    const long long cell0 = g.idx(i, j, 0), sxy = g.sxy;

    const double dt = dt_from_maxima(P, st->maxima[parity]);
    const double dtdx = dt / P.dx, dtdy = dt / P.dy, dtdz = dt / P.dz;
    const R gam = (R)P.gamma, gr = (R)(P.gamma / (P.gamma - 1.0));
    const int lim = P.limiter;
    const bool grav = P.nmasses > 0 || phi != nullptr;
    // Cooling mode 1: the reference's one-step formula on every cell.  Modes
    // 2, 3 and 4: cells with dt <= 0.01 t_cool take the one substep the
    // converged integrator would take; the others are queued for the dense
    // pass after this kernel, which also forms their CFL maxima.
    const bool cool = P.cooling == 1 && P.ntable > 1;
    const bool coolpass = P.cooling >= 2 && P.ntable > 1;
    unsigned *const qseg = coolpass ? cqueue + (long long)L * (TX * TY * ZC) : nullptr;
    const double g1 = P.gamma - 1.0;
    const double eta = P.eta;
    if (blockIdx.x == 0 && tid == 0) {
        st->time[parity ^ 1] = st->time[parity] + dt;
        st->dt = dt;
        st->it = st->it + 1;
    }
    if (tid == 0) {
        s_wmask = P.nwinds ? winds_mask(P, x0 - 2, x0 + TX + 1, y0 - 2, y0 + TY + 1, z0 - 2, z1 + 2)
                           : 0ull;
        s_nq = 0;
        s_ni = 0;
    }
    if constexpr (sizeof(R) == 4) {
        if (cool || coolpass) stage_table(P, sTab, tid, NT);  // visible after the plane's first barrier
    }

    // z pipeline of U^{1/2}: primitives at k, k+1, conserved at k, k+1, the
    // slope of cell k (limiter, HLLC.f90:574-582); U^n at k-1, k; the z flux
    // at k-1/2.
    R wc[5], wp[5], uc[5], up[5], hz[5], hm[5] = {}, nm[5] = {}, nc[5];
    R pf_q[5], pf_n[5], pf_h[5] = {};
    float pf_g[3] = {};  // point-mass acceleration of the next plane (FP32 field)
    // The march starts one plane below the tile (k = z0 - 1): that step only
    // forms the z flux through the tile's lower face, with the code of every
    // later step, so the two tiles that share the face compute the same flux.
    {
        R wmm[5], t5[5];
        load_col(Uh, pl, cc, z0 - 2, t5);
        prims(pg, t5, wmm);
        load_col(Uh, pl, cc, z0 - 1, uc);
        prims(pg, uc, wc);
        load_col(Uh, pl, cc, z0, up);
        prims(pg, up, wp);
#pragma unroll
        for (int e = 0; e < 5; ++e) hz[e] = avg<LIM>(lim, wc[e] - wmm[e], wp[e] - wc[e]);
        load_col(U, pl, cc, z0 - 1, nc);
    }
    load_col(Uh, pl, cc, z0 + 1, pf_q);
    load_col(U, pl, cc, z0, pf_n);

    float mx = 0.f, my = 0.f, mz = 0.f;
    double mxd = 0.0, myd = 0.0, mzd = 0.0;
    for (int k = z0 - 1; k <= z1; ++k) {
        R uq[5], wq[5], np[5], hr[5], hp[5];
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            uq[e] = pf_q[e];
            np[e] = pf_n[e];
            hr[e] = pf_h[e];
        }
        const float gk[3] = {pf_g[0], pf_g[1], pf_g[2]};
        // The plane index is block-uniform; readfirstlane lets the compiler
        // form the plane addresses in scalar registers.
        const int ku = __builtin_amdgcn_readfirstlane(k);
        if (k < z1) {  // prefetch plane k+1 (planes z0-2 .. z1+2 <= nz+2)
            load_col(Uh, pl, cc, ku + 3, pf_q);
            load_col(U, pl, cc, ku + 2, pf_n);
            if (ht.kind) load_col(hsrc, pl, ch, ku + 1, pf_h);
            load_gacc(gacc, pl, cc, ku + 1, pf_g);
        }
        prims(pg, uq, wq);
        R hzp[5];
        {
            R wl[5], wr[5];
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                hzp[e] = avg<LIM>(lim, wp[e] - wc[e], wq[e] - wp[e]);
                wl[e] = plus_half(wc[e], hz[e]);
                wr[e] = minus_half(wp[e], hzp[e]);
            }
            riemann_t<2, SOL>(gam, gr, wl, wr, uc, up, hp);
        }
        if (k >= z0) {
            // Phases of a plane, each closed by a barrier: (1) store sW, sU, sN;
            // (2) face states from sW; (3) face fluxes from the states and sU into
            // sF, sG; (4) the update from sF, sG, sN.  The previous plane read
            // sW in its phase 2 and sU in its phase 3, both closed by its later
            // barriers; its phase 4 reads the other sN buffer and sF, sG, which
            // this plane rewrites only after its own first barrier.
            auto &sN = sN2[(k - z0) & 1];
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                sW[e][ty + 2][tx + 2] = wc[e];
                sU[e][ty + 1][tx + 1] = uc[e];
                sN[e][ty + 1][tx + 1] = nc[e];
            }
            if (ht.kind == 1) {
                R wh[5];
                prims(pg, hr, wh);
#pragma unroll
                for (int e = 0; e < 5; ++e) sW[e][ht.ly + 2][ht.lx + 2] = wh[e];
                if (ht.lx >= -1 && ht.lx <= TX && ht.ly >= -1 && ht.ly <= TY) {
#pragma unroll
                    for (int e = 0; e < 5; ++e) sU[e][ht.ly + 1][ht.lx + 1] = hr[e];
                }
            } else if (ht.kind == 2) {
#pragma unroll
                for (int e = 0; e < 5; ++e) sN[e][ht.ly + 1][ht.lx + 1] = hr[e];
            }
            lds_barrier();
            // Face states: one limited slope per cell and axis (limiter,
            // HLLC.f90:574-582: the slope of cell i is average(P_i - P_{i-1},
            // P_{i+1} - P_i) for both of its faces).
            auto xstate = [&](int lx, int ly) {
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    const R a = sW[e][ly + 2][lx + 1], b = sW[e][ly + 2][lx + 2], c = sW[e][ly + 2][lx + 3];
                    const R h = avg<LIM>(lim, b - a, c - b);
                    sXm[e][ly][lx + 1] = minus_half(b, h);
                    sXp[e][ly][lx + 1] = plus_half(b, h);
                }
            };
            auto ystate = [&](int lx, int ly) {
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    const R a = sW[e][ly + 1][lx + 2], b = sW[e][ly + 2][lx + 2], c = sW[e][ly + 3][lx + 2];
                    const R h = avg<LIM>(lim, b - a, c - b);
                    sYm[e][ly + 1][lx] = minus_half(b, h);
                    sYp[e][ly + 1][lx] = plus_half(b, h);
                }
            };
            // The tile cell's own slopes: its centre value is in registers, and
            // its lower face states (read only by this thread) stay there too.
            R xm[5], ym[5];
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                const R b = wc[e];
                const R hx = avg<LIM>(lim, b - sW[e][ty + 2][tx + 1], sW[e][ty + 2][tx + 3] - b);
                xm[e] = minus_half(b, hx);
                sXp[e][ty][tx + 1] = plus_half(b, hx);
                const R hy = avg<LIM>(lim, b - sW[e][ty + 1][tx + 2], sW[e][ty + 3][tx + 2] - b);
                ym[e] = minus_half(b, hy);
                sYp[e][ty + 1][tx] = plus_half(b, hy);
            }
            if (tid < 2 * TX) ystate(tid % TX, tid < TX ? -1 : TY);
            else if (tid < 2 * TX + 2 * TY) {
                const int h = tid - 2 * TX;
                xstate((h & 1) ? TX : -1, h >> 1);
            }
            lds_barrier();
            // Faces: the lower x and y face of every tile cell, then the upper y
            // faces of the last row and the upper x faces of the last column.
            // The upper cell's state (wr) and conserved value (ur) of a tile
            // cell's own lower faces come from registers; the fluxes of those two
            // faces stay in registers for the update.
            auto xface = [&](int fx, int fy, const R *wr, const R *ur, R f[5]) {  // cells fx-1 | fx
                R wl[5], ul[5];
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    wl[e] = sXp[e][fy][fx];
                    ul[e] = sU[e][fy + 1][fx];
                }
                riemann_t<0, SOL>(gam, gr, wl, wr, ul, ur, f);
#pragma unroll
                for (int e = 0; e < 5; ++e) sF[e][fy][fx] = f[e];
            };
            auto yface = [&](int fx, int fy, const R *wr, const R *ur, R f[5]) {  // rows fy-1 | fy
                R wl[5], ul[5];
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    wl[e] = sYp[e][fy][fx];
                    ul[e] = sU[e][fy][fx + 1];
                }
                riemann_t<1, SOL>(gam, gr, wl, wr, ul, ur, f);
#pragma unroll
                for (int e = 0; e < 5; ++e) sG[e][fy][fx] = f[e];
            };
            // The faces of the upper row and right column, whose upper cell is a halo cell.
            auto edge = [&](int fx, int fy, bool xdir) {
                R wr[5], ur[5], f[5];
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    wr[e] = xdir ? sXm[e][fy][fx + 1] : sYm[e][fy + 1][fx];
                    ur[e] = sU[e][fy + 1][fx + 1];
                }
                if (xdir) xface(fx, fy, wr, ur, f);
                else yface(fx, fy, wr, ur, f);
            };
            R fxm[5], gym[5];
            xface(tx, ty, xm, uc, fxm);
            yface(tx, ty, ym, uc, gym);
            if (tid < TX) edge(tid, TY, false);
            else if (tid < TX + TY) edge(TX, tid - TX, true);
            lds_barrier();
            if (active) {
                double v[5];
#pragma unroll
                for (int e = 0; e < 5; ++e)
                    v[e] = godunov1<R>(nc[e], sF[e][ty][tx + 1], fxm[e], sG[e][ty + 1][tx], gym[e], hp[e],
                                       hm[e], dtdx, dtdy, dtdz);
                if (grav) {
                    double s[5];
                    gravity_src<R>(P, i, j, k, wc, gacc ? gk : nullptr, phi, s);
#pragma unroll
                    for (int e = 0; e < 5; ++e) v[e] = v[e] + dt * s[e];
                }
                const unsigned long long wmask = s_wmask;
                // ghosts of U^{n+1} = boundary fill of the pre-wind UP, then winds
                if (gedge || k <= 2 || k >= nz - 1) write_face_ghosts(P, wmask, Unew, i, j, k, v, false);
                double o[5];
#pragma unroll
                for (int e = 0; e < 5; ++e) o[e] = v[e];
                wind_override(P, wmask, i, j, k, o);
                // step: U^{n+1} = UP + eta * Lap(U^n)  (HydroCore.f90:184-193)
                R oq[5];
#pragma unroll
                for (int e = 0; e < 5; ++e) {
                    const R lap = sN[e][ty + 1][tx + 2] + sN[e][ty + 1][tx] + sN[e][ty + 2][tx + 1] +
                                  sN[e][ty][tx + 1] + np[e] + nm[e] - R(6) * nc[e];
                    oq[e] = (R)(o[e] + eta * (double)lap);
                }
                // calc_prims + cooling (cooling.f90:64-117) + time_step maxima.
                // Cooling modes 2 and 3: a cell with dt > 0.01 t_cool is queued
                // for the dense pass, which cools it and forms its maxima; the
                // others take the one step of apply_cooling, the converged
                // integrator's single substep for them.
                bool queued = false;
                float qdtc = 0.f;  // dt / t_cool of a queued cell
                if constexpr (sizeof(R) == 8) {
                    const double rho = oq[0];
                    const double vx = oq[1] / oq[0], vy = oq[2] / oq[0], vz = oq[3] / oq[0];
                    double p = pressure64(gamma_d, rho, oq[1], oq[2], oq[3], oq[4]);
                    if (cool || coolpass) {
                        bool cooled;
                        double cdt = 0.0;
                        const double pn = cool_exact(P, rho, p, dt, &cooled, &cdt);
                        queued = coolpass && cooled && cdt > (double)COOL_EPS;
                        qdtc = (float)cdt;
                        if (cooled && !queued) {
                            const double ek = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
                            oq[4] = ek + pn / g1;
                            p = pn;
                        }
                    }
                    if (!queued) {
                        const double cs = sqrt(gamma_d * p / rho);
                        mxd = fmax(mxd, fabs(vx) + cs);
                        myd = fmax(myd, fabs(vy) + cs);
                        mzd = fmax(mzd, fabs(vz) + cs);
                    }
                } else {
                    const float rinv = __builtin_amdgcn_rcpf(oq[0]);
                    const float vx = oq[1] * rinv, vy = oq[2] * rinv, vz = oq[3] * rinv;
                    const double m1 = oq[1], m2 = oq[2], m3 = oq[3];
                    const double ek = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) * inv64(oq[0]);
                    double pd = g1 * ((double)oq[4] - ek);
                    float p = pd > 1e-20 ? (float)pd : 1e-20f;
                    if (cool || coolpass) {
                        bool cooled;
                        float cdt = 0.f;
                        const float pn = cool_fast_tab(cc32, sTab, oq[0], p, (float)dt, &cooled, &cdt);
                        queued = coolpass && cooled && cdt > COOL_EPS;
                        qdtc = cdt;
                        if (cooled && !queued) {
                            oq[4] = (float)(ek + (double)pn * (double)cc32.inv_gm1);
                            p = pn;
                        }
                    }
                    if (!queued) {
                        const float cs = __builtin_amdgcn_sqrtf((float)gamma_d * p * rinv);
                        mx = fmaxf(mx, fabsf(vx) + cs);
                        my = fmaxf(my, fabsf(vy) + cs);
                        mz = fmaxf(mz, fabsf(vz) + cs);
                    }
                }
                store_col(Unew, pl, cc, ku, oq);
                if (coolpass) {  // append to the tile's queue segment: one LDS atomic per wave
                    const bool net = queued && netmode && qdtc >= net_min;
                    const bool itg = queued && !net;
                    const int lane = tid & 63;
                    const unsigned long long nm = __ballot(net), im = __ballot(itg);
                    if (nm) {
                        const int leader = __ffsll((long long)nm) - 1;
                        unsigned qb = 0;
                        if (lane == leader) qb = atomicAdd(&s_nq, (unsigned)__popcll(nm));
                        qb = __shfl(qb, leader);
                        if (net) qseg[qb + __popcll(nm & ((1ull << lane) - 1ull))] = (unsigned)(cell0 + sxy * k);
                    }
                    if (im) {
                        const int leader = __ffsll((long long)im) - 1;
                        unsigned qb = 0;
                        if (lane == leader) qb = atomicAdd(&s_ni, (unsigned)__popcll(im));
                        qb = __shfl(qb, leader);
                        if (itg)
                            qseg[TX * TY * ZC - 1 - (qb + __popcll(im & ((1ull << lane) - 1ull)))] =
                                (unsigned)(cell0 + sxy * k);
                    }
                }
            }
        }
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            hm[e] = hp[e];
            hz[e] = hzp[e];
            wc[e] = wp[e];
            wp[e] = wq[e];
            uc[e] = up[e];
            up[e] = uq[e];
            nm[e] = nc[e];
            nc[e] = np[e];
        }
    }
    if constexpr (sizeof(R) == 8) block_max3<NT>(mxd, myd, mzd, st->maxima[parity ^ 1]);
    else block_max3<NT>((double)mx, (double)my, (double)mz, st->maxima[parity ^ 1]);
    // block_max3 ends with a block barrier: every append precedes it.  cseg
    // holds the network counts of the tiles, then their integrator counts.
    if (coolpass && tid == 0) {
        cseg[L] = s_nq;
        cseg[T.ntiles + L] = s_ni;
    }
}

}  // namespace tlq
