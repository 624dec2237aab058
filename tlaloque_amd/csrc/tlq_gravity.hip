// Self-gravity kernels: the multigrid V-cycle of MultiGrid
// (source/sources.f90:99-205) and the multipole boundary values of
// Moments / phi_[xyz]boundaries (sources.f90:599-787), FP64.
#include "tlq_gravity.hpp"

#include <cmath>
#include <cstdlib>

namespace tlq {

namespace {

constexpr int RB_TX = 64, RB_TY = 4;   // red-black relaxation block: i-pairs x rows
constexpr int NT = 256;

__device__ __forceinline__ long long sidx(int nx, int ny, int i, int j, int k) {
    return (long long)i + (long long)(nx + 2) * ((long long)j + (long long)(ny + 2) * k);
}

// Block sum of one double per thread (fixed tree: deterministic).
template <int THREADS>
__device__ __forceinline__ double block_sum(double v, double *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int w = 0; w < THREADS / 64; ++w) s += red[w];
    __syncthreads();
    return s;  // valid in thread 0
}

// Level-0 source 4 pi G rho (Residue(0) = 4.0*PI*Gconst*dens, sources.f90:137).
template <class R>
__device__ __forceinline__ double src0(const R *U, const Grid &g, double fpg, int i, int j, int k) {
    return fpg * (double)U[g.idx(i, j, k)];
}

__global__ void k_sg_fill(double *__restrict__ a, long long n, double v) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x)
        a[q] = v;
}

// Ghost cells of a correction level before a relaxation (sources.f90:337-344,
// serial branch).  The six whole-plane copies x-, y-, z-, x+, y+, z+ (each
// ghost plane = minus the first interior plane) leave every ghost cell at
// (-1)^(number of axes it is a ghost in) times the interior cell reached by
// moving each ghost index one step inwards: each copy negates and moves one
// axis inwards, and a cell's latest copy reads a cell whose other ghost axes
// were copied earlier.  Threads cover the x faces (all j, k), the y faces
// (interior i) and the z faces (interior i, j): every ghost cell once.
// With e2, the same values also go to the second buffer of the level; with
// copy_only, the ghost cells are copied from e to e2 unchanged.
__global__ void k_sg_bc(double *__restrict__ e, double *__restrict__ e2, int nx, int ny, int nz,
                        int copy_only) {
    const long long nxf = 2ll * (ny + 2) * (nz + 2), nyf = 2ll * nx * (nz + 2), nzf = 2ll * nx * ny;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int i, j, k;
    if (t < nxf) {
        const long long q = t >> 1;
        i = (t & 1) ? nx + 1 : 0;
        j = (int)(q % (ny + 2));
        k = (int)(q / (ny + 2));
    } else if (t < nxf + nyf) {
        const long long u = t - nxf, q = u >> 1;
        j = (u & 1) ? ny + 1 : 0;
        i = 1 + (int)(q % nx);
        k = (int)(q / nx);
    } else if (t < nxf + nyf + nzf) {
        const long long u = t - nxf - nyf, q = u >> 1;
        k = (u & 1) ? nz + 1 : 0;
        i = 1 + (int)(q % nx);
        j = 1 + (int)(q / nx);
    } else {
        return;
    }
    const long long c = sidx(nx, ny, i, j, k);
    if (copy_only) {
        e2[c] = e[c];
        return;
    }
    int ci = i, cj = j, ck = k, ng = 0;
    if (ci == 0) { ci = 1; ++ng; } else if (ci == nx + 1) { ci = nx; ++ng; }
    if (cj == 0) { cj = 1; ++ng; } else if (cj == ny + 1) { cj = ny; ++ng; }
    if (ck == 0) { ck = 1; ++ng; } else if (ck == nz + 1) { ck = nz; ++ng; }
    const double v = e[sidx(nx, ny, ci, cj, ck)];
    const double gv = (ng & 1) ? -v : v;
    e[c] = gv;
    if (e2) e2[c] = gv;
}

// One red-black half sweep: cells with (i + j + k) % 2 == color take the
// Gauss-Seidel update of poisson_solver (sources.f90:362-365, w = 1):
// E = (1/6)(E(i+1) + E(i-1) + E(j+1) + E(j-1) + E(k+1) + E(k-1) - s dxl^2).
// Thread (t, j, k) owns the pair i = 2t+1, 2t+2 and updates the one of the
// colour.  With `partial`, each block writes its sum of |change|.
template <class R, bool FROM_U>
__global__ __launch_bounds__(RB_TX *RB_TY) void k_sg_relax(
    double *__restrict__ e, const double *__restrict__ rs, const R *__restrict__ U, Grid g,
    double fpg, int nx, int ny, int nz, double dxl, int color, double *__restrict__ partial) {
    __shared__ double red[RB_TX * RB_TY / 64];
    const int t = blockIdx.x * RB_TX + threadIdx.x;
    const int j = blockIdx.y * RB_TY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    const int i = 2 * t + 1 + ((color + 1 + j + k) & 1);
    double d = 0.0;
    if (i <= nx && j <= ny) {
        const long long sy = nx + 2, sz = (long long)(nx + 2) * (ny + 2);
        const long long c = sidx(nx, ny, i, j, k);
        double s;
        if constexpr (FROM_U) s = src0<R>(U, g, fpg, i, j, k);
        else s = rs[c];
        const double old = e[c];
        const double w = 1.0;
        const double nv = ((1.0 / 6.0) * (e[c + 1] + e[c - 1] + e[c + sy] + e[c - sy] + e[c + sz] +
                                          e[c - sz] - s * (dxl * dxl))) * w + (1.0 - w) * old;
        e[c] = nv;
        d = fabs(nv - old);
    }
    if (partial) {
        const double b = block_sum<RB_TX * RB_TY>(d, red);
        if (threadIdx.x == 0 && threadIdx.y == 0)
            partial[blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z)] = b;
    }
}

// One full red-black iteration in one pass, reading `ein` and writing the
// interior of `eout` (the ghost cells of eout must equal ein's): colour 0
// from the old values of its neighbours, then colour 1 from the new colour-0
// values -- the same updates, in the same order of terms, as the two half
// sweeps of k_sg_relax, at one read and one write of E per iteration.
// A block owns a ZX x ZY column tile and marches planes k0..k1; at step t it
// computes colour 0 of plane t+1 on the tile and a ring of 1, one colour-0
// cell per thread (the ring values are recomputed by the neighbour tiles,
// identically), then colour 1 of plane t on the tile, and writes plane t.  The old plane t+3 (ring 2 around the
// tile) and the sources of the next step are loaded into registers during
// step t, and the barriers wait on LDS only, so the loads stay in flight.
// Block ids are remapped so that each of the 8 XCDs (blocks go to them round
// robin) takes a contiguous range of tiles: tiles that share ring cells then
// share one L2.
constexpr int ZX = 32, ZY = 8, ZZ = 64;

__device__ __forceinline__ int mod3(int t) { return ((t % 3) + 3) % 3; }

template <class R, bool FROM_U>
__global__ __launch_bounds__(ZX *ZY) void k_sg_rbz(
    const double *__restrict__ ein, double *__restrict__ eout, const double *__restrict__ rs,
    const R *__restrict__ U, Grid g, double fpg, int nx, int ny, int nz, double dxl,
    double *__restrict__ partial) {
    constexpr int OX = ZX + 4, OY = ZY + 4, NX = ZX + 2, NY = ZY + 2, T = ZX * ZY;
    static_assert(OX * OY <= 2 * T, "two ring-2 cells per thread");
    __shared__ double O[4][OY][OX];   // old planes, ring 2 (four: see the loop's end)
    __shared__ double N[3][NY][NX];   // new planes, ring 1
    __shared__ double red[T / 64];
    const int tid = threadIdx.x;
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z;
    long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    if (nb % 8 == 0) b = (b % 8) * (nb / 8) + b / 8;
    const int bx = (int)(b % gridDim.x), by = (int)((b / gridDim.x) % gridDim.y),
              bz = (int)(b / ((long long)gridDim.x * gridDim.y));
    const int x0 = bx * ZX, y0 = by * ZY;
    const int k0 = 1 + bz * ZZ, k1 = min(k0 + ZZ - 1, nz);
    const double h2 = dxl * dxl;
    auto src = [&](int i, int j, int k) -> double {
        if constexpr (FROM_U) return src0<R>(U, g, fpg, i, j, k);
        else return rs[sidx(nx, ny, i, j, k)];
    };
    // Per-lane bounds of this thread's cells, fixed along the march, and per
    // wavefront: when every lane of a wavefront is in bounds, its loads run
    // unmasked behind a wavefront-uniform branch.
    bool inO[2], allO[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int q = tid + s * T;
        const int i = x0 + 1 + q % OX - 2, j = y0 + 1 + q / OX - 2;
        inO[s] = q < OX * OY && i >= 0 && i <= nx + 1 && j >= 0 && j <= ny + 1;
        allO[s] = __ballot(inO[s]) == ~0ull;
    }
    // Colour 0 of a plane on the tile and its ring of 1: the NC0 colour-0
    // cells of the NX x NY region, one per thread (tid < NC0).  Thread tid
    // takes row ly = tid / (NX/2) - 1 and the p-th colour-0 cell of that row,
    // p = tid % (NX/2): lx = 2p - 1 when ly + kr is odd, 2p when it is even
    // (x0 and y0 are even, so the colour of (lx, ly) on plane kr is that of
    // lx + ly + kr).
    constexpr int HX = NX / 2, NC0 = HX * NY;
    static_assert(NX % 2 == 0 && ZX % 2 == 0 && ZY % 2 == 0 && NC0 <= T, "one colour-0 cell per thread");
    const bool hasR = tid < NC0;
    const int lyr = tid / HX - 1, lxo = 2 * (tid % HX) - 1;  // lx for ly + kr odd; lxo + 1 for even
    auto red_in = [&](int lx) {
        const int i = x0 + 1 + lx, j = y0 + 1 + lyr;
        return hasR && i >= 1 && i <= nx && j >= 1 && j <= ny;
    };
    const bool inRo = red_in(lxo), inRe = red_in(lxo + 1);
    const bool allR = __ballot(inRo && inRe) == ~0ull, anyR = __ballot(hasR) != 0ull;
    auto red_lx = [&](int kr) { return lxo + ((lyr + kr + 1) & 1); };
    const int ib = x0 + 1 + tid % ZX, jb = y0 + 1 + tid / ZX;
    const bool inB = ib <= nx && jb <= ny;
    const bool allB = __ballot(inB) == ~0ull;
    // old plane k (0 <= k <= nz+1), this thread's two ring-2 cells
    auto fetch_plane = [&](int k, double (&r)[2]) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int q = tid + s * T;
            const int lx = q % OX - 2, ly = q / OX - 2;
            const int i = x0 + 1 + lx, j = y0 + 1 + ly;
            if (allO[s]) r[s] = ein[sidx(nx, ny, i, j, k)];
            else r[s] = inO[s] ? ein[sidx(nx, ny, i, j, k)] : 0.0;
        }
    };
    auto store_plane = [&](int k, const double (&r)[2]) {
        double(*o)[OX] = O[k & 3];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int q = tid + s * T;
            if (q < OX * OY) o[q / OX][q % OX] = r[s];
        }
    };
    // source of this thread's colour-0 cell on plane kr (0 outside the interior)
    auto fetch_src_red = [&](int kr) -> double {
        const int lx = red_lx(kr);
        const int i = x0 + 1 + lx, j = y0 + 1 + lyr;
        double v = 0.0;
        if (kr >= 1 && kr <= nz) {
            if (allR) v = src(i, j, kr);
            else if (anyR && ((lx == lxo) ? inRo : inRe)) v = src(i, j, kr);
        }
        return v;
    };
    // source of this thread's tile cell on plane t (colour 1, interior)
    auto fetch_src_blk = [&](int t) -> double {
        double v = 0.0;
        if (t >= 1 && t <= nz) {
            if (allB) v = src(ib, jb, t);
            else if (inB) v = src(ib, jb, t);
        }
        return ((ib + jb + t) & 1) == 1 ? v : 0.0;
    };
    double pf[2], sr, sb;
    if (k0 - 2 >= 0) {
        fetch_plane(k0 - 2, pf);
        store_plane(k0 - 2, pf);
    }
    fetch_plane(k0 - 1, pf);
    store_plane(k0 - 1, pf);
    fetch_plane(k0, pf);        // plane t+2 of the first step (k0 <= nz)
    sr = fetch_src_red(k0 - 1);
    sb = fetch_src_blk(k0 - 2);
    double dsum = 0.0;
    // N's buffers of planes t-1, t, t+1 (mod 3), rotated along the march
    int nm = mod3(k0 - 3), n0 = mod3(k0 - 2), np = mod3(k0 - 1);
    for (int t = k0 - 2; t <= k1; ++t) {
        if (t + 2 <= nz + 1) store_plane(t + 2, pf);
        const double src_r = sr, sbc = sb;
        if (t + 3 <= nz + 1) fetch_plane(t + 3, pf);
        sr = fetch_src_red(t + 2);
        sb = fetch_src_blk(t + 1);
        lds_barrier();
        // colour 0 of plane kr = t+1 on the tile and its ring of 1
        const int kr = t + 1;
        {
            double(*om)[OX] = O[(kr - 1) & 3];
            double(*oc)[OX] = O[kr & 3];
            double(*op)[OX] = O[(kr + 1) & 3];
            double(*nn)[NX] = N[np];
            if (hasR) {
                const int lx = red_lx(kr), ly = lyr;
                const int i = x0 + 1 + lx, j = y0 + 1 + ly;
                const int a = lx + 2, b = ly + 2;
                const double old = oc[b][a];
                const double nv =
                    ((1.0 / 6.0) * (oc[b][a + 1] + oc[b][a - 1] + oc[b + 1][a] + oc[b - 1][a] +
                                    op[b][a] + om[b][a] - src_r * h2)) * 1.0 +
                    (1.0 - 1.0) * old;
                // outside the interior (ghost cells) the old value stands
                const bool upd = i >= 1 && i <= nx && j >= 1 && j <= ny && kr >= 1 && kr <= nz;
                const double v = upd ? nv : old;
                const bool in_tile = lx >= 0 && lx < ZX && ly >= 0 && ly < ZY && kr >= k0 && kr <= k1;
                dsum += upd && in_tile ? fabs(v - old) : 0.0;
                nn[ly + 1][lx + 1] = v;
            }
        }
        lds_barrier();
        // colour 1 of plane t on the tile, and the write of plane t
        if (t >= k0) {
            const int lx = tid % ZX, ly = tid / ZX;
            const int i = x0 + 1 + lx, j = y0 + 1 + ly;
            if (i <= nx && j <= ny) {
                double(*nc)[NX] = N[n0];
                const int a = lx + 1, b = ly + 1;
                const double cur = nc[b][a];
                const double old = O[t & 3][ly + 2][lx + 2];
                const double nv =
                    ((1.0 / 6.0) * (nc[b][a + 1] + nc[b][a - 1] + nc[b + 1][a] + nc[b - 1][a] +
                                    N[np][b][a] + N[nm][b][a] - sbc * h2)) * 1.0 +
                    (1.0 - 1.0) * old;
                // every lane forms the update; a colour-0 cell keeps its new
                // value from N (a select: no exec-mask change on the scalar unit)
                const bool upd = ((i + j + t) & 1) == 1;
                const double v = upd ? nv : cur;
                dsum += upd ? fabs(v - old) : 0.0;
                eout[sidx(nx, ny, i, j, t)] = v;
            }
        }
        {
            const int r = nm;
            nm = n0;
            n0 = np;
            np = r;
        }
        // No barrier before the next step: its store of old plane t+3 goes to
        // the buffer of plane t-1, which no thread reads after the first
        // barrier of this step, and its colour-0 writes of new plane t+2 (the
        // buffer of t-1) come after its own first barrier, which every
        // thread reaches only once it has read plane t-1 here.
    }
    if (partial) {
        const double bs = block_sum<T>(dsum, red);
        if (tid == 0) partial[b] = bs;
    }
}

// resid + restriction (sources.f90:270-300, 207-228) fused: each coarse cell
// averages the residual source - Laplacian(E) of its 8 children, in the
// reference's order of the 8 terms.
template <class R, bool FROM_U>
__global__ void k_sg_restrict(const double *__restrict__ ef, const double *__restrict__ rf,
                              const R *__restrict__ U, Grid g, double fpg, int nxf, int nyf,
                              double dxl, double *__restrict__ rc, int nxc, int nyc, int nzc) {
    const long long n = (long long)nxc * nyc * nzc;
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int ic = 1 + (int)(q % nxc), jc = 1 + (int)((q / nxc) % nyc), kc = 1 + (int)(q / ((long long)nxc * nyc));
    const int a = 2 * ic - 1, b = 2 * jc - 1, cc = 2 * kc - 1;
    const long long sy = nxf + 2, sz = (long long)(nxf + 2) * (nyf + 2);
    auto res = [&](int i, int j, int k) -> double {
        const long long c = sidx(nxf, nyf, i, j, k);
        double s;
        if constexpr (FROM_U) s = src0<R>(U, g, fpg, i, j, k);
        else s = rf[c];
        const double gr = (ef[c - 1] + ef[c + 1] + ef[c - sy] + ef[c + sy] + ef[c - sz] + ef[c + sz] -
                           6.0 * ef[c]) /
                          (dxl * dxl);
        return s - gr;
    };
    const double sum = res(a, b, cc) + res(a + 1, b, cc) + res(a, b + 1, cc) + res(a, b, cc + 1) +
                       res(a + 1, b + 1, cc) + res(a + 1, b, cc + 1) + res(a, b + 1, cc + 1) +
                       res(a + 1, b + 1, cc + 1);
    rc[sidx(nxc, nyc, ic, jc, kc)] = (1.0 / 8.0) * sum;
}

// prolongation + E = E + ret (sources.f90:230-268, 189-195) fused: the
// 27/9/3/1 weights of the coarse cell and its neighbours (ghosts included)
// added onto the 8 children.
__global__ void k_sg_prolong(const double *__restrict__ ec, int nxc, int nyc, int nzc,
                             double *__restrict__ ef, int nxf, int nyf) {
    const long long n = (long long)nxc * nyc * nzc;
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = 1 + (int)(q % nxc), j = 1 + (int)((q / nxc) % nyc), k = 1 + (int)(q / ((long long)nxc * nyc));
    auto V = [&](int ii, int jj, int kk) { return ec[sidx(nxc, nyc, ii, jj, kk)]; };
    const int a = 2 * i - 1, b = 2 * j - 1, c = 2 * k - 1;
#pragma unroll
    for (int dc = 0; dc < 2; ++dc)
#pragma unroll
        for (int db = 0; db < 2; ++db)
#pragma unroll
            for (int da = 0; da < 2; ++da) {
                const int si = da ? 1 : -1, sj = db ? 1 : -1, sk = dc ? 1 : -1;
                const double s = 27.0 * V(i, j, k) + 9.0 * V(i + si, j, k) + 9.0 * V(i, j + sj, k) +
                                 9.0 * V(i, j, k + sk) + 3.0 * V(i + si, j + sj, k) +
                                 3.0 * V(i + si, j, k + sk) + 3.0 * V(i, j + sj, k + sk) +
                                 V(i + si, j + sj, k + sk);
                const long long f = sidx(nxf, nyf, a + da, b + db, c + dc);
                ef[f] = ef[f] + (1.0 / 64.0) * s;
            }
}

// The coarsest level's solve (poisson_solver, Solve, sources.f90:311-377):
// Gauss-Seidel to Dif <= 0.001, here red-black in one workgroup, capped.
__global__ __launch_bounds__(NT) void k_sg_coarse(double *__restrict__ e,
                                                  const double *__restrict__ r, int nx, int ny,
                                                  int nz, double dxl, int cap) {
    __shared__ double red[NT / 64];
    __shared__ double s_dif;
    const int n = nx * ny * nz;
    const long long sy = nx + 2, sz = (long long)(nx + 2) * (ny + 2);
    for (int iter = 0; iter < cap; ++iter) {
        double d = 0.0;
        for (int color = 0; color < 2; ++color) {
            for (int q = threadIdx.x; q < n; q += NT) {
                const int i = 1 + q % nx, j = 1 + (q / nx) % ny, k = 1 + q / (nx * ny);
                if (((i + j + k) & 1) != color) continue;
                const long long c = sidx(nx, ny, i, j, k);
                const double old = e[c];
                const double nv = ((1.0 / 6.0) * (e[c + 1] + e[c - 1] + e[c + sy] + e[c - sy] +
                                                  e[c + sz] + e[c - sz] - r[c] * (dxl * dxl))) * 1.0 +
                                  (1.0 - 1.0) * old;
                e[c] = nv;
                d += fabs(nv - old);
            }
            __syncthreads();
        }
        const double b = block_sum<NT>(d, red);
        if (threadIdx.x == 0) s_dif = b;
        __syncthreads();
        const double dif = s_dif;
        __syncthreads();
        if (!(dif > 0.001)) break;
    }
}

// Sums of n partials (count interleaved sums, stride apart) in a fixed
// order: block b sums its contiguous share into out[b * count + q].
__global__ __launch_bounds__(NT) void k_sg_sum(const double *__restrict__ partial, long long n,
                                               int stride, int count, double *__restrict__ out) {
    __shared__ double red[NT / 64];
    const long long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long long lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    for (int q = 0; q < count; ++q) {
        double s = 0.0;
        for (long long p = lo + threadIdx.x; p < hi; p += NT) s += partial[p * stride + q];
        const double b = block_sum<NT>(s, red);
        if (threadIdx.x == 0) out[(long long)blockIdx.x * count + q] = b;
    }
}

// Legendre / aLegendre (sources.f90:9-97).
__device__ __forceinline__ double legendre(double x, int l) {
    switch (l) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return 0.5 * (3.0 * (x * x) - 1.0);
    case 3: return 0.5 * (5.0 * (x * x * x) - 3.0 * x);
    default: return (1.0 / 8.0) * (35.0 * (x * x * x * x) - 30.0 * (x * x) + 3.0);
    }
}
__device__ __forceinline__ double alegendre(double x, int l, int m) {
    const double q = 1.0 - x * x, s = sqrt(q);
    switch (l * 8 + m) {
    case 9: return -s;
    case 17: return -3.0 * x * s;
    case 18: return 3.0 * q;
    case 25: return (3.0 / 2.0) * (1.0 - 5.0 * (x * x)) * s;
    case 26: return 15.0 * x * q;
    case 27: return -15.0 * (q * s);
    case 33: return -(5.0 / 2.0) * (7.0 * (x * x * x) - 3.0 * x) * s;
    case 34: return (15.0 / 2.0) * (7.0 * (x * x) - 1.0) * q;
    case 35: return -105.0 * x * (q * s);
    case 36: return 105.0 * (q * q);
    default: return legendre(x, l);
    }
}

// Spherical coordinates of a cell centre (pangle, aangle, distance,
// sources.f90:556-597).
__device__ __forceinline__ void sph(const Problem &P, int i, int j, int k, double &r, double &ct,
                                    double &fi) {
    const double x = xcoord(P, i), y = ycoord(P, j), z = zcoord(P, k);
    r = sqrt(x * x + y * y + z * z);
    ct = cos(acos(z / r));
    fi = atan2(y, x);
    if (fi < 0) fi = fi + 2. * PI_REF;
}

// Moment slot of (l, m): 0..4 c0(l); 5 + idx cm(l,m); 15 + idx sm(l,m),
// idx = l(l-1)/2 + m - 1.
__device__ __forceinline__ int mslot(int l, int m) { return l * (l - 1) / 2 + m - 1; }

// Moments (sources.f90:599-676) with the sums the MPI build reduces: per
// block, the 25 partial sums over its cells (dV = dx dy dz).
template <class R>
__global__ __launch_bounds__(NT) void k_sg_moments(const Problem *__restrict__ Pp, const R *__restrict__ U,
                                                   double *__restrict__ partial) {
    __shared__ double red[NT / 64];
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const long long n = (long long)g.nx * g.ny * g.nz;
    double acc[SG_NMOM];
#pragma unroll
    for (int q = 0; q < SG_NMOM; ++q) acc[q] = 0.0;
    const double dv = P.dx * P.dy * P.dz;
    for (long long q = blockIdx.x * (long long)NT + threadIdx.x; q < n; q += (long long)gridDim.x * NT) {
        const int i = 1 + (int)(q % g.nx), j = 1 + (int)((q / g.nx) % g.ny),
                  k = 1 + (int)(q / ((long long)g.nx * g.ny));
        double r, ct, fi;
        sph(P, i, j, k, r, ct, fi);
        const double d = (double)U[g.idx(i, j, k)] * dv;
        acc[0] += d;
        double rl = 1.0;
#pragma unroll
        for (int l = 1; l <= 4; ++l) {
            rl *= r;
            acc[l] += legendre(ct, l) * rl * d;
#pragma unroll
            for (int m = 1; m <= l; ++m) {
                const double a = alegendre(ct, l, m) * rl * d;
                acc[5 + mslot(l, m)] += a * cos(m * fi);
                acc[15 + mslot(l, m)] += a * sin(m * fi);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < SG_NMOM; ++q) {
        const double b = block_sum<NT>(acc[q], red);
        if (threadIdx.x == 0) partial[(long long)blockIdx.x * SG_NMOM + q] = b;
    }
}

// phi_[xyz]boundaries (sources.f90:678-787) on the six faces (face cells
// only): the m > 0 terms carry the factor 2 written there on top of the
// 2 (l-m)!/(l+m)! of the moments.
__global__ void k_sg_faces(const Problem *__restrict__ Pp, const double *__restrict__ mom,
                           double gconst, double *__restrict__ phi) {
    const Problem &P = *Pp;
    const int nx = P.g.nx, ny = P.g.ny, nz = P.g.nz;
    const long long nxf = 2ll * ny * nz, nyf = 2ll * nx * nz, nzf = 2ll * nx * ny;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int i, j, k;
    if (t < nxf) {
        const long long q = t >> 1;
        i = (t & 1) ? nx + 1 : 0;
        j = 1 + (int)(q % ny);
        k = 1 + (int)(q / ny);
    } else if (t < nxf + nyf) {
        const long long u = t - nxf, q = u >> 1;
        j = (u & 1) ? ny + 1 : 0;
        i = 1 + (int)(q % nx);
        k = 1 + (int)(q / nx);
    } else if (t < nxf + nyf + nzf) {
        const long long u = t - nxf - nyf, q = u >> 1;
        k = (u & 1) ? nz + 1 : 0;
        i = 1 + (int)(q % nx);
        j = 1 + (int)(q / nx);
    } else {
        return;
    }
    double rb, ct, fi;
    sph(P, i, j, k, rb, ct, fi);
    double phi_ = 0.0;
    phi_ = phi_ - mom[0] / rb;
    double rl1 = rb;
    for (int l = 1; l <= 4; ++l) {
        rl1 *= rb;
        phi_ = phi_ - mom[l] * legendre(ct, l) / rl1;
        for (int m = 1; m <= l; ++m)
            phi_ = phi_ - 2.0 * (mom[5 + mslot(l, m)] * cos(m * fi) + mom[15 + mslot(l, m)] * sin(m * fi)) *
                              alegendre(ct, l, m) / rl1;
    }
    phi[sidx(nx, ny, i, j, k)] = gconst * phi_;
}

// cm, sm = c * sums, c = 2 Gamma(l-m+1) / Gamma(l+m+1) (sources.f90:625, 647-648).
__global__ void k_sg_moment_factors(double *__restrict__ mom) {
    if (threadIdx.x != 0) return;
    for (int l = 1; l <= 4; ++l)
        for (int m = 1; m <= l; ++m) {
            double num = 1.0, den = 1.0;
            for (int q = 2; q <= l - m; ++q) num *= q;
            for (int q = 2; q <= l + m; ++q) den *= q;
            const double c = 2.0 * num / den;
            mom[5 + mslot(l, m)] = c * mom[5 + mslot(l, m)];
            mom[15 + mslot(l, m)] = c * mom[15 + mslot(l, m)];
        }
}

inline unsigned nblocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

dim3 relax_grid(const SgLevel &L) {
    return dim3(nblocks((L.nx + 1) / 2, RB_TX), nblocks(L.ny, RB_TY), (unsigned)L.nz);
}

long long relax_blocks(const SgLevel &L) {
    const dim3 gd = relax_grid(L);
    return (long long)gd.x * gd.y * gd.z;
}

constexpr int MOM_BLOCKS = 1024;
constexpr int SUM_BLOCKS = 256;
// Levels of at least RBZ_MIN_CELLS cells take the one-pass iteration, smaller
// ones the two half sweeps (the same potential bit for bit).  A one-pass march
// is latency-bound on a level with few tiles (a 128^3 level has 128), where
// the half sweeps, with many more blocks, are faster.  The environment
// variable TLQ_MG_ONEPASS_MIN_CELLS, read when a solver is set up, overrides
// the threshold.
constexpr long long RBZ_MIN_CELLS = 4194304;

long long onepass_min_cells() {
    const char *e = std::getenv("TLQ_MG_ONEPASS_MIN_CELLS");
    if (e && *e) {
        char *end = nullptr;
        const long long v = std::strtoll(e, &end, 10);
        if (end && *end == 0 && v > 0) return v;
    }
    return RBZ_MIN_CELLS;
}

bool use_rbz(const SgState &S, const SgLevel &L) {
    return (long long)L.nx * L.ny * L.nz >= S.onepass_min;
}

dim3 rbz_grid(const SgLevel &L) {
    return dim3(nblocks(L.nx, ZX), nblocks(L.ny, ZY), nblocks(L.nz, ZZ));
}

long long rbz_blocks(const SgLevel &L) {
    const dim3 gd = rbz_grid(L);
    return (long long)gd.x * gd.y * gd.z;
}

}  // namespace

int sg_levels_auto(int nx, int ny, int nz) {
    int L = 0;
    while (L < SG_MAX_LEVELS && nx % (2 << L) == 0 && ny % (2 << L) == 0 && nz % (2 << L) == 0) ++L;
    return L;
}

long long sg_coarse_cells(int nx, int ny, int nz, int levels) {
    return (long long)(nx >> levels) * (ny >> levels) * (nz >> levels);
}

hipError_t sg_alloc(SgState &S, int nx, int ny, int nz, int levels, double xspan, hipStream_t stream) {
    S.levels = levels;
    S.onepass_min = onepass_min_cells();
    for (int l = 0; l <= levels; ++l) {
        SgLevel &L = S.lv[l];
        L.nx = nx >> l;
        L.ny = ny >> l;
        L.nz = nz >> l;
        L.dxl = xspan / (double)L.nx;
        const size_t n = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
        hipError_t e = hipMalloc(&L.e, n * sizeof(double));
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_sg_fill, dim3(std::min<long long>(nblocks(n, NT), 65535)), dim3(NT), 0,
                           stream, L.e, (long long)n, 0.0);
        if (use_rbz(S, L)) {
            e = hipMalloc(&L.e2, n * sizeof(double));
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_sg_fill, dim3(std::min<long long>(nblocks(n, NT), 65535)), dim3(NT),
                               0, stream, L.e2, (long long)n, 0.0);
        }
        if (l > 0) {
            e = hipMalloc(&L.r, n * sizeof(double));
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_sg_fill, dim3(std::min<long long>(nblocks(n, NT), 65535)), dim3(NT),
                               0, stream, L.r, (long long)n, 0.0);
        }
    }
    S.phi0 = S.lv[0].e;
    long long np = (long long)MOM_BLOCKS * SG_NMOM;
    for (int l = 0; l <= levels; ++l)
        np = std::max(np, use_rbz(S, S.lv[l]) ? rbz_blocks(S.lv[l]) : 2 * relax_blocks(S.lv[l]));
    S.npartial = np;
    hipError_t e = hipMalloc(&S.partial, (size_t)S.npartial * sizeof(double));
    if (e != hipSuccess) return e;
    e = hipMalloc(&S.partial2, (size_t)SUM_BLOCKS * SG_NMOM * sizeof(double));
    if (e != hipSuccess) return e;
    e = hipMalloc(&S.scal, (1 + SG_NMOM) * sizeof(double));
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

void sg_free(SgState &S) {
    for (int l = 0; l <= S.levels; ++l) {
        if (S.lv[l].e) (void)hipFree(S.lv[l].e);
        if (S.lv[l].e2) (void)hipFree(S.lv[l].e2);
        if (S.lv[l].r) (void)hipFree(S.lv[l].r);
        S.lv[l].e = S.lv[l].e2 = S.lv[l].r = nullptr;
    }
    if (S.partial) (void)hipFree(S.partial);
    if (S.partial2) (void)hipFree(S.partial2);
    if (S.scal) (void)hipFree(S.scal);
    S.partial = S.partial2 = S.scal = S.phi0 = nullptr;
}

namespace {

// Fixed-order sum of n partials (count interleaved sums) into out[0..count).
void sum_partials(hipStream_t s, SgState &S, const double *partial, long long n, int count,
                  double *out) {
    hipLaunchKernelGGL(k_sg_sum, dim3(SUM_BLOCKS), dim3(NT), 0, s, partial, n, count, count,
                       S.partial2);
    hipLaunchKernelGGL(k_sg_sum, dim3(1), dim3(NT), 0, s, S.partial2, (long long)SUM_BLOCKS, count,
                       count, out);
}

// poisson_solver's Relax (sources.f90:378-476): it-1 = 5 iterations, each a
// colour-0 then a colour-1 sweep -- on large levels one pass per iteration
// into the second buffer (the buffers then swap), else two in-place half
// sweeps.  With want_dif, the last iteration's sum of |change| goes to
// S.scal[0].
template <class R>
void relax(hipStream_t s, const Grid &g, const R *U, double fpg, SgState &S, int l, bool want_dif) {
    SgLevel &L = S.lv[l];
    if (use_rbz(S, L)) {
        const dim3 gd = rbz_grid(L);
        for (int it = 1; it <= 5; ++it) {
            double *part = (want_dif && it == 5) ? S.partial : nullptr;
            if (l == 0) {
                hipEvent_t a = nullptr, b = nullptr;
                if (S.time_relax0) {
                    hipEventCreate(&a);
                    hipEventCreate(&b);
                    hipEventRecord(a, s);
                }
                hipLaunchKernelGGL((k_sg_rbz<R, true>), gd, dim3(ZX * ZY), 0, s, L.e, L.e2, nullptr,
                                   U, g, fpg, L.nx, L.ny, L.nz, L.dxl, part);
                if (S.time_relax0) {
                    hipEventRecord(b, s);
                    S.relax0_ev.emplace_back(a, b);
                }
            } else
                hipLaunchKernelGGL((k_sg_rbz<R, false>), gd, dim3(ZX * ZY), 0, s, L.e, L.e2, L.r, U,
                                   g, fpg, L.nx, L.ny, L.nz, L.dxl, part);
            std::swap(L.e, L.e2);
        }
        if (want_dif) sum_partials(s, S, S.partial, rbz_blocks(L), 1, S.scal);
        return;
    }
    const dim3 gd = relax_grid(L), bd(RB_TX, RB_TY);
    const long long nb = (long long)gd.x * gd.y * gd.z;
    for (int it = 1; it <= 5; ++it)
        for (int color = 0; color < 2; ++color) {
            double *part = (want_dif && it == 5) ? S.partial + color * nb : nullptr;
            if (l == 0)
                hipLaunchKernelGGL((k_sg_relax<R, true>), gd, bd, 0, s, L.e, nullptr, U, g, fpg, L.nx,
                                   L.ny, L.nz, L.dxl, color, part);
            else
                hipLaunchKernelGGL((k_sg_relax<R, false>), gd, bd, 0, s, L.e, L.r, U, g, fpg, L.nx,
                                   L.ny, L.nz, L.dxl, color, part);
        }
    if (want_dif) sum_partials(s, S, S.partial, 2 * nb, 1, S.scal);
}

void bc(hipStream_t s, SgLevel &L) {
    const long long n = 2ll * (L.ny + 2) * (L.nz + 2) + 2ll * L.nx * (L.nz + 2) + 2ll * L.nx * L.ny;
    hipLaunchKernelGGL(k_sg_bc, dim3(nblocks(n, NT)), dim3(NT), 0, s, L.e, L.e2, L.nx, L.ny, L.nz, 0);
}

// Level 0's ghost cells into its second buffer (they hold PHI's boundary
// values, which the one-pass iterations read from either buffer).
void copy_shell(hipStream_t s, SgLevel &L) {
    if (!L.e2) return;
    const long long n = 2ll * (L.ny + 2) * (L.nz + 2) + 2ll * L.nx * (L.nz + 2) + 2ll * L.nx * L.ny;
    hipLaunchKernelGGL(k_sg_bc, dim3(nblocks(n, NT)), dim3(NT), 0, s, L.e, L.e2, L.nx, L.ny, L.nz, 1);
}

}  // namespace

template <class R>
hipError_t sg_solve(hipStream_t s, const Grid &g, const R *U, SgState &S, int it,
                    const Problem *dP) {
    const double fpg = 4.0 * PI_REF * S.gconst;
    const int LV = S.levels;
    SgLevel &L0 = S.lv[0];
    if (S.boundary == 2 && it % 10 == 0) {   // sources.f90:114-136
        hipLaunchKernelGGL((k_sg_moments<R>), dim3(MOM_BLOCKS), dim3(NT), 0, s, dP, U, S.partial);
        sum_partials(s, S, S.partial, (long long)MOM_BLOCKS, SG_NMOM, S.scal + 1);
        hipLaunchKernelGGL(k_sg_moment_factors, dim3(1), dim3(64), 0, s, S.scal + 1);
        const long long nf = 2ll * (L0.ny * (long long)L0.nz + L0.nx * (long long)L0.nz +
                                    L0.nx * (long long)L0.ny);
        hipLaunchKernelGGL(k_sg_faces, dim3(nblocks(nf, NT)), dim3(NT), 0, s, dP, S.scal + 1,
                           S.gconst, L0.e);
    }
    copy_shell(s, L0);
    double dif = 1.e28;
    int iter = 0;
    while (dif > S.tol && iter < 31) {   // sources.f90:140 (tol 10)
        ++iter;
        for (int l = 1; l <= LV; ++l) {
            const long long n = (long long)(S.lv[l].nx + 2) * (S.lv[l].ny + 2) * (S.lv[l].nz + 2);
            hipLaunchKernelGGL(k_sg_fill, dim3(std::min<long long>(nblocks(n, NT), 65535)), dim3(NT), 0,
                               s, S.lv[l].e, n, 0.0);
        }
        relax<R>(s, g, U, fpg, S, 0, LV == 0);
        for (int l = 1; l <= LV; ++l) {
            SgLevel &F = S.lv[l - 1], &C = S.lv[l];
            const long long nc = (long long)C.nx * C.ny * C.nz;
            if (l == 1)
                hipLaunchKernelGGL((k_sg_restrict<R, true>), dim3(nblocks(nc, NT)), dim3(NT), 0, s, F.e,
                                   nullptr, U, g, fpg, F.nx, F.ny, F.dxl, C.r, C.nx, C.ny, C.nz);
            else
                hipLaunchKernelGGL((k_sg_restrict<R, false>), dim3(nblocks(nc, NT)), dim3(NT), 0, s,
                                   F.e, F.r, U, g, fpg, F.nx, F.ny, F.dxl, C.r, C.nx, C.ny, C.nz);
            bc(s, C);
            if (l == LV)
                hipLaunchKernelGGL(k_sg_coarse, dim3(1), dim3(NT), 0, s, C.e, C.r, C.nx, C.ny, C.nz,
                                   C.dxl, SG_COARSE_CAP);
            else
                relax<R>(s, g, U, fpg, S, l, false);
        }
        for (int l = LV - 1; l >= 0; --l) {
            SgLevel &F = S.lv[l], &C = S.lv[l + 1];
            const long long nc = (long long)C.nx * C.ny * C.nz;
            hipLaunchKernelGGL(k_sg_prolong, dim3(nblocks(nc, NT)), dim3(NT), 0, s, C.e, C.nx, C.ny,
                               C.nz, F.e, F.nx, F.ny);
            if (l > 0) bc(s, F);
            relax<R>(s, g, U, fpg, S, l, l == 0);
        }
        hipError_t e = hipMemcpyAsync(&dif, S.scal, sizeof(double), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        for (auto &ev : S.relax0_ev) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, ev.first, ev.second);
            S.relax0_ms += ms;
            ++S.relax0_n;
            hipEventDestroy(ev.first);
            hipEventDestroy(ev.second);
        }
        S.relax0_ev.clear();
    }
    if (L0.e != S.phi0) {   // an odd number of one-pass iterations: PHI is in the second buffer
        const size_t n = (size_t)(L0.nx + 2) * (L0.ny + 2) * (L0.nz + 2);
        hipError_t e = hipMemcpyAsync(S.phi0, L0.e, n * sizeof(double), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        std::swap(L0.e, L0.e2);
    }
    S.cycles = iter;
    S.dif = dif;
    return hipGetLastError();
}

template hipError_t sg_solve<float>(hipStream_t, const Grid &, const float *, SgState &, int,
                                    const Problem *);
template hipError_t sg_solve<double>(hipStream_t, const Grid &, const double *, SgState &, int,
                                     const Problem *);

}  // namespace tlq
