// tlq_launch_impl.hpp -- definitions of the launchers (included once per
// precision translation unit).
#pragma once

#include "tlq_fused.hpp"
#include "tlq_launch.hpp"

namespace tlq {

static inline unsigned grid_for(long long n, int block) {
    long long b = (n + block - 1) / block;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;  // grid-stride loops cover the rest
    return (unsigned)b;
}

template <class R>
hipError_t Launch<R>::calc_prims(hipStream_t s, const Problem *P, const R *U, R *W, long long n) {
    hipLaunchKernelGGL(k_calc_prims<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, U, W);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::maxima(hipStream_t s, const Problem *P, const R *A, int prims,
                             unsigned long long *out, long long n) {
    unsigned nb = grid_for(n, 256 * 16);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL((k_maxima<R, 256>), dim3(nb), dim3(256), 0, s, P, A, prims, out);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::fluxes(hipStream_t s, const Problem *P, const R *U, const R *W, R *F, R *G,
                             R *H, int choice, long long n) {
    hipLaunchKernelGGL(k_fluxes<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, U, W, F, G, H,
                       choice);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::godunov(hipStream_t s, const Problem *P, R *out, const R *U, const R *F,
                              const R *G, const R *H, const R *S, double dt, long long n) {
    hipLaunchKernelGGL(k_godunov<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, out, U, F, G, H, S,
                       dt);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::bc_plane(hipStream_t s, const Problem *P, R *U, int axis, int dst, int src,
                               int negate, long long n) {
    hipLaunchKernelGGL(k_bc_plane<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, U, axis, dst, src,
                       negate);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::winds(hipStream_t s, const Problem *P, R *U, int i0, int j0, int k0, int ni,
                            int nj, int nk) {
    const long long n = (long long)ni * nj * nk;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_winds<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, U, i0, j0, k0, ni,
                       nj, nk);
    return hipGetLastError();
}
template <class R> hipError_t Launch<R>::copy(hipStream_t s, R *dst, const R *src, long long n) {
    hipLaunchKernelGGL(k_copy<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::step(hipStream_t s, const Problem *P, R *U, const R *UP, const R *UPP,
                           long long n) {
    hipLaunchKernelGGL(k_step<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, U, UP, UPP);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::cooling(hipStream_t s, const Problem *P, R *W, R *U, double dt, long long n) {
    hipLaunchKernelGGL(k_cooling<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, W, U, dt);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::gravity(hipStream_t s, const Problem *P, const R *W, R *S, long long n) {
    hipLaunchKernelGGL(k_gravity<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, W, S);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::aos_to_soa(hipStream_t s, const Problem *P, const double *aos, R *soa, int k0,
                                 int nk, long long n) {
    hipLaunchKernelGGL(k_aos_to_soa<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, aos, soa, k0,
                       nk);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::soa_to_aos(hipStream_t s, const Problem *P, const R *soa, double *aos, int k0,
                                 int nk, long long n) {
    hipLaunchKernelGGL(k_soa_to_aos<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, soa, aos, k0,
                       nk);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::check(hipStream_t s, const Problem *P, const R *U, unsigned long long *bad,
                            long long n) {
    hipLaunchKernelGGL(k_check<R>, dim3(grid_for(n, 256)), dim3(256), 0, s, P, U, bad);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::stage1(hipStream_t s, const Problem *P, const Tiling &T, const R *U, R *Uh,
                             StepState *st, int parity, int solver) {
    using TS = typename TileShape<R>::S1;
    const dim3 grid(8 * T.per_xcd), block(TS::TX * TS::TY);
    if (solver == SOLVER_HLL)
        hipLaunchKernelGGL((k_stage1<R, TS::TX, TS::TY, TS::ZC, SOLVER_HLL, TS::WAVES>), grid, block, 0, s, P,
                           T, U, Uh, st, parity);
    else
        hipLaunchKernelGGL((k_stage1<R, TS::TX, TS::TY, TS::ZC, SOLVER_HLLC, TS::WAVES>), grid, block, 0, s,
                           P, T, U, Uh, st, parity);
    return hipGetLastError();
}
template <class R>
hipError_t Launch<R>::stage2(hipStream_t s, const Problem *P, const Tiling &T, const R *U,
                             const R *Uh, R *Unew, StepState *st, int parity, int limiter,
                             int solver, unsigned *cqueue, unsigned *cseg) {
    using TS = typename TileShape<R>::S2;
    const dim3 grid(8 * T.per_xcd), block(TS::TX * TS::TY);
    // van Leer (the reference default) is compiled in; other limiters are
    // selected at run time inside one kernel.
    const bool vl = limiter == LIM_VANLEER;
    if (solver == SOLVER_HLL) {
        if (vl)
            hipLaunchKernelGGL((k_stage2<R, TS::TX, TS::TY, TS::ZC, LIM_VANLEER, SOLVER_HLL, TS::WAVES>), grid,
                               block, 0, s, P, T, U, Uh, Unew, st, parity, cqueue, cseg);
        else
            hipLaunchKernelGGL((k_stage2<R, TS::TX, TS::TY, TS::ZC, 0, SOLVER_HLL, TS::WAVES>), grid, block, 0,
                               s, P, T, U, Uh, Unew, st, parity, cqueue, cseg);
    } else {
        if (vl)
            hipLaunchKernelGGL((k_stage2<R, TS::TX, TS::TY, TS::ZC, LIM_VANLEER, SOLVER_HLLC, TS::WAVES>),
                               grid, block, 0, s, P, T, U, Uh, Unew, st, parity, cqueue, cseg);
        else
            hipLaunchKernelGGL((k_stage2<R, TS::TX, TS::TY, TS::ZC, 0, SOLVER_HLLC, TS::WAVES>), grid, block, 0,
                               s, P, T, U, Uh, Unew, st, parity, cqueue, cseg);
    }
    return hipGetLastError();
}

}  // namespace tlq
