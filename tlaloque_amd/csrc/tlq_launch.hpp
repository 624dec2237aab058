// tlq_launch.hpp -- kernel launchers, explicitly instantiated per precision in
// tlq_fp64.hip (built without FMA contraction, for bitwise agreement with the
// reference's arithmetic) and tlq_fp32.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "tlq_device.hpp"

namespace tlq {

struct StepState;
struct Tiling;

// Tile shapes of the fused kernels per precision.
template <class R> struct TileShape;
// Tile shape of each fused kernel: TX x TY columns, ZC planes, and the
// minimum number of waves per SIMD the kernel is compiled for (its register
// budget follows from it).  Stage 1 (first-order fluxes, fewer live values)
// runs small blocks at three waves per SIMD; stage 2 is register-bound at two.
template <int TX_, int TY_, int ZC_, int WAVES_> struct Tile {
    static constexpr int TX = TX_, TY = TY_, ZC = ZC_, WAVES = WAVES_;
};
#ifndef TLQ_F32_S1
#define TLQ_F32_S1 32, 8, 32, 3
#endif
#ifndef TLQ_F32_S2
#define TLQ_F32_S2 64, 8, 64, 1
#endif
template <> struct TileShape<float> {
    using S1 = Tile<TLQ_F32_S1>;
    using S2 = Tile<TLQ_F32_S2>;
};
template <> struct TileShape<double> {
    using S1 = Tile<32, 8, 64, 1>;
    using S2 = Tile<32, 8, 64, 1>;
};

// Tiles of a mesh for one tile shape.
inline Tiling make_tiling(int nx, int ny, int nz, int tx, int ty, int zc) {
    Tiling T;
    T.nbx = (nx + tx - 1) / tx;
    T.nby = (ny + ty - 1) / ty;
    T.nbz = (nz + zc - 1) / zc;
    T.ntiles = T.nbx * T.nby * T.nbz;
    T.per_xcd = (T.ntiles + 7) / 8;
    return T;
}

template <class R> struct Launch {
    static hipError_t calc_prims(hipStream_t s, const Problem *P, const R *U, R *W, long long ncell);
    static hipError_t maxima(hipStream_t s, const Problem *P, const R *A, int prims,
                             unsigned long long *out, long long ninterior);
    static hipError_t fluxes(hipStream_t s, const Problem *P, const R *U, const R *W, R *F, R *G,
                             R *H, int choice, long long nfaces);
    static hipError_t godunov(hipStream_t s, const Problem *P, R *out, const R *U, const R *F,
                              const R *G, const R *H, const R *S, double dt, long long ninterior);
    static hipError_t bc_plane(hipStream_t s, const Problem *P, R *U, int axis, int dst, int src,
                               int negate, long long nplane);
    static hipError_t winds(hipStream_t s, const Problem *P, R *U, int i0, int j0, int k0, int ni,
                            int nj, int nk);
    static hipError_t copy(hipStream_t s, R *dst, const R *src, long long n);
    static hipError_t step(hipStream_t s, const Problem *P, R *U, const R *UP, const R *UPP,
                           long long ncell);
    static hipError_t cooling(hipStream_t s, const Problem *P, R *W, R *U, double dt,
                              long long ninterior);
    static hipError_t gravity(hipStream_t s, const Problem *P, const R *W, R *S, long long ninterior);
    static hipError_t aos_to_soa(hipStream_t s, const Problem *P, const double *aos, R *soa, int k0,
                                 int nk, long long ncell);
    static hipError_t soa_to_aos(hipStream_t s, const Problem *P, const R *soa, double *aos, int k0,
                                 int nk, long long ncell);
    static hipError_t check(hipStream_t s, const Problem *P, const R *U, unsigned long long *bad,
                            long long ninterior);
    static hipError_t stage1(hipStream_t s, const Problem *P, const Tiling &T, const R *U, R *Uh,
                             StepState *st, int parity, int solver);
    // cqueue/cseg: the cooling queue (modes 2 and 3), one segment of
    // TX*TY*ZC cells per tile, and the per-tile counts; null otherwise.
    static hipError_t stage2(hipStream_t s, const Problem *P, const Tiling &T, const R *U,
                             const R *Uh, R *Unew, StepState *st, int parity, int limiter,
                             int solver, unsigned *cqueue, unsigned *cseg);
};

}  // namespace tlq
