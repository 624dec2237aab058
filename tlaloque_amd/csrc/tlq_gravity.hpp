// Self-gravity on the GPU: the multigrid Poisson solve of MultiGrid
// (source/sources.f90:99-205) and its multipole boundary values
// (sources.f90:599-787), all in FP64.
//
// Level l holds the potential/correction E_l with one ghost layer,
// (nx_l+2)(ny_l+2)(nz_l+2) doubles, i fastest, nx_l = nx / 2^l; level 0's E
// is PHI(0:nx+1, 0:ny+1, 0:nz+1) itself (globals.f90:19).  Levels l >= 1 also
// hold the restricted residual R_l (same shape).  Level 0's source 4 pi G rho
// is read from U's density when needed, never stored.
//
// The V-cycle keeps the reference's structure: per cycle the correction
// levels are zeroed, each level is pre-relaxed on the way down (5 iterations)
// and post-relaxed on the way up (5 iterations), the coarsest level is solved
// to Dif <= 0.001, correction levels take minus-interior ghosts before each
// relaxation, and cycles repeat while Dif > 10 and fewer than 31 have run
// (Dif: the sum of |change| over level 0's last iteration).  Relaxation is
// red-black Gauss-Seidel (the reference's sweep order is sequential), so a
// cycle is not the reference's bit for bit; the converged potential is the
// same discrete solution.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "tlq_device.hpp"

namespace tlq {

constexpr int SG_MAX_LEVELS = 16;
constexpr int SG_NMOM = 25;             // c0(0:4), cm(l,m), sm(l,m) for 1 <= m <= l <= 4
constexpr int SG_COARSE_MAX = 4096;     // cells of the coarsest level (one workgroup)
constexpr int SG_COARSE_CAP = 20000;    // iterations cap of the coarsest solve

struct SgLevel {
    int nx, ny, nz;
    double dxl;    // (xr - xl) / nx_l, as poisson_solver and resid compute it
    double *e;     // E_l (level 0: PHI); the current buffer
    double *e2;    // second buffer of the one-pass iterations (large levels), or null
    double *r;     // R_l (levels >= 1)
};

struct SgState {
    int levels = 0;
    int boundary = 0;       // 1: faces kept as set; 2: multipole every 10 steps
    double gconst = 0.0;
    double tol = 10.0;           // V-cycles while Dif > tol (sources.f90:140 uses 10)
    // Levels of at least this many cells relax with the one-pass iteration,
    // smaller ones with two in-place half sweeps (set by sg_alloc).
    long long onepass_min = 0;
    SgLevel lv[SG_MAX_LEVELS + 1]{};
    double *partial = nullptr;   // per-block partial sums
    long long npartial = 0;
    double *partial2 = nullptr;  // second stage of the sums
    double *phi0 = nullptr;      // level 0's first buffer: PHI as the hydro kernels see it
    double *scal = nullptr;      // [0] Dif, [1..25] moments
    int cycles = 0;              // V-cycles of the last solve
    double dif = 0.0;            // its final Dif
    // Timing of level 0's relaxation launches (HIP events on the solve's
    // stream) when time_relax0 is set: summed into relax0_ms / relax0_n.
    bool time_relax0 = false;
    double relax0_ms = 0.0;
    long long relax0_n = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> relax0_ev;
};

// Most halvings that keep nx, ny, nz whole.
int sg_levels_auto(int nx, int ny, int nz);
// Cells of the coarsest level for `levels` halvings.
long long sg_coarse_cells(int nx, int ny, int nz, int levels);
hipError_t sg_alloc(SgState &S, int nx, int ny, int nz, int levels, double xspan,
                    hipStream_t stream);
void sg_free(SgState &S);

// One MultiGrid(dens, PHI) call at step count `it`, dens = U's density
// (dP: the device Problem, for the multipole faces).  Synchronises the
// stream once per V-cycle (the Dif test).
template <class R>
hipError_t sg_solve(hipStream_t s, const Grid &g, const R *U, SgState &S, int it,
                    const Problem *dP);

}  // namespace tlq
